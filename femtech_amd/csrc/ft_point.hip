// ft_point.hip -- FemTech's per-element, per-Gauss-point routines on the GPU.
//
// The reference evaluates the explicit step one element and one Gauss point
// at a time through free functions that read and write global arrays
// (ShapeFunction_C3D8, CalculateDeformationGradient, DeterminateF, InverseF,
// StressUpdate and the material routines, StrainDisplacementMatrix,
// InternalForceUpdate, CalculateCharacteristicLength_C3D8,
// CalculateTimeStep, inverse3x3Matrix).  The engine never needs them (its
// element kernel fuses the whole path in registers), but drivers may call
// them directly.  ft_point_eval evaluates any batch of such calls on the
// device, one thread per call, in FP64; the FemTech surface
// (ft_femtech.cpp) gathers each call's inputs from its host arrays and
// scatters the outputs back, exactly where the reference writes them.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <mutex>

#include "../../include/femtech_amd.h"
#include "ft_device.h"

extern "C" __attribute__((visibility("hidden"))) void ft_internal_set_error(const char *msg);

namespace {

using namespace ft;

constexpr int kBlockP = 256;

// Record sizes (doubles) of each operation: inputs, outputs.
__host__ __device__ constexpr int in_size(int op) {
  return op == FT_PT_SHAPE ? 27 : op == FT_PT_DEFGRAD ? 48 : op == FT_PT_DETF ? 9 : op == FT_PT_INVF ? 10
       : op == FT_PT_STRESS ? 20 : op == FT_PT_BMAT ? 12 : op == FT_PT_FINT ? 64 : op == FT_PT_CHARLEN ? 24
       : op == FT_PT_TIMESTEP ? 27 : op == FT_PT_INV3 ? 9 : 0;
}
__host__ __device__ constexpr int out_size(int op) {
  return op == FT_PT_SHAPE ? 33 : op == FT_PT_DEFGRAD ? 9 : op == FT_PT_DETF ? 1 : op == FT_PT_INVF ? 9
       : op == FT_PT_STRESS ? 34 : op == FT_PT_BMAT ? 18 : op == FT_PT_FINT ? 192 : op == FT_PT_CHARLEN ? 1
       : op == FT_PT_TIMESTEP ? 1 : op == FT_PT_INV3 ? 10 : 0;
}

// Determinant and adjugate of a 3x3 matrix m stored column-major (m[3 c + r]).
__device__ inline double adj3(const double *m, double *adj) {
  auto a = [&](int r, int c) { return m[3 * c + r]; };
  // adj[3 c + r] = cofactor (c, r)
  adj[0] = a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1);
  adj[3] = a(0, 2) * a(2, 1) - a(0, 1) * a(2, 2);
  adj[6] = a(0, 1) * a(1, 2) - a(0, 2) * a(1, 1);
  adj[1] = a(1, 2) * a(2, 0) - a(1, 0) * a(2, 2);
  adj[4] = a(0, 0) * a(2, 2) - a(0, 2) * a(2, 0);
  adj[7] = a(0, 2) * a(1, 0) - a(0, 0) * a(1, 2);
  adj[2] = a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0);
  adj[5] = a(0, 1) * a(2, 0) - a(0, 0) * a(2, 1);
  adj[8] = a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0);
  return a(0, 0) * adj[0] + a(0, 1) * adj[1] + a(0, 2) * adj[2];
}

// C3D8 reference coordinates of the local nodes (ShapeFunction_C3D8.cpp:22-29).
__device__ inline double node_sign(int a, int k) {
  const int s[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                       {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  return (double)s[a][k];
}

// ShapeFunction_C3D8.cpp:4-128 at the point chi: N_a, dN_a/dX and det J,
// with J = dx/dchi from the element's node coordinates xe[3 a + j].
// in: chi[3], xe[24]; out: shp[8], dshp[24] (node-major, d/dX d/dY d/dZ), detJ.
__device__ void op_shape(const double *in, double *out) {
  const double *chi = in, *xe = in + 3;
  double dn[8][3];
  for (int a = 0; a < 8; ++a) {
    double f[3];
    for (int k = 0; k < 3; ++k) f[k] = 1.0 + node_sign(a, k) * chi[k];
    out[a] = f[0] * f[1] * f[2] / 8.0;
    dn[a][0] = node_sign(a, 0) * f[1] * f[2] / 8.0;
    dn[a][1] = node_sign(a, 1) * f[0] * f[2] / 8.0;
    dn[a][2] = node_sign(a, 2) * f[0] * f[1] / 8.0;
  }
  // J[3 k + j] = dx_j / dchi_k (column-major, rows = x components)
  double J[9], adj[9];
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k) {
      double s = 0.0;
      for (int a = 0; a < 8; ++a) s += xe[3 * a + j] * dn[a][k];
      J[3 * k + j] = s;
    }
  const double det = adj3(J, adj);
  // dN/dX_j = sum_k dN/dchi_k (J^-1)[k][j], (J^-1)[k][j] = adj[3 j + k] / det
  for (int a = 0; a < 8; ++a)
    for (int j = 0; j < 3; ++j)
      out[8 + 3 * a + j] = (dn[a][0] * adj[3 * j + 0] + dn[a][1] * adj[3 * j + 1] + dn[a][2] * adj[3 * j + 2]) / det;
  out[32] = det;
}

// CalculateDeformationGradient.cpp:11-25: F_ij = sum_a x_a,i dN_a/dX_j,
// stored column-major.  in: x[24] (X + u per local node), dshp[24].
__device__ void op_defgrad(const double *in, double *out) {
  const double *x = in, *dn = in + 24;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int a = 0; a < 8; ++a) s += x[3 * a + i] * dn[3 * a + j];
      out[3 * j + i] = s;
    }
}

__device__ void op_detf(const double *in, double *out) {
  double adj[9];
  out[0] = adj3(in, adj);
}

// InverseF.cpp: the matrix whose rows are F's consecutive triples, inverted
// with the given determinant (detF, as the reference reads it).
__device__ void op_invf(const double *in, double *out) {
  double At[9], adj[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) At[3 * c + r] = in[3 * r + c];  // A[r][c] = F[3 r + c]
  (void)adj3(At, adj);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) out[3 * r + c] = adj[3 * c + r] / in[9];
}

// StressUpdate.cpp:5-29 (material given explicitly, so the individual material
// routines are served as well).  in: F[9], material id, props[9], dt;
// out: pk2[6] (Voigt 11, 22, 33, 23, 13, 12), history[27] (Hn_1, Hn_2, S0n
// as full 3x3 matrices, materials 5 and 7: read and updated in place), and
// a flag (1: unknown material).
__device__ void op_stress(const double *in, double *out) {
  const double *Fm = in;
  const int mat = (int)in[9];
  const double *pr = in + 10;
  const double dt = in[19];
  PartDev pp;
  memset(&pp, 0, sizeof(pp));
  pp.mat = mat;
  pp.rho = pr[0];
  pp.mu = pr[1];
  pp.lambda = pr[2];
  pp.k1 = pr[3];
  pp.k2 = pr[4];
  pp.g1 = pr[5];
  pp.t1 = pr[6];
  pp.g2 = pr[7];
  pp.t2 = pr[8];
  pp.alpha = pr[3];
  pp.kappa = pr[5];
  pp.visco = mat == 5 || mat == 7;
  double H[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) H[i][j] = Fm[3 * j + i] - (i == j ? 1.0 : 0.0);
  // history: full symmetric 3x3 (column-major) <-> Voigt (00, 11, 22, 12, 02, 01)
  const int vi[6] = {0, 4, 8, 7, 6, 3};
  double *hist27 = out + 6;
  double h[18];
  for (int t = 0; t < 3; ++t)
    for (int k = 0; k < 6; ++k) h[6 * t + k] = hist27[9 * t + vi[k]];
  ViscoCoef vc = {0, 0, 0, 0};
  if (pp.visco) {  // HGOIsotropicViscoelastic.cpp:126-131, the global dt
    const double rt1 = dt / pp.t1, rt2 = dt / pp.t2;
    vc.c11 = exp(-rt1);
    vc.c12 = exp(-rt2);
    vc.c21 = pp.g1 * (1 - vc.c11) / rt1;
    vc.c22 = pp.g2 * (1 - vc.c12) / rt2;
  }
  Sym<double> S;
  double Jm1;
  if (!pk2_stress<double, -1>(pp, H, h, pp.visco != 0, vc, S, Jm1)) {
    out[33] = 1.0;  // unknown material
    return;
  }
  for (int k = 0; k < 6; ++k) out[k] = S.v[k];
  if (pp.visco) {
    for (int t = 0; t < 3; ++t) {
      double *m = hist27 + 9 * t;
      const double *v = h + 6 * t;
      m[0] = v[0];
      m[4] = v[1];
      m[8] = v[2];
      m[5] = m[7] = v[3];
      m[2] = m[6] = v[4];
      m[1] = m[3] = v[5];
    }
  }
  out[33] = 0.0;
}

// StrainDisplacementMatrix.cpp:25-77: Belytschko's B0 of node I (6 x 3,
// column-major) from dN_I/dX and F.  in: dshp_I[3], F[9].
__device__ void op_bmat(const double *in, double *out) {
  const double *dn = in, *Fm = in + 3;
  auto F = [&](int r, int c) { return Fm[3 * c + r]; };  // F(r, c) = F_{r+1, c+1}
  // column d of B: strain-displacement terms for displacement component d
  for (int d = 0; d < 3; ++d) {
    double *col = out + 6 * d;
    col[0] = dn[0] * F(d, 0);
    col[1] = dn[1] * F(d, 1);
    col[2] = dn[2] * F(d, 2);
    col[3] = dn[1] * F(d, 2) + dn[2] * F(d, 1);
    col[4] = dn[0] * F(d, 2) + dn[2] * F(d, 0);
    col[5] = dn[0] * F(d, 1) + dn[1] * F(d, 0);
  }
}

// InternalForceUpdate.cpp:4-28: B (6 x 24) of all nodes, fintGQ = B^T S and
// force += w detJ fintGQ.  in: dshp[24], F[9], pk2[6], w detJ, force[24];
// out: force[24], fintGQ[24], B[144].
__device__ void op_fint(const double *in, double *out) {
  const double *dn = in, *Fm = in + 24, *S = in + 33, pre = in[39];
  const double *force = in + 40;
  double *B = out + 48;
  for (int a = 0; a < 8; ++a) {
    double rec[12];
    for (int k = 0; k < 3; ++k) rec[k] = dn[3 * a + k];
    for (int k = 0; k < 9; ++k) rec[3 + k] = Fm[k];
    op_bmat(rec, B + 18 * a);
  }
  for (int c = 0; c < 24; ++c) {
    double s = 0.0;
    for (int r = 0; r < 6; ++r) s += B[6 * c + r] * S[r];
    out[24 + c] = s;
    out[c] = force[c] + pre * s;
  }
}

// CalculateCharacteristicLength_C3D8.cpp:3-30 (volume over the largest face
// area, Geometry.cpp:3-64) from the deformed node positions x[24].
__device__ double char_length64(const double *x) {
  double c[7][3];
  for (int d = 0; d < 3; ++d) {
    double v[8];
    for (int p = 0; p < 8; ++p) v[p] = p == 0 ? 0.0 : x[3 * p2n(p) + d] - x[3 * p2n(0) + d];
    wht8(v);
    for (int m = 0; m < 7; ++m) c[m][d] = v[m + 1] * 0.125;
  }
  double amax = 0.0;
  for (int f = 0; f < 6; ++f) {
    const double a = hex_part(c, f);
    if (a > amax) amax = a;
  }
  return hex_part(c, 6) / amax;
}

// CalculateTimeStep.cpp:7-21: Le / ce, ce from (rho, mu, lambda) in the
// reference's operation order.  in: x[24], props[3].
__device__ double time_step64(const double *in) {
  const double le = char_length64(in);
  const double rho = in[24], mu = in[25], lambda = in[26];
  const double nu = 0.5 * lambda / (lambda + mu);
  const double ce = sqrt(lambda * (1.0 / nu - 1.0) / rho);
  return le / ce;
}

// math.cpp:8-28: inverse of a column-major 3x3 and its determinant.
__device__ void op_inv3(const double *in, double *out) {
  double adj[9];
  const double det = adj3(in, adj);
  const double inv = 1.0 / det;
  for (int k = 0; k < 9; ++k) out[k] = adj[k] * inv;
  out[9] = det;
}

__global__ __launch_bounds__(kBlockP) void k_point(int op, int n, const double *in, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *a = in + (size_t)in_size(op) * i;
  double *b = out + (size_t)out_size(op) * i;
  switch (op) {
    case FT_PT_SHAPE: op_shape(a, b); break;
    case FT_PT_DEFGRAD: op_defgrad(a, b); break;
    case FT_PT_DETF: op_detf(a, b); break;
    case FT_PT_INVF: op_invf(a, b); break;
    case FT_PT_STRESS: op_stress(a, b); break;
    case FT_PT_BMAT: op_bmat(a, b); break;
    case FT_PT_FINT: op_fint(a, b); break;
    case FT_PT_CHARLEN: b[0] = char_length64(a); break;
    case FT_PT_TIMESTEP: b[0] = time_step64(a); break;
    case FT_PT_INV3: op_inv3(a, b); break;
    default: break;
  }
}

// Device staging for the calls, grown on demand and kept (one per process).
std::mutex g_mu;
double *g_buf = nullptr;
size_t g_len = 0;

int set_error(int code, const char *msg) {
  ft_internal_set_error(msg);
  return code;
}

}  // namespace

extern "C" int ft_point_eval(int op, int n, const double *in, double *out) {
  if (op < FT_PT_SHAPE || op > FT_PT_INV3) return set_error(FT_ERR_ARG, "unknown point operation");
  if (n < 0 || (n > 0 && (!in || !out))) return set_error(FT_ERR_ARG, "invalid point arguments");
  if (n == 0) return 0;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(FT_ERR_HIP, "no HIP device available");
  const size_t ni = (size_t)in_size(op) * n, no = (size_t)out_size(op) * n;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_len < ni + no) {
    if (g_buf) (void)hipFree(g_buf);
    g_buf = nullptr;
    g_len = 0;
    if (hipMalloc(&g_buf, (ni + no) * sizeof(double)) != hipSuccess)
      return set_error(FT_ERR_ALLOC, "point staging allocation failed");
    g_len = ni + no;
  }
  double *din = g_buf, *dout = g_buf + ni;
  // The stress operation updates the Prony history in place: it arrives in
  // the output record (out[6..32]), which is therefore an input as well.
  if (hipMemcpy(din, in, ni * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      (op == FT_PT_STRESS && hipMemcpy(dout, out, no * sizeof(double), hipMemcpyHostToDevice) != hipSuccess))
    return set_error(FT_ERR_HIP, "point upload failed");
  hipLaunchKernelGGL(k_point, dim3((n + kBlockP - 1) / kBlockP), dim3(kBlockP), 0, 0, op, n, din, dout);
  if (hipGetLastError() != hipSuccess) return set_error(FT_ERR_HIP, "point kernel launch failed");
  if (hipMemcpy(out, dout, no * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return set_error(FT_ERR_HIP, "point download failed");
  if (op == FT_PT_STRESS)
    for (int i = 0; i < n; ++i)
      if (out[(size_t)out_size(op) * i + 33] != 0.0) return set_error(FT_ERR_MATERIAL, "Unknown material type");
  return 0;
}
