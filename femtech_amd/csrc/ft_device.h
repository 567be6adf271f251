// ft_device.h -- device-side element mechanics shared by the element kernels.
//
// One thread evaluates one C3D8 element: the trilinear position and
// displacement fields are taken to their Walsh-Hadamard (hourglass) basis, the
// 2x2x2 Gauss-point gradients come from that basis, the material law runs in
// registers, and the nodal forces come back through the transposed transform.
// Reference semantics: ShapeFunction_C3D8.cpp:4-128 (dN/dxi, J0),
// CalculateDeformationGradient.cpp:11-25 (F), StressUpdate.cpp:5-29 and the
// material files (S), InternalForceUpdate.cpp:4-28 with
// StrainDisplacementMatrix.cpp:27-76 (f_a = sum_gp w detJ0 B_a^T S),
// CalculateCharacteristicLength_C3D8.cpp:3-30 with Geometry.cpp:3-64 (Le).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace ft {

// Gauss abscissa as written in GaussQuadrature3D.cpp:19-49 and the face
// quadrature abscissa of Geometry.cpp:50 (sqrt(3)/3 evaluated in FP64).
constexpr double kGp = 0.577350269189626;
constexpr double kFaceQ = 0.57735026918962573;  // sqrt(3.0)/3.0 rounded to FP64

struct PartDev {
  int mat;     // material id (StressUpdate.cpp:7-27)
  int visco;   // 1 if the part carries Prony history
  double rho, mu, lambda, k1, k2, g1, t1, g2, t2;
  double ce;   // dilatational wave speed, CalculateTimeStep.cpp:13-19
  double inv_ce;
  double alpha;  // Ogden exponent (materials 6, 7)
  double kappa;  // fibre dispersion (material 8)
};

template <typename R> struct Num;
template <> struct Num<float> {
  // log(1 + x): 7-term series for |x| < 1/16 (truncation < 4e-11 relative),
  // the library routine otherwise
  __host__ __device__ static float log1p_(float x) {
    if (fabsf(x) < 0.0625f) {
      float p = 1.0f / 7.0f;
      p = p * x - 1.0f / 6.0f;
      p = p * x + 1.0f / 5.0f;
      p = p * x - 1.0f / 4.0f;
      p = p * x + 1.0f / 3.0f;
      p = p * x - 0.5f;
      p = p * x + 1.0f;
      return p * x;
    }
    return log1pf(x);
  }
  __host__ __device__ static float fma_(float a, float b, float c) { return fmaf(a, b, c); }
  // exp(x) - 1: 8-term series for |x| < 1/16 (truncation < 1e-13
  // relative), the library routine otherwise
  __host__ __device__ static float expm1_(float x) {
    if (fabsf(x) < 0.0625f) {
      float p = 1.0f / 5040.0f;
      p = p * x + 1.0f / 720.0f;
      p = p * x + 1.0f / 120.0f;
      p = p * x + 1.0f / 24.0f;
      p = p * x + 1.0f / 6.0f;
      p = p * x + 0.5f;
      p = p * x + 1.0f;
      return p * x;
    }
    return expm1f(x);
  }
  __host__ __device__ static float exp_(float x) { return expf(x); }
  // hardware square root (1 ulp) in the FP32 element path
  __host__ __device__ static float sqrt_(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
  }
  // hardware reciprocal (1 ulp) in the FP32 element path
  __host__ __device__ static float rcp_(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
  }
  // hardware reciprocal square root (1 ulp) in the FP32 element path
  __host__ __device__ static float rsqrt_(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
  }
  // x / c for a constant c: FP32 multiplies by the rounded reciprocal
  __host__ __device__ static float div_c(float x, double c) { return x * float(1.0 / c); }
};
template <> struct Num<double> {
  __host__ __device__ static double log1p_(double x) { return log1p(x); }
  __host__ __device__ static double fma_(double a, double b, double c) { return fma(a, b, c); }
  __host__ __device__ static double expm1_(double x) { return expm1(x); }
  __host__ __device__ static double exp_(double x) { return exp(x); }
  __host__ __device__ static double sqrt_(double x) { return sqrt(x); }
  __host__ __device__ static double rcp_(double x) { return 1.0 / x; }
  __host__ __device__ static double rsqrt_(double x) { return 1.0 / sqrt(x); }
  __host__ __device__ static double div_c(double x, double c) { return x / c; }
};

// Prony history in HBM, blocked by 64 elements: component c = 18 gp + k of
// element e at [((e / 64) * 144 + c) * 64 + e % 64].  A wavefront's 64
// elements read one contiguous 256-byte run per component (coalesced), and
// all 144 runs of the wavefront lie in one contiguous 36 KB block (FP32).
constexpr int kHistBlock = 64;
__host__ __device__ inline size_t hist_index(int c, int e) {
  return ((size_t)(e / kHistBlock) * 144 + c) * kHistBlock + e % kHistBlock;
}
// values in the history array of ne elements (whole blocks)
inline size_t hist_len(int ne) { return 144 * (((size_t)ne + kHistBlock - 1) / kHistBlock * kHistBlock); }

// Position p = i + 2j + 4k of the 2x2x2 node cube (i: xi, j: eta, k: zeta
// sign bit) -> C3D8 local node number.
__host__ __device__ constexpr int p2n(int p) {
  return p == 0 ? 0 : p == 1 ? 1 : p == 2 ? 3 : p == 3 ? 2 : p == 4 ? 4 : p == 5 ? 5 : p == 6 ? 7 : 6;
}

// Unnormalised Walsh-Hadamard transform over the three sign bits.
template <typename R>
__host__ __device__ inline void wht8(R v[8]) {
#pragma unroll
  for (int b = 1; b < 8; b <<= 1) {
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      if (p & b) continue;
      R lo = v[p], hi = v[p | b];
      v[p] = hi + lo;
      v[p | b] = hi - lo;
    }
  }
}

// Unnormalised Walsh-Hadamard coefficients of component d of a nodal field
// (v: the 8 nodal values in cube-position order, FP64, overwritten), rounded
// to R once: c[m][d] for m = xi, eta, xi.eta, zeta, xi.zeta, eta.zeta,
// xi.eta.zeta (the sum, index 0, is not needed).  Taken in FP64, the
// differences of nearby node positions lose nothing to their magnitude.
template <typename R>
__host__ __device__ inline void field_coeffs(double v[8], R c[7][3], int d) {
  wht8(v);
#pragma unroll
  for (int m = 0; m < 7; ++m) c[m][d] = R(v[m + 1]);
}

// Inverse direction used for forces: coefficients -> nodal values with sign
// convention (bit set -> +1, clear -> -1).
template <typename R>
__host__ __device__ inline void iwht8(R c[8]) {
#pragma unroll
  for (int b = 1; b < 8; b <<= 1) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (q & b) continue;
      R lo = c[q], hi = c[q | b];
      c[q] = lo - hi;
      c[q | b] = lo + hi;
    }
  }
}

// Symmetric 3x3 stored as (00, 11, 22, 12, 02, 01) -- Voigt order
// (11, 22, 33, 23, 13, 12) of CompressibleNeoHookean.cpp:45-55.
template <typename R> struct Sym { R v[6]; };

template <typename R>
__host__ __device__ inline R sym_at(const Sym<R> &s, int i, int j) {
  if (i == j) return s.v[i];
  int k = 3 - i - j;  // (1,2)->0 ... map: pair {1,2}->3, {0,2}->4, {0,1}->5
  return s.v[3 + k];
}

// det(I + H) - 1 without forming I + H (no cancellation at small strain).
template <typename R>
__host__ __device__ inline R det_ip1_m1(const R H[3][3]) {
  R tr = H[0][0] + H[1][1] + H[2][2];
  R m2 = (H[0][0] * H[1][1] - H[0][1] * H[1][0]) + (H[0][0] * H[2][2] - H[0][2] * H[2][0]) +
         (H[1][1] * H[2][2] - H[1][2] * H[2][1]);
  R d3 = H[0][0] * (H[1][1] * H[2][2] - H[1][2] * H[2][1]) -
         H[0][1] * (H[1][0] * H[2][2] - H[1][2] * H[2][0]) +
         H[0][2] * (H[1][0] * H[2][1] - H[1][1] * H[2][0]);
  return tr + m2 + d3;
}

// Green-Lagrange E = (H + H^T + H^T H) / 2.
template <typename R>
__host__ __device__ inline Sym<R> green_lagrange(const R H[3][3]) {
  Sym<R> E;
  const int I[6] = {0, 1, 2, 1, 0, 0}, J[6] = {0, 1, 2, 2, 2, 1};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    int i = I[k], j = J[k];
    R hth = H[0][i] * H[0][j] + H[1][i] * H[1][j] + H[2][i] * H[2][j];
    E.v[k] = R(0.5) * (H[i][j] + H[j][i] + hth);
  }
  return E;
}

// adj(C) for C = I + 2E (symmetric).
template <typename R>
__host__ __device__ inline Sym<R> adj_c(const Sym<R> &E) {
  R c00 = R(1) + R(2) * E.v[0], c11 = R(1) + R(2) * E.v[1], c22 = R(1) + R(2) * E.v[2];
  R c12 = R(2) * E.v[3], c02 = R(2) * E.v[4], c01 = R(2) * E.v[5];
  Sym<R> A;
  A.v[0] = c11 * c22 - c12 * c12;
  A.v[1] = c00 * c22 - c02 * c02;
  A.v[2] = c00 * c11 - c01 * c01;
  A.v[3] = c01 * c02 - c00 * c12;
  A.v[4] = c01 * c12 - c02 * c11;
  A.v[5] = c02 * c12 - c01 * c22;
  return A;
}

// S = (A T) * s for symmetric A, T whose product is symmetric (they commute:
// both are functions of E).
template <typename R>
__host__ __device__ inline Sym<R> sym_mul(const Sym<R> &A, const Sym<R> &T, R s) {
  Sym<R> S;
  const int I[6] = {0, 1, 2, 1, 0, 0}, J[6] = {0, 1, 2, 2, 2, 1};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    int i = I[k], j = J[k];
    R acc = sym_at(A, i, 0) * sym_at(T, 0, j) + sym_at(A, i, 1) * sym_at(T, 1, j) +
            sym_at(A, i, 2) * sym_at(T, 2, j);
    S.v[k] = acc * s;
  }
  return S;
}

// Whether an FP32 Jacobi rotation of the off-diagonal term apq of the
// diagonal pair (app, aqq) changes anything at FP32 round-off.
__host__ __device__ inline bool jacobi_active(float apq, float app, float aqq) {
  return fabsf(apq) > 1e-30f && fabsf(apq) > 1e-12f * (fabsf(app) + fabsf(aqq));
}

// Eigenvalues w and eigenvectors (columns of V) of a symmetric 3x3 matrix by
// cyclic Jacobi rotations, a fixed number of sweeps (converged to round-off
// for strain-sized matrices; repeated eigenvalues need no special case).
template <typename R, int SWEEPS>
__host__ __device__ inline void sym_eig3(const Sym<R> &m, R w[3], R V[3][3]) {
  R a[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      a[i][j] = sym_at(m, i, j);
      V[i][j] = i == j ? R(1) : R(0);
    }
#pragma unroll 1
  for (int sweep = 0; sweep < SWEEPS; ++sweep) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int p = r == 2 ? 1 : 0, q = r == 0 ? 1 : 2, o = 3 - p - q;
      const R apq = a[p][q];
      if constexpr (sizeof(R) == 4) {
        // The FP32 reciprocal of a vanishing (denormal) 2 apq is infinite,
        // and 0 * inf would poison the rotation: skip rotations of an
        // off-diagonal term below 1e-12 of the diagonal pair (far below FP32
        // round-off) or below 1e-30 (strains of no consequence).
        if (!jacobi_active(apq, a[p][p], a[q][q])) continue;
      } else {
        if (apq == R(0)) continue;
      }
      // (FP64 divides; FP32 uses the 1-ulp reciprocal and reciprocal square
      // root: the rotation stays orthogonal to round-off)
      const R theta = (a[q][q] - a[p][p]) * Num<R>::rcp_(R(2) * apq);
      // t = tan of the rotation angle, the smaller root (|t| <= 1)
      R t = Num<R>::rcp_(fabs(theta) + Num<R>::sqrt_(theta * theta + R(1)));
      if (theta < R(0)) t = -t;
      if (!(fabs(theta) < R(1e18))) t = R(0.5) * Num<R>::rcp_(theta);  // theta^2 would overflow
      const R c = Num<R>::rsqrt_(t * t + R(1)), sn = t * c;
      a[p][p] -= t * apq;
      a[q][q] += t * apq;
      a[p][q] = a[q][p] = R(0);
      const R aop = a[o][p], aoq = a[o][q];
      a[o][p] = a[p][o] = c * aop - sn * aoq;
      a[o][q] = a[q][o] = sn * aop + c * aoq;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const R vp = V[k][p], vq = V[k][q];
        V[k][p] = c * vp - sn * vq;
        V[k][q] = sn * vp + c * vq;
      }
    }
    if constexpr (sizeof(R) == 4) {
      // converged to FP32 round-off: the remaining sweeps would not rotate
      if (!jacobi_active(a[0][1], a[0][0], a[1][1]) && !jacobi_active(a[0][2], a[0][0], a[2][2]) &&
          !jacobi_active(a[1][2], a[1][1], a[2][2]))
        break;
    }
  }
  w[0] = a[0][0];
  w[1] = a[1][1];
  w[2] = a[2][2];
}

// b - I = H + H^T + H H^T from H = grad_X u (no cancellation at small strain).
template <typename R>
__host__ __device__ inline Sym<R> b_minus_identity(const R H[3][3]) {
  const int I[6] = {0, 1, 2, 1, 0, 0}, J[6] = {0, 1, 2, 2, 2, 1};
  Sym<R> bm;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int i = I[k], j = J[k];
    bm.v[k] = H[i][j] + H[j][i] + (H[i][0] * H[j][0] + H[i][1] * H[j][1] + H[i][2] * H[j][2]);
  }
  return bm;
}

// 1-term Ogden (materials 6 and 7; no reference counterpart): deviatoric
// energy (2 mu / alpha^2) sum_a (lbar_a^alpha - 1) on the isochoric principal
// stretches, volumetric part as the reference's HGO (K/2 (J^2 - 1) in tau,
// K = lambda + 2 mu / 3).  Kirchhoff stress tau = sum_a tau_a n_a n_a + p I
// with tau_a = (2 mu / alpha) (lbar_a^alpha - mean_b lbar_b^alpha); the
// stretches come from the eigenvalues of b - I (ln lbar_a = log1p(.) / 2 -
// ln J / 3, lbar^alpha - 1 by expm1), so small strains keep full precision.
// Returns p separately (the Prony update acts on the deviatoric part).
template <typename R>
__host__ __device__ inline void ogden_tau(const PartDev &pp, const Sym<R> &bm, R Jm1, Sym<R> &tau, R &p) {
  const R mu = R(pp.mu), alpha = R(pp.alpha);
  const R K = R(pp.lambda) + Num<R>::div_c(R(2) * mu, 3.0);
  const R lnJ3 = Num<R>::div_c(Num<R>::log1p_(Jm1), 3.0);
  R w[3], V[3][3];
  sym_eig3<R, sizeof(R) == 4 ? 4 : 6>(bm, w, V);
  R e[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) e[a] = Num<R>::expm1_(alpha * (R(0.5) * Num<R>::log1p_(w[a]) - lnJ3));
  const R mean = Num<R>::div_c(e[0] + e[1] + e[2], 3.0);
  const R f = R(2) * mu * Num<R>::rcp_(alpha);
  R t[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) t[a] = f * (e[a] - mean);
  const int I[6] = {0, 1, 2, 1, 0, 0}, J[6] = {0, 1, 2, 2, 2, 1};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int i = I[k], j = J[k];
    tau.v[k] = t[0] * V[i][0] * V[j][0] + t[1] * V[i][1] * V[j][1] + t[2] * V[i][2] * V[j][2];
  }
  p = R(0.5) * K * Jm1 * (Jm1 + R(2));
}

struct ViscoCoef { double c11, c12, c21, c22; };

// PK2 stress of the reference's material routines, restated in terms of the
// displacement gradient H = F - I.  hist: 18 values (H1, H2, S0 as Voigt
// 6-vectors) for material 5 or nullptr.  MAT >= 0 fixes the material at
// compile time; MAT < 0 dispatches on pp.mat.  Returns false for an unknown
// id.
template <typename R, int MAT = -1>
__host__ __device__ inline bool pk2_stress(const PartDev &pp, const R H[3][3], R (&hist)[18], bool use_hist,
                                  const ViscoCoef &vc, Sym<R> &S, R &Jm1) {
  Jm1 = det_ip1_m1(H);
  const int mat = MAT >= 0 ? MAT : pp.mat;
  if (mat == 0) {  // rigid: StressUpdate.cpp:8-9
#pragma unroll
    for (int k = 0; k < 6; ++k) S.v[k] = R(0);
    return true;
  }
  const R mu = R(pp.mu), lambda = R(pp.lambda);
  if (mat == 3) {
    // LinearElastic.cpp:14-70: P = lambda tr(eps) I + 2 mu eps, S = F^-1 P,
    // unsymmetrised, upper triangle kept.
    R trE = H[0][0] + H[1][1] + H[2][2];
    R P[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) P[i][j] = mu * (H[i][j] + H[j][i]);
    P[0][0] += lambda * trE;
    P[1][1] += lambda * trE;
    P[2][2] += lambda * trE;
    R F[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) F[i][j] = H[i][j] + (i == j ? R(1) : R(0));
    R A[3][3];  // adj(F)
    A[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
    A[0][1] = F[0][2] * F[2][1] - F[0][1] * F[2][2];
    A[0][2] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
    A[1][0] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
    A[1][1] = F[0][0] * F[2][2] - F[0][2] * F[2][0];
    A[1][2] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
    A[2][0] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
    A[2][1] = F[0][1] * F[2][0] - F[0][0] * F[2][1];
    A[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
    R iJ = Num<R>::rcp_(R(1) + Jm1);
    const int I[6] = {0, 1, 2, 1, 0, 0}, J[6] = {0, 1, 2, 2, 2, 1};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      int i = I[k], j = J[k];
      S.v[k] = (A[i][0] * P[0][j] + A[i][1] * P[1][j] + A[i][2] * P[2][j]) * iJ;
    }
    return true;
  }
  if (mat == 6 || mat == 7) {
    // 1-term Ogden (ogden_tau), pulled back: S = F^-1 tau F^-T.  Material 7
    // adds the two-term Prony update of HGOIsotropicViscoelastic.cpp:94-159 on
    // the isochoric part Sdev = F^-1 dev(tau) F^-T.
    Sym<R> td;
    R p;
    ogden_tau(pp, b_minus_identity(H), Jm1, td, p);
    R Fi[3][3];  // adj(F) / J
    {
      R F[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[i][j] = H[i][j] + (i == j ? R(1) : R(0));
      const R iJ = Num<R>::rcp_(R(1) + Jm1);
      Fi[0][0] = (F[1][1] * F[2][2] - F[1][2] * F[2][1]) * iJ;
      Fi[0][1] = (F[0][2] * F[2][1] - F[0][1] * F[2][2]) * iJ;
      Fi[0][2] = (F[0][1] * F[1][2] - F[0][2] * F[1][1]) * iJ;
      Fi[1][0] = (F[1][2] * F[2][0] - F[1][0] * F[2][2]) * iJ;
      Fi[1][1] = (F[0][0] * F[2][2] - F[0][2] * F[2][0]) * iJ;
      Fi[1][2] = (F[0][2] * F[1][0] - F[0][0] * F[1][2]) * iJ;
      Fi[2][0] = (F[1][0] * F[2][1] - F[1][1] * F[2][0]) * iJ;
      Fi[2][1] = (F[0][1] * F[2][0] - F[0][0] * F[2][1]) * iJ;
      Fi[2][2] = (F[0][0] * F[1][1] - F[0][1] * F[1][0]) * iJ;
    }
    const int I[6] = {0, 1, 2, 1, 0, 0}, J[6] = {0, 1, 2, 2, 2, 1};
    Sym<R> Sdev, Ci;  // F^-1 dev(tau) F^-T and C^-1 = F^-1 F^-T
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int i = I[k], j = J[k];
      R sd = R(0), ci = R(0);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        ci += Fi[i][a] * Fi[j][a];
#pragma unroll
        for (int b = 0; b < 3; ++b) sd += Fi[i][a] * sym_at(td, a, b) * Fi[j][b];
      }
      Sdev.v[k] = sd;
      Ci.v[k] = ci;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) S.v[k] = Sdev.v[k] + p * Ci.v[k];
    if (mat == 7 && use_hist) {
      const R c11 = R(vc.c11), c12 = R(vc.c12), c21 = R(vc.c21), c22 = R(vc.c22);
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        R d = Sdev.v[k] - hist[12 + k];
        R h1 = c11 * hist[k] + c21 * d;
        R h2 = c12 * hist[6 + k] + c22 * d;
        hist[k] = h1;
        hist[6 + k] = h2;
        hist[12 + k] = Sdev.v[k];
        S.v[k] = S.v[k] + h1 + h2;
      }
    }
    return true;
  }
  Sym<R> E = green_lagrange(H);
  R trE = E.v[0] + E.v[1] + E.v[2];
  if (mat == 2) {  // StVenantKirchhoff.cpp:24-41
#pragma unroll
    for (int k = 0; k < 6; ++k) S.v[k] = R(2) * mu * E.v[k];
    S.v[0] += lambda * trE;
    S.v[1] += lambda * trE;
    S.v[2] += lambda * trE;
    return true;
  }
  const R lnJ = Num<R>::log1p_(Jm1);
  const R J = R(1) + Jm1;
  const R iJ2 = Num<R>::rcp_(J * J);
  Sym<R> Ac = adj_c(E);  // C^-1 = Ac / J^2
  if (mat == 1) {
    // CompressibleNeoHookean.cpp:26-55: S = mu (I - C^-1) + lambda ln J C^-1
    //                                     = C^-1 (2 mu E + lambda ln J I)
    Sym<R> T;
#pragma unroll
    for (int k = 0; k < 6; ++k) T.v[k] = R(2) * mu * E.v[k];
    T.v[0] += lambda * lnJ;
    T.v[1] += lambda * lnJ;
    T.v[2] += lambda * lnJ;
    S = sym_mul(Ac, T, iJ2);
    return true;
  }
  if (mat == 4 || mat == 5) {
    // HGOIsotropic.cpp:37-90 with kappa = 1/3:
    //   S = J F^-1 sigma F^-T = C^-1 [J^-2/3 (mu + fib) dev(2E) + K/2 (J^2-1) I]
    const R kappa = R(1.0 / 3.0);
    const R K = lambda + Num<R>::div_c(R(2) * mu, 3.0);
    const R Jm23m1 = Num<R>::expm1_(R(-2.0 / 3.0) * lnJ);
    const R Jm23 = R(1) + Jm23m1;
    const R Ea = kappa * (R(3) * Jm23m1 + R(2) * Jm23 * trE);
    R fib = R(0);
    if (Ea > R(0)) fib = R(2) * R(pp.k1) * Num<R>::exp_(R(pp.k2) * Ea * Ea) * Ea * kappa;
    const R a = Jm23 * (mu + fib);
    const R tr3 = Num<R>::div_c(R(2) * trE, 3.0);
    Sym<R> Tdev;
#pragma unroll
    for (int k = 0; k < 6; ++k) Tdev.v[k] = a * R(2) * E.v[k];
    Tdev.v[0] -= a * tr3;
    Tdev.v[1] -= a * tr3;
    Tdev.v[2] -= a * tr3;
    Sym<R> Sdev = sym_mul(Ac, Tdev, iJ2);
    const R pvol = R(0.5) * K * Jm1 * (Jm1 + R(2)) * iJ2;
    S.v[0] = Sdev.v[0] + pvol * Ac.v[0];
    S.v[1] = Sdev.v[1] + pvol * Ac.v[1];
    S.v[2] = Sdev.v[2] + pvol * Ac.v[2];
    S.v[3] = Sdev.v[3] + pvol * Ac.v[3];
    S.v[4] = Sdev.v[4] + pvol * Ac.v[4];
    S.v[5] = Sdev.v[5] + pvol * Ac.v[5];
    if (mat == 5 && use_hist) {
      // HGOIsotropicViscoelastic.cpp:94-159: Sdev = S - (C:S)/3 C^-1 is the
      // isochoric part above; H_j <- c1j H_j + c2j (Sdev - S0); S += H1 + H2.
      const R c11 = R(vc.c11), c12 = R(vc.c12), c21 = R(vc.c21), c22 = R(vc.c22);
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        R d = Sdev.v[k] - hist[12 + k];
        R h1 = c11 * hist[k] + c21 * d;
        R h2 = c12 * hist[6 + k] + c22 * d;
        hist[k] = h1;
        hist[6 + k] = h2;
        hist[12 + k] = Sdev.v[k];
        S.v[k] = S.v[k] + h1 + h2;
      }
    }
    return true;
  }
  return false;
}

// Kirchhoff stress tau = F S F^T = J sigma of the neo-Hookean (MAT 1) and
// isotropic HGO (MAT 4) laws, from H = grad_X u:
//   neo-Hookean (CompressibleNeoHookean.cpp:26-55): tau = mu (b - I) + lambda ln J I
//   HGO (HGOIsotropic.cpp:37-90, kappa = 1/3):
//     tau = J^-2/3 (mu + fib) dev(b) + K/2 (J^2 - 1) I,  K = lambda + 2 mu / 3
// with b - I = H + H^T + H H^T formed directly (no cancellation at small
// strain).  Equal to F S F^T of pk2_stress in exact arithmetic.
//
// Also the 1-term Ogden law (MAT 6, ogden_tau) and the anisotropic HGO law
// (MAT 8; no reference counterpart): W = mu/2 (I1bar - 3) + K/4 (J^2 - 1 -
// 2 ln J) + k1/(2 k2) (exp(k2 E^2) - 1) for E = kappa (I1bar - 3) +
// (1 - 3 kappa)(I4bar - 1) > 0, I4bar = J^-2/3 |F a|^2 for the unit fibre a:
//   tau = dev[(mu + 2 psi kappa) bbar + 2 psi (1 - 3 kappa) abar abar] + K/2 (J^2 - 1) I
// with psi = k1 E exp(k2 E^2); kappa = 1/3 gives the HGO law above.
template <typename R, int MAT>
__host__ __device__ inline void kirchhoff_stress(const PartDev &pp, const R H[3][3], R Jm1, const R fib[3],
                                                 Sym<R> &tau) {
  const R mu = R(pp.mu), lambda = R(pp.lambda);
  const Sym<R> bm = b_minus_identity(H);
  if (MAT == 6) {
    R p;
    ogden_tau(pp, bm, Jm1, tau, p);
    tau.v[0] += p;
    tau.v[1] += p;
    tau.v[2] += p;
    return;
  }
  const R lnJ = Num<R>::log1p_(Jm1);
  if (MAT == 8) {
    const R kappa = R(pp.kappa);
    const R K = lambda + Num<R>::div_c(R(2) * mu, 3.0);
    const R Jm23m1 = Num<R>::expm1_(R(-2.0 / 3.0) * lnJ);
    const R Jm23 = R(1) + Jm23m1;
    const R trb = bm.v[0] + bm.v[1] + bm.v[2];
    R fa[3], ha[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      ha[i] = H[i][0] * fib[0] + H[i][1] * fib[1] + H[i][2] * fib[2];
      fa[i] = fib[i] + ha[i];
    }
    // |F a|^2 - 1 = 2 a.Ha + |Ha|^2
    const R s4 = R(2) * (fib[0] * ha[0] + fib[1] * ha[1] + fib[2] * ha[2]) +
                 (ha[0] * ha[0] + ha[1] * ha[1] + ha[2] * ha[2]);
    const R I4m1 = Jm23m1 * (R(1) + s4) + s4;
    const R I1m3 = R(3) * Jm23m1 + Jm23 * trb;
    const R E = kappa * I1m3 + (R(1) - R(3) * kappa) * I4m1;
    R psi = R(0);
    if (E > R(0)) psi = R(pp.k1) * E * Num<R>::exp_(R(pp.k2) * E * E);
    const R c1 = Jm23 * (mu + R(2) * psi * kappa);
    const R c4 = Jm23 * R(2) * psi * (R(1) - R(3) * kappa);
    const R p = R(0.5) * K * Jm1 * (Jm1 + R(2)) - Num<R>::div_c(c1 * trb, 3.0) - Num<R>::div_c(c4 * (R(1) + s4), 3.0);
    const int I[6] = {0, 1, 2, 1, 0, 0}, J[6] = {0, 1, 2, 2, 2, 1};
#pragma unroll
    for (int k = 0; k < 6; ++k) tau.v[k] = c1 * bm.v[k] + c4 * fa[I[k]] * fa[J[k]];
    tau.v[0] += p;
    tau.v[1] += p;
    tau.v[2] += p;
    return;
  }
  if (MAT == 1) {
#pragma unroll
    for (int k = 0; k < 6; ++k) tau.v[k] = mu * bm.v[k];
    tau.v[0] += lambda * lnJ;
    tau.v[1] += lambda * lnJ;
    tau.v[2] += lambda * lnJ;
  } else {
    const R kappa = R(1.0 / 3.0);
    const R K = lambda + Num<R>::div_c(R(2) * mu, 3.0);
    const R Jm23m1 = Num<R>::expm1_(R(-2.0 / 3.0) * lnJ);
    const R Jm23 = R(1) + Jm23m1;
    const R trb = bm.v[0] + bm.v[1] + bm.v[2];  // tr(b - I) = tr(C - I)
    const R Ea = kappa * (R(3) * Jm23m1 + Jm23 * trb);
    R fib = R(0);
    if (Ea > R(0)) fib = R(2) * R(pp.k1) * Num<R>::exp_(R(pp.k2) * Ea * Ea) * Ea * kappa;
    const R a = Jm23 * (mu + fib);
    const R p = R(0.5) * K * Jm1 * (Jm1 + R(2)) - Num<R>::div_c(a * trb, 3.0);
#pragma unroll
    for (int k = 0; k < 6; ++k) tau.v[k] = a * bm.v[k];
    tau.v[0] += p;
    tau.v[1] += p;
    tau.v[2] += p;
  }
}

template <typename R>
__host__ __device__ inline R norm_cross(const R a[3], const R b[3]) {
  // normOfCrossProduct, math.cpp:30-43
  R z = a[0] * b[1] - a[1] * b[0];
  R x = a[1] * b[2] - a[2] * b[1];
  R y = -a[0] * b[2] + a[2] * b[0];
  return Num<R>::sqrt_(x * x + y * y + z * z);
}

template <typename R>
__host__ __device__ inline R triple(const R s[3], const R a[3], const R b[3]) {
  // tripleProduct, math.cpp:52-56
  return s[2] * (a[0] * b[1] - a[1] * b[0]) + s[0] * (a[1] * b[2] - a[2] * b[1]) -
         s[1] * (a[0] * b[2] - a[2] * b[0]);
}

template <typename R>
__host__ __device__ inline void cross3(const R a[3], const R b[3], R c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = -a[0] * b[2] + a[2] * b[0];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename R>
__host__ __device__ inline R norm3(const R x[3]) {
  return Num<R>::sqrt_(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
}

// areaHexahedronFace (Geometry.cpp:29-64) from a face's twist cD and edge
// vectors c1, c2 (each the quarter-sum of the reference's corner formula).
// The 2x2 rule sums |v1 x v2| with v1 = q_j cD + c1, v2 = q_i cD + c2; since
// cD x cD = 0 that is |c1 x c2 + q_i (c1 x cD) + q_j (cD x c2)|, so three cross
// products serve all four points.
// `tol` is the reference's 1e-6, times the scale of the vectors (see
// char_length).
template <typename R>
__host__ __device__ inline R face_area(const R cD[3], const R c1[3], const R c2[3], const R tol = R(1e-6)) {
  R w0[3];
  cross3(c1, c2, w0);
  if ((cD[0] < tol) && (cD[1] < tol) && (cD[2] < tol)) return R(4) * norm3(w0);
  R w1[3], w2[3];
  cross3(c1, cD, w1);
  cross3(cD, c2, w2);
  const R t = R(kFaceQ);
  R area = R(0);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    R b[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) b[k] = w0[k] + (i ? t : -t) * w1[k];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      R x[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) x[k] = b[k] + (j ? t : -t) * w2[k];
      area += norm3(x);
    }
  }
  return area;
}

// Characteristic length V / max face area (CalculateCharacteristicLength_
// C3D8.cpp:3-30) from the unnormalised Walsh-Hadamard coefficients of the
// deformed position field (8 x the coefficients of its monomial expansion):
// d[m][comp] for m = xi, eta, xi.eta, zeta, xi.zeta, eta.zeta, xi.eta.zeta.
// These are the reference's q0..q5 of volumeHexahedron as they stand (q0..q5
// = xi.eta, xi.zeta, xi, eta.zeta, eta, zeta).  The face vectors are 8 x the
// normalised ones, so the areas come out 64 x too large and the tolerance
// 8 x; both factors are powers of two and exact.
template <typename R>
__host__ __device__ inline R char_length(const R d[7][3]) {
  const R *c1 = d[0], *c2 = d[1], *c3 = d[2], *c4 = d[3], *c5 = d[4], *c6 = d[5], *c7 = d[6];
  const R *q0 = c3, *q1 = c5, *q2 = c1, *q3 = c6, *q4 = c2, *q5 = c4;
  // FP64 keeps the reference's divisions; FP32 multiplies by reciprocals
  // (its minimum element is recomputed in FP64 where it matters).
  const R t123 = triple(q0, q4, q3) + triple(q2, q0, q1) + triple(q1, q3, q5), t4 = triple(q2, q4, q5);
  const R vol = sizeof(R) == 8 ? t123 / R(192) + t4 / R(64) : t123 * R(1.0 / 192) + t4 * R(1.0 / 64);
  // Faces in the reference's order (CalculateCharacteristicLength_C3D8.cpp:12-13):
  // zeta-, zeta+, xi-, xi+, eta-, eta+.
  R amax = R(0);
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    R cD[3], e1[3], e2[3];
    const R s = (f & 1) ? R(1) : R(-1);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (f < 2) {
        cD[k] = c3[k] + s * c7[k];
        e1[k] = c1[k] + s * c5[k];
        e2[k] = c2[k] + s * c6[k];
      } else if (f < 4) {
        cD[k] = c6[k] + s * c7[k];
        e1[k] = c2[k] + s * c3[k];
        e2[k] = c4[k] + s * c5[k];
      } else {
        cD[k] = c5[k] + s * c7[k];
        e1[k] = c1[k] + s * c3[k];
        e2[k] = c4[k] + s * c6[k];
      }
    }
    R fa = face_area(cD, e1, e2, R(8) * R(1e-6));
    if (fa > amax) amax = fa;
  }
  return sizeof(R) == 8 ? vol / amax * R(64) : vol * Num<R>::rcp_(amax) * R(64);
}

// A lower bound of char_length(d) that needs one square root instead of 24:
// the same volume over an upper bound of the largest face area.  Each face's
// 2x2 rule sums |v1 x v2| <= |v1| |v2| over its four points, and by
// Cauchy-Schwarz sum |v1| |v2| <= sqrt(sum |v1|^2 sum |v2|^2), where (the
// odd terms cancel between the points q = +-1/sqrt(3)) sum |v1|^2 = 4 (|c1|^2
// + |cD|^2 / 3) and sum |v2|^2 = 4 (|c2|^2 + |cD|^2 / 3); the parallelogram
// branch, 4 |c1 x c2|, is below the same value.  Within 1 % of the exact
// length on near-cubic elements (0.992 median, 0.89 worst over the 1M-hex
// jittered block).
template <typename R>
__host__ __device__ inline R char_length_lower(const R d[7][3]) {
  const R *c1 = d[0], *c2 = d[1], *c3 = d[2], *c4 = d[3], *c5 = d[4], *c6 = d[5], *c7 = d[6];
  const R *q0 = c3, *q1 = c5, *q2 = c1, *q3 = c6, *q4 = c2, *q5 = c4;
  const R t123 = triple(q0, q4, q3) + triple(q2, q0, q1) + triple(q1, q3, q5), t4 = triple(q2, q4, q5);
  const R vol = t123 * R(1.0 / 192) + t4 * R(1.0 / 64);
  // The two faces of a pair share their vectors up to the sign of one term,
  // |a +- b|^2 = |a|^2 +- 2 a.b + |b|^2: seven squared norms and nine dot
  // products serve all six faces (pairs zeta-/+, xi-/+, eta-/+).
  auto dot = [](const R *a, const R *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  const R n1 = dot(c1, c1), n2 = dot(c2, c2), n3 = dot(c3, c3), n4 = dot(c4, c4), n5 = dot(c5, c5),
          n6 = dot(c6, c6), n7 = dot(c7, c7);
  const R pairs[3][9] = {{n3, dot(c3, c7), n1, dot(c1, c5), n5, n2, dot(c2, c6), n6, n7},
                         {n6, dot(c6, c7), n2, dot(c2, c3), n3, n4, dot(c4, c5), n5, n7},
                         {n5, dot(c5, c7), n1, dot(c1, c3), n3, n4, dot(c4, c6), n6, n7}};
  R bmax = R(0);
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    const R *p = pairs[g];
#pragma unroll
    for (int sg = 0; sg < 2; ++sg) {
      const R s2 = sg ? R(2) : R(-2);
      const R nd = p[0] + s2 * p[1] + p[8];  // |cD|^2
      const R ne1 = p[2] + s2 * p[3] + p[4];  // |e1|^2
      const R ne2 = p[5] + s2 * p[6] + p[7];  // |e2|^2
      const R b = (ne1 + nd * R(1.0 / 3)) * (ne2 + nd * R(1.0 / 3));
      if (b > bmax) bmax = b;
    }
  }
  // area <= 4 sqrt(bmax) in the 8x units of char_length (x 64 below)
  return vol * Num<R>::rsqrt_(bmax) * R(16);
}

// One piece of the characteristic length of a deformed hex from its
// normalised coefficients d[m][comp]: face area k (k < 6, faces in the order
// of CalculateCharacteristicLength_C3D8.cpp:12-13) or the volume (k == 6).
template <typename R>
__host__ __device__ inline R hex_part(const R d[7][3], int k) {
  if (k == 6) {
    R q0[3], q1[3], q2[3], q3[3], q4[3], q5[3];
    for (int j = 0; j < 3; ++j) {
      q0[j] = R(8) * d[2][j];
      q1[j] = R(8) * d[4][j];
      q2[j] = R(8) * d[0][j];
      q3[j] = R(8) * d[5][j];
      q4[j] = R(8) * d[1][j];
      q5[j] = R(8) * d[3][j];
    }
    return (triple(q0, q4, q3) + triple(q2, q0, q1) + triple(q1, q3, q5)) / R(192) +
           triple(q2, q4, q5) / R(64);
  }
  const R s = (k & 1) ? R(1) : R(-1);
  const int g = k >> 1;
  R cD[3], e1[3], e2[3];
  for (int j = 0; j < 3; ++j) {
    if (g == 0) {
      cD[j] = d[2][j] + s * d[6][j];
      e1[j] = d[0][j] + s * d[4][j];
      e2[j] = d[1][j] + s * d[5][j];
    } else if (g == 1) {
      cD[j] = d[5][j] + s * d[6][j];
      e1[j] = d[1][j] + s * d[2][j];
      e2[j] = d[3][j] + s * d[4][j];
    } else {
      cD[j] = d[4][j] + s * d[6][j];
      e1[j] = d[0][j] + s * d[2][j];
      e2[j] = d[3][j] + s * d[5][j];
    }
  }
  return face_area(cD, e1, e2);
}

// f_a,i = sum_gp sum_k Q_ik dN_a/dxi_k accumulated in the monomial basis of
// the element (1, xi, eta, xi.eta, zeta, xi.zeta, eta.zeta, xi.eta.zeta; the
// constant term is zero).  w holds this Gauss point's weights of the seven
// monomials: the signs of its coordinates times the scale of the monomial
// class (linear, bilinear, trilinear), so h needs no scaling afterwards.
template <typename R> struct GpWeights { R l, x, y, z, yz, xz, xy; };
template <typename R>
__host__ __device__ inline void accumulate_force_basis(R h[7], R q0, R q1, R q2, const GpWeights<R> &w) {
  h[0] = Num<R>::fma_(w.l, q0, h[0]);
  h[1] = Num<R>::fma_(w.l, q1, h[1]);
  h[3] = Num<R>::fma_(w.l, q2, h[3]);
  h[2] = Num<R>::fma_(w.y, q0, h[2]);
  h[2] = Num<R>::fma_(w.x, q1, h[2]);
  h[4] = Num<R>::fma_(w.z, q0, h[4]);
  h[4] = Num<R>::fma_(w.x, q2, h[4]);
  h[5] = Num<R>::fma_(w.z, q1, h[5]);
  h[5] = Num<R>::fma_(w.y, q2, h[5]);
  h[6] = Num<R>::fma_(w.yz, q0, h[6]);
  h[6] = Num<R>::fma_(w.xz, q1, h[6]);
  h[6] = Num<R>::fma_(w.xy, q2, h[6]);
}

// c with its sign flipped when neg (0 or 1).  Integer arithmetic on a
// Gauss-point index that is uniform over the wavefront stays on the scalar
// unit, and the result is a scalar operand of the vector instructions.
template <typename R> struct SignBit;
template <> struct SignBit<float> {
  using U = unsigned;
  static constexpr int kShift = 31;
};
template <> struct SignBit<double> {
  using U = unsigned long long;
  static constexpr int kShift = 63;
};
template <typename R>
__host__ __device__ inline R flip_sign(R c, unsigned neg) {
  using U = typename SignBit<R>::U;
  return __builtin_bit_cast(R, __builtin_bit_cast(U, c) ^ ((U)neg << SignBit<R>::kShift));
}

// ---------------------------------------------------------------------------
// Constitutive surrogate of materials 7 and 8 (no reference counterpart).
// The network replaces the isochoric part S_iso of the PK2 stress (Ogden's
// eigen-decomposition, the fibre law); the volumetric part p C^-1 and the
// Prony recurrence on S_iso stay analytic.  Inputs and outputs are in the
// material frame: the fibre and a right-handed completion (material 8), the
// reference axes (material 7).
// ---------------------------------------------------------------------------
constexpr int kSurrFeat = 16;   // features per sample (12 used)
constexpr int kSurrOut = 8;     // outputs per sample (6 used)
// normalisation block: in mean [16], in scale [16], box low [16], box high
// [16] (raw features), out mean [8], out scale [8]
constexpr int kSurrNorm = 80;

template <typename R>
__host__ __device__ inline void material_frame(int mat, const R fib[3], R Q[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Q[i][j] = i == j ? R(1) : R(0);
  if (mat != 8 || !fib) return;
  const R a[3] = {fib[0], fib[1], fib[2]};
  // the reference axis least aligned with the fibre
  const R ax = fabs(a[0]), ay = fabs(a[1]), az = fabs(a[2]);
  const int k = (ax <= ay && ax <= az) ? 0 : (ay <= az ? 1 : 2);
  R b[3] = {k == 1 ? -a[2] : (k == 2 ? a[1] : R(0)), k == 0 ? a[2] : (k == 2 ? -a[0] : R(0)),
            k == 0 ? -a[1] : (k == 1 ? a[0] : R(0))};  // a x e_k
  const R nb = Num<R>::sqrt_(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  b[0] /= nb;
  b[1] /= nb;
  b[2] /= nb;
  const R c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Q[i][0] = a[i];
    Q[i][1] = b[i];
    Q[i][2] = c[i];
  }
}

// Q^T S Q (to = true: reference -> material frame) or Q S Q^T.
template <typename R>
__host__ __device__ inline Sym<R> rotate_sym(const Sym<R> &S, const R Q[3][3], bool to_material) {
  const int I[6] = {0, 1, 2, 1, 0, 0}, J[6] = {0, 1, 2, 2, 2, 1};
  Sym<R> out;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int i = I[k], j = J[k];
    R acc = R(0);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const R qa = to_material ? Q[a][i] : Q[i][a];
        const R qb = to_material ? Q[b][j] : Q[j][b];
        acc += qa * sym_at(S, a, b) * qb;
      }
    out.v[k] = acc;
  }
  return out;
}

// F^-1 = adj(I + H) / J.
template <typename R>
__host__ __device__ inline void inverse_f(const R H[3][3], R Jm1, R Fi[3][3]) {
  R F[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) F[i][j] = H[i][j] + (i == j ? R(1) : R(0));
  const R iJ = Num<R>::rcp_(R(1) + Jm1);
  Fi[0][0] = (F[1][1] * F[2][2] - F[1][2] * F[2][1]) * iJ;
  Fi[0][1] = (F[0][2] * F[2][1] - F[0][1] * F[2][2]) * iJ;
  Fi[0][2] = (F[0][1] * F[1][2] - F[0][2] * F[1][1]) * iJ;
  Fi[1][0] = (F[1][2] * F[2][0] - F[1][0] * F[2][2]) * iJ;
  Fi[1][1] = (F[0][0] * F[2][2] - F[0][2] * F[2][0]) * iJ;
  Fi[1][2] = (F[0][2] * F[1][0] - F[0][0] * F[1][2]) * iJ;
  Fi[2][0] = (F[1][0] * F[2][1] - F[1][1] * F[2][0]) * iJ;
  Fi[2][1] = (F[0][1] * F[2][0] - F[0][0] * F[2][1]) * iJ;
  Fi[2][2] = (F[0][0] * F[1][1] - F[0][1] * F[1][0]) * iJ;
}

// Isochoric PK2 stress S_iso = F^-1 tau_iso F^-T of material 7 (Ogden) or
// 8 (anisotropic HGO) and the volumetric factor p = K/2 (J^2 - 1): the
// elastic PK2 stress is S_iso + p C^-1 (as in pk2_stress and
// kirchhoff_stress).
template <typename R, int MAT>
__host__ __device__ inline void iso_pk2(const PartDev &pp, const R H[3][3], R Jm1, const R fib[3], Sym<R> &Siso,
                                        R &p) {
  Sym<R> tau;
  if (MAT == 7) {
    ogden_tau(pp, b_minus_identity(H), Jm1, tau, p);
  } else {
    kirchhoff_stress<R, 8>(pp, H, Jm1, fib, tau);
    const R K = R(pp.lambda) + Num<R>::div_c(R(2) * R(pp.mu), 3.0);
    p = R(0.5) * K * Jm1 * (Jm1 + R(2));
    tau.v[0] -= p;
    tau.v[1] -= p;
    tau.v[2] -= p;
  }
  R Fi[3][3];
  inverse_f(H, Jm1, Fi);
  const int I[6] = {0, 1, 2, 1, 0, 0}, J[6] = {0, 1, 2, 2, 2, 1};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int i = I[k], j = J[k];
    R acc = R(0);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) acc += Fi[i][a] * sym_at(tau, a, b) * Fi[j][b];
    Siso.v[k] = acc;
  }
}

// Raw features of a Gauss point, from Dbar = Cbar - I in the material frame
// (Cbar = J^-2/3 C formed from H without cancellation) and its Frobenius
// norm s: the direction Dbar / s (Voigt 11, 22, 33, 23, 13, 12), s, J - 1,
// Ibar1 - 3, Ibar2 - 3, Ibar4 - 1 (frame axis 1: the fibre), the law (0:
// material 7, 1: material 8).  The network predicts S_iso / s, so small
// strains keep the network's relative precision.
constexpr double kSurrSmin = 1e-12;
template <typename R>
__host__ __device__ inline void surr_features(const R H[3][3], R Jm1, const R Q[3][3], int mat, R f[kSurrFeat]) {
  R D0[3][3];  // C - I = H + H^T + H^T H
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) D0[i][j] = H[i][j] + H[j][i] + (H[0][i] * H[0][j] + H[1][i] * H[1][j] + H[2][i] * H[2][j]);
  Sym<R> D;
  {
    const int I[6] = {0, 1, 2, 1, 0, 0}, J[6] = {0, 1, 2, 2, 2, 1};
#pragma unroll
    for (int k = 0; k < 6; ++k) D.v[k] = D0[I[k]][J[k]];
  }
  D = rotate_sym(D, Q, true);
  const R Jm23m1 = Num<R>::expm1_(R(-2.0 / 3.0) * Num<R>::log1p_(Jm1));
  const R Jm23 = R(1) + Jm23m1;
#pragma unroll
  for (int k = 0; k < 6; ++k) D.v[k] *= Jm23;
  D.v[0] += Jm23m1;
  D.v[1] += Jm23m1;
  D.v[2] += Jm23m1;
  const R tr = D.v[0] + D.v[1] + D.v[2];
  const R tr2 = D.v[0] * D.v[0] + D.v[1] * D.v[1] + D.v[2] * D.v[2] +
                R(2) * (D.v[3] * D.v[3] + D.v[4] * D.v[4] + D.v[5] * D.v[5]);
  const R sn = Num<R>::sqrt_(tr2);
  const R is = sn > R(kSurrSmin) ? R(1) / sn : R(0);
#pragma unroll
  for (int k = 0; k < 6; ++k) f[k] = D.v[k] * is;
  f[6] = sn;
  f[7] = Jm1;
  f[8] = tr;
  f[9] = R(2) * tr + R(0.5) * (tr * tr - tr2);
  f[10] = D.v[0];
  f[11] = mat == 8 ? R(1) : R(0);
#pragma unroll
  for (int k = 12; k < kSurrFeat; ++k) f[k] = R(0);
}

// Per-element inputs of the surrogate stress: the network outputs of its 8
// Gauss points (normalised, stride gp_stride floats between Gauss points),
// the normalisation block, and whether this element is audited.
struct SurrIn {
  const float *y;
  size_t gp_stride;
  const float *norm;
  bool audit;
};

// Per-thread tallies: audited squared error of S_iso, squared norms of the
// analytic S_iso and full PK2, fallback Gauss points, surrogate Gauss
// points, audited Gauss points.
struct SurrTally {
  double v[6];
};

// PK2 stress of material 7 or 8 with the network's S_iso (the analytic law
// outside the training box) plus p C^-1 and, for material 7, the Prony
// update on S_iso (HGOIsotropicViscoelastic.cpp:94-159 semantics).
template <typename R, int MAT>
__host__ __device__ inline void surr_pk2(const PartDev &pp, const R H[3][3], const R fib[3], const SurrIn &in, int gp,
                                         R (&hist)[18], bool use_hist, const ViscoCoef &vc, Sym<R> &S, R &Jm1,
                                         SurrTally &t) {
  Jm1 = det_ip1_m1(H);
  R Q[3][3];
  material_frame<R>(MAT, fib, Q);
  R f[kSurrFeat];
  surr_features<R>(H, Jm1, Q, MAT, f);
  bool inside = true;
#pragma unroll
  for (int k = 0; k < 12; ++k) inside = inside && f[k] >= R(in.norm[32 + k]) && f[k] <= R(in.norm[48 + k]);
  const R K = R(pp.lambda) + Num<R>::div_c(R(2) * R(pp.mu), 3.0);
  R p = R(0.5) * K * Jm1 * (Jm1 + R(2));
  Sym<R> Siso;
  if (inside) {
    const float *y = in.y + (size_t)gp * in.gp_stride;
    Sym<R> Sm;
#pragma unroll
    for (int k = 0; k < 6; ++k) Sm.v[k] = (R(y[k]) * R(in.norm[72 + k]) + R(in.norm[64 + k])) * f[6];
    Siso = rotate_sym(Sm, Q, false);
  } else {
    iso_pk2<R, MAT>(pp, H, Jm1, fib, Siso, p);
    t.v[3] += 1.0;
  }
  t.v[4] += 1.0;
  R Fi[3][3];
  inverse_f(H, Jm1, Fi);
  Sym<R> Ci;
  {
    const int I[6] = {0, 1, 2, 1, 0, 0}, J[6] = {0, 1, 2, 2, 2, 1};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int i = I[k], j = J[k];
      Ci.v[k] = Fi[i][0] * Fi[j][0] + Fi[i][1] * Fi[j][1] + Fi[i][2] * Fi[j][2];
    }
  }
  if (in.audit) {
    Sym<R> Sref;
    R pr;
    iso_pk2<R, MAT>(pp, H, Jm1, fib, Sref, pr);
    double e2 = 0.0, r2 = 0.0, f2 = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double w = k < 3 ? 1.0 : 2.0;  // Frobenius norm of the symmetric tensor
      const double d = (double)Siso.v[k] - (double)Sref.v[k];
      const double full = (double)Sref.v[k] + (double)pr * (double)Ci.v[k];
      e2 += w * d * d;
      r2 += w * (double)Sref.v[k] * (double)Sref.v[k];
      f2 += w * full * full;
    }
    t.v[0] += e2;
    t.v[1] += r2;
    t.v[2] += f2;
    t.v[5] += 1.0;
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) S.v[k] = Siso.v[k] + p * Ci.v[k];
  if (MAT == 7 && use_hist) {
    const R c11 = R(vc.c11), c12 = R(vc.c12), c21 = R(vc.c21), c22 = R(vc.c22);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      R d = Siso.v[k] - hist[12 + k];
      R h1 = c11 * hist[k] + c21 * d;
      R h2 = c12 * hist[6 + k] + c22 * d;
      hist[k] = h1;
      hist[6 + k] = h2;
      hist[12 + k] = Siso.v[k];
      S.v[k] = S.v[k] + h1 + h2;
    }
  }
}

// Status bits of element_eval (FT_STATUS_* of the C ABI plus an internal
// "unknown material" bit).
constexpr int kFlagNonFinite = 1;
constexpr int kFlagInverted = 2;
constexpr int kFlagMaterial = 8;

// One element from the unnormalised Walsh-Hadamard coefficients of its
// reference coordinates (cx) and displacements (cu): c[m][d] = sum_p
// sign_m(p) v_d(p) over the 8 nodes in cube-position order p (see p2n), m =
// xi, eta, xi.eta, zeta, xi.zeta, eta.zeta, xi.eta.zeta -- 8 x the
// coefficients of the field's monomial expansion.  The factor 8 is never
// divided out: it cancels in H = Ju J0^-1, the force accumulators come out
// 64 x too large and the final constants take that out, and char_length
// takes it out itself (powers of two: every scaled result is exact).  cx and
// cu are overwritten.  Writes the 24 internal-force components in C3D8 node
// order (node-major, xyz) and lowers dt_e when DT is set and the element is
// not skipped.  `hist` is an accessor with load(gp, k) / store(gp, k, v) for
// the 18 Prony values of a Gauss point (materials 5 and 7).
// Per-Gauss-point fields of FemTech's element layer for one element, in the
// reference's layout (ShapeFunctions.cpp:174-205 with 8 Gauss points): F
// (9 per point, column-major), det F (1) and the PK2 stress (6, Voigt).
struct GpSink {
  double *F, *detF, *pk2;
};

template <typename R, bool FORCE, bool DT, int MAT, class Hist, bool SURR = false, bool GPOUT = false>
__host__ __device__ inline void element_eval(R cx[7][3], R cu[7][3], const PartDev &pp,
                                             bool skip_dt, double dt_step, Hist &hist,
                                             bool update_hist, R outv[24], double &my_dt,
                                             int &flags, const R fib[3] = nullptr,
                                             const SurrIn *sin = nullptr, SurrTally *tally = nullptr,
                                             const GpSink *gps = nullptr, double dt_exact_below = 1e300,
                                             bool check_finite = true) {
  if (DT && !skip_dt) {
    // StableTimeStep.cpp:11-30: dt_e = Le / ce on the deformed geometry.
    R dcf[7][3];
#pragma unroll
    for (int m = 0; m < 7; ++m)
#pragma unroll
      for (int d = 0; d < 3; ++d) dcf[m][d] = cx[m][d] + cu[m][d];
    // FP32 run loop: an element whose lower bound is above dt_exact_below
    // (a margin over the last minimum) cannot set the step; it reports the
    // bound instead of its exact length (the bound is below it).
    double dlb = 0.0;
    if (sizeof(R) == 4 && dt_exact_below < 1e300) dlb = (double)(char_length_lower(dcf) * R(pp.inv_ce));
    if (dlb > dt_exact_below) {
      if (dlb < my_dt) my_dt = dlb;
    } else {
      const R le = char_length(dcf);
      // FP64 keeps the reference's division; FP32 multiplies by 1 / ce
      const double dte = sizeof(R) == 8 ? (double)le / pp.ce : (double)(le * R(pp.inv_ce));
      if (dte < my_dt) my_dt = dte;  // NaN never lowers the minimum
    }
  }
  const int mat = MAT >= 0 ? MAT : pp.mat;
  if (FORCE && mat == 0 && !GPOUT) {  // rigid: no stress, no internal force
#pragma unroll
    for (int k = 0; k < 24; ++k) outv[k] = R(0);
  } else if (FORCE) {
    ViscoCoef vc = {0, 0, 0, 0};
    if (mat == 5 || mat == 7) {  // HGOIsotropicViscoelastic.cpp:126-131, global dt
      const double rt1 = dt_step / pp.t1, rt2 = dt_step / pp.t2;
      vc.c11 = exp(-rt1);
      vc.c12 = exp(-rt2);
      vc.c21 = pp.g1 * (1 - vc.c11) / rt1;
      vc.c22 = pp.g2 * (1 - vc.c12) / rt2;
    }
    const R g = R(kGp), g2 = R(kGp * kGp);
    R h[3][7];  // force accumulators in the monomial basis
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int m = 0; m < 7; ++m) h[d][m] = R(0);
    // Half sums: on the Gauss-point plane zeta = sz the columns of
    // J = d(field)/d(xi) are
    //   J[.][0] = a0 + sy b0,  a0 = c0 + sz g c4,  b0 = g c2 + sz g^2 c6,
    //   J[.][1] = a1 + sx b0,  a1 = c1 + sz g c5,
    //   J[.][2] = c3 + sx g c4 + sy g c5 + sx sy g^2 c6.
    // a0, a1, b0 replace c0, c1, c2 in place (plane zeta = +) and move to the
    // plane zeta = - once, so the loop holds no more values than before.
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      R a0 = Num<R>::fma_(g, cx[4][d], cx[0][d]), a1 = Num<R>::fma_(g, cx[5][d], cx[1][d]);
      R b0 = Num<R>::fma_(g2, cx[6][d], g * cx[2][d]);
      cx[0][d] = a0, cx[1][d] = a1, cx[2][d] = b0;
      a0 = Num<R>::fma_(g, cu[4][d], cu[0][d]), a1 = Num<R>::fma_(g, cu[5][d], cu[1][d]);
      b0 = Num<R>::fma_(g2, cu[6][d], g * cu[2][d]);
      cu[0][d] = a0, cu[1][d] = a1, cu[2][d] = b0;
    }
    // Gauss points in the reference's order (GaussQuadrature3D.cpp:19-49):
    // (-,-,+), (+,-,+), (+,+,+), (-,+,+), then the same at zeta = -.  The loop
    // is kept rolled so one Gauss point's temporaries are live at a time.
#pragma unroll 1
    for (int gp = 0; gp < 8; ++gp) {
      if (gp == 4) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          cx[0][d] = Num<R>::fma_(R(-2) * g, cx[4][d], cx[0][d]);
          cx[1][d] = Num<R>::fma_(R(-2) * g, cx[5][d], cx[1][d]);
          cx[2][d] = Num<R>::fma_(R(-2) * g2, cx[6][d], cx[2][d]);
          cu[0][d] = Num<R>::fma_(R(-2) * g, cu[4][d], cu[0][d]);
          cu[1][d] = Num<R>::fma_(R(-2) * g, cu[5][d], cu[1][d]);
          cu[2][d] = Num<R>::fma_(R(-2) * g2, cu[6][d], cu[2][d]);
        }
      }
      // sign bits of xi, eta, zeta at this Gauss point
      const unsigned nx = ((gp ^ (gp >> 1)) & 1) ^ 1, ny = ((gp >> 1) & 1) ^ 1, nz = (gp >> 2) & 1;
      const R sx = flip_sign(R(1), nx), sy = flip_sign(R(1), ny);
      const R gx = flip_sign(g, nx), gy = flip_sign(g, ny), gxy = flip_sign(g2, nx ^ ny);
      // monomial weights: 1/8 (linear), g/8 (bilinear), g^2/8 (trilinear),
      // and 1/64 for the factor 8 of the coefficients (adj(8 J) = 64 adj(J))
      const R w1 = R(0.125 / 64), wg = g * R(0.125 / 64), wg2 = g2 * R(0.125 / 64);
      const GpWeights<R> gw = {w1, flip_sign(wg, nx), flip_sign(wg, ny), flip_sign(wg, nz),
                               flip_sign(wg2, ny ^ nz), flip_sign(wg2, nx ^ nz), flip_sign(wg2, nx ^ ny)};
      // J[d][k] = d(field_d)/d(xi_k) at this Gauss point (8 x).
      R J0[3][3], Ju[3][3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        J0[d][0] = Num<R>::fma_(sy, cx[2][d], cx[0][d]);
        J0[d][1] = Num<R>::fma_(sx, cx[2][d], cx[1][d]);
        J0[d][2] = Num<R>::fma_(gxy, cx[6][d], Num<R>::fma_(gy, cx[5][d], Num<R>::fma_(gx, cx[4][d], cx[3][d])));
        Ju[d][0] = Num<R>::fma_(sy, cu[2][d], cu[0][d]);
        Ju[d][1] = Num<R>::fma_(sx, cu[2][d], cu[1][d]);
        Ju[d][2] = Num<R>::fma_(gxy, cu[6][d], Num<R>::fma_(gy, cu[5][d], Num<R>::fma_(gx, cu[4][d], cu[3][d])));
      }
      // adj(J0): J0^-1 = A / det0 (ShapeFunction_C3D8.cpp:95-123)
      R A[3][3];
      A[0][0] = J0[1][1] * J0[2][2] - J0[1][2] * J0[2][1];
      A[0][1] = J0[0][2] * J0[2][1] - J0[0][1] * J0[2][2];
      A[0][2] = J0[0][1] * J0[1][2] - J0[0][2] * J0[1][1];
      A[1][0] = J0[1][2] * J0[2][0] - J0[1][0] * J0[2][2];
      A[1][1] = J0[0][0] * J0[2][2] - J0[0][2] * J0[2][0];
      A[1][2] = J0[0][2] * J0[1][0] - J0[0][0] * J0[1][2];
      A[2][0] = J0[1][0] * J0[2][1] - J0[1][1] * J0[2][0];
      A[2][1] = J0[0][1] * J0[2][0] - J0[0][0] * J0[2][1];
      A[2][2] = J0[0][0] * J0[1][1] - J0[0][1] * J0[1][0];
      const R det0 = J0[0][0] * A[0][0] + J0[0][1] * A[1][0] + J0[0][2] * A[2][0];
      if (!(det0 > R(0))) flags |= kFlagInverted;
      const R idet0 = Num<R>::rcp_(det0);
      // H = Ju J0^-1 = grad_X u  (F = I + H, CalculateDeformationGradient.cpp:15-25)
      R H[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          H[i][j] = (Ju[i][0] * A[0][j] + Ju[i][1] * A[1][j] + Ju[i][2] * A[2][j]) * idet0;
      if constexpr ((MAT == 1 || MAT == 4 || MAT == 6 || MAT == 8) && !SURR && !GPOUT) {
        // Spatial form: f_a = sum_gp det0 P J0^-T dN_a/dxi with P = tau F^-T
        // and F J0 = Jx (the deformed Jacobian), so Q = det0 tau Jx^-T =
        // tau adj(Jx)^T / J.
        const R Jm1 = det_ip1_m1(H);
        if (!(Jm1 > R(-1))) flags |= kFlagInverted;
        Sym<R> tau;  // tau / J
        const R iJ = Num<R>::rcp_(R(1) + Jm1);
        if constexpr (MAT == 1) {
          // kirchhoff_stress's neo-Hookean tau = mu (b - I) + lambda ln J I,
          // with 1 / J applied to the two coefficients
          const Sym<R> bm = b_minus_identity(H);
          const R m = R(pp.mu) * iJ, l = R(pp.lambda) * Num<R>::log1p_(Jm1) * iJ;
#pragma unroll
          for (int k = 0; k < 6; ++k) tau.v[k] = m * bm.v[k];
          tau.v[0] += l;
          tau.v[1] += l;
          tau.v[2] += l;
        } else {
          kirchhoff_stress<R, MAT>(pp, H, Jm1, fib, tau);
#pragma unroll
          for (int k = 0; k < 6; ++k) tau.v[k] *= iJ;
        }
        R X[3][3], B[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) X[i][j] = J0[i][j] + Ju[i][j];
        B[0][0] = X[1][1] * X[2][2] - X[1][2] * X[2][1];
        B[0][1] = X[0][2] * X[2][1] - X[0][1] * X[2][2];
        B[0][2] = X[0][1] * X[1][2] - X[0][2] * X[1][1];
        B[1][0] = X[1][2] * X[2][0] - X[1][0] * X[2][2];
        B[1][1] = X[0][0] * X[2][2] - X[0][2] * X[2][0];
        B[1][2] = X[0][2] * X[1][0] - X[0][0] * X[1][2];
        B[2][0] = X[1][0] * X[2][1] - X[1][1] * X[2][0];
        B[2][1] = X[0][1] * X[2][0] - X[0][0] * X[2][1];
        B[2][2] = X[0][0] * X[1][1] - X[0][1] * X[1][0];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const R t0 = sym_at(tau, i, 0), t1 = sym_at(tau, i, 1), t2 = sym_at(tau, i, 2);
          const R q0 = t0 * B[0][0] + t1 * B[0][1] + t2 * B[0][2];
          const R q1 = t0 * B[1][0] + t1 * B[1][1] + t2 * B[1][2];
          const R q2 = t0 * B[2][0] + t1 * B[2][1] + t2 * B[2][2];
          accumulate_force_basis(h[i], q0, q1, q2, gw);
        }
        continue;
      }
      // history of this Gauss point, in registers (fixed-size array, static
      // indices: no private-memory copy)
      R hloc[18];
      const bool use_hist = (mat == 5 || mat == 7) && hist.valid();
#pragma unroll
      for (int k = 0; k < 18; ++k) hloc[k] = use_hist ? hist.load(gp, k) : R(0);
      Sym<R> S;
      R Jm1;
      if constexpr (SURR) {
        surr_pk2<R, MAT>(pp, H, fib, *sin, gp, hloc, use_hist, vc, S, Jm1, *tally);
      } else {
        if (!pk2_stress<R, MAT>(pp, H, hloc, use_hist, vc, S, Jm1)) flags |= kFlagMaterial;
      }
      if (!(Jm1 > R(-1))) flags |= kFlagInverted;
      if (use_hist && update_hist) {
#pragma unroll
        for (int k = 0; k < 18; ++k) hist.store(gp, k, hloc[k]);
      }
      if constexpr (GPOUT) {
        // F = I + H (CalculateDeformationGradient.cpp:14 layout), det F, S
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) gps->F[9 * gp + 3 * j + i] = (double)H[i][j] + (i == j ? 1.0 : 0.0);
        gps->detF[gp] = 1.0 + (double)Jm1;
        for (int k = 0; k < 6; ++k) gps->pk2[6 * gp + k] = (double)S.v[k];
      }
      // P = F S (first Piola), Q = w detJ0 P J0^-T = P A^T  (w = 1)
      R P[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          P[i][j] = sym_at(S, i, j) + H[i][0] * sym_at(S, 0, j) + H[i][1] * sym_at(S, 1, j) +
                    H[i][2] * sym_at(S, 2, j);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const R q0 = P[i][0] * A[0][0] + P[i][1] * A[0][1] + P[i][2] * A[0][2];
        const R q1 = P[i][0] * A[1][0] + P[i][1] * A[1][1] + P[i][2] * A[1][2];
        const R q2 = P[i][0] * A[2][0] + P[i][1] * A[2][1] + P[i][2] * A[2][2];
        // f_a,i = sum_gp sum_k Q_ik dN_a/dxi_k in the monomial basis
        accumulate_force_basis(h[i], q0, q1, q2, gw);
      }
    }
    // Back to the nodes: f_p = sum_q c_q prod(signs) (h carries the scale).
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      R c[8];
      c[0] = R(0);
      c[1] = h[d][0];
      c[2] = h[d][1];
      c[4] = h[d][3];
      c[3] = h[d][2];
      c[5] = h[d][4];
      c[6] = h[d][5];
      c[7] = h[d][6];
      iwht8(c);
#pragma unroll
      for (int p = 0; p < 8; ++p) outv[3 * p2n(p) + d] = c[p];
    }
    // (the run loop's node update checks the node sums instead: every
    // element force reaches one, and a non-finite value keeps the sum
    // non-finite)
    if (check_finite) {
      bool finite = true;
#pragma unroll
      for (int k = 0; k < 24; ++k) finite = finite && isfinite(outv[k]);
      if (!finite) flags |= kFlagNonFinite;
    }
  }
}

}  // namespace ft
