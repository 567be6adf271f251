/*
 * oracle.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * FP64 CPU restatement of FemTech's explicit C3D8 step.  It follows the
 * reference's per-Gauss-point algorithm and operation order so that it agrees
 * with the reference build to round-off; it is not the product and nothing in
 * the product links it.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   gauss/shape functions  golden vectors from the reference build + patch
 *                          tests (sum dN/dX = 0, sum X (x) dN/dX = I,
 *                          sum detJ = V)
 *   stress (materials 1-5) golden vectors from the reference build + closed
 *                          forms (S(I) = 0, objectivity, uniaxial NH/SVK)
 *   element force          golden vectors + sum_a f_a = 0 + rigid rotation
 *   char. length / dt      golden vectors (jittered faces exercise the
 *                          parallelogram shortcut of Geometry.cpp:46)
 *   lumped mass            golden vectors + sum m = rho V, rho V/8 per corner
 *   run loop + energy      golden trajectories from the reference build +
 *                          the reference's Abaqus report (ex9, 10 elements)
 *   motions 2 and 3        no reference counterpart: pinned by the closed-form
 *                          kinematics (finite differences of u(t) give v, a;
 *                          |r| preserved).  Motion 3 is "parity unpinned"
 *                          against the reference, which has no such loading.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kGauss = 0.577350269189626;  // GaussQuadrature3D.cpp:19-49

/* inverse3x3Matrix, math.cpp:8-28 (column-major in and out). */
void inverse3(const double *mat, double *inv, double *det) {
  double d = mat[0] * (mat[4] * mat[8] - mat[5] * mat[7]) -
             mat[3] * (mat[1] * mat[8] - mat[7] * mat[2]) +
             mat[6] * (mat[1] * mat[5] - mat[4] * mat[2]);
  double id = 1 / d;
  inv[0] = (mat[4] * mat[8] - mat[5] * mat[7]) * id;
  inv[3] = (mat[6] * mat[5] - mat[3] * mat[8]) * id;
  inv[6] = (mat[3] * mat[7] - mat[6] * mat[4]) * id;
  inv[1] = (mat[7] * mat[2] - mat[1] * mat[8]) * id;
  inv[4] = (mat[0] * mat[8] - mat[6] * mat[2]) * id;
  inv[7] = (mat[1] * mat[6] - mat[0] * mat[7]) * id;
  inv[2] = (mat[1] * mat[5] - mat[2] * mat[4]) * id;
  inv[5] = (mat[2] * mat[3] - mat[0] * mat[5]) * id;
  inv[8] = (mat[0] * mat[4] - mat[1] * mat[3]) * id;
  *det = d;
}

/* C = op(A) * op(B) for 3x3 column-major, summation over k in order, as the
 * BLAS dgemm calls of the material routines compute them.  ta/tb: transpose. */
void gemm3(bool ta, bool tb, double alpha, const double *A, const double *B,
           double *C) {
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) {
        double a = ta ? A[k + 3 * i] : A[i + 3 * k];
        double b = tb ? B[j + 3 * k] : B[k + 3 * j];
        s += a * b;
      }
      C[i + 3 * j] = alpha * s;
    }
}

/* InverseF, InverseF.cpp:42-66: returns F^{-1} column-major using the stored
 * det F. */
void inverseF(const double *F, double detA, double *fInv) {
  double A[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) A[r][c] = F[3 * r + c];
  fInv[0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / detA;
  fInv[1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / detA;
  fInv[2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / detA;
  fInv[3] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / detA;
  fInv[4] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / detA;
  fInv[5] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / detA;
  fInv[6] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / detA;
  fInv[7] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / detA;
  fInv[8] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / detA;
}

void voigt_from(const double *S, double *S6) {
  // Voigt (11,22,33,23,13,12) from column-major [0,4,8,7,6,3]
  // (CompressibleNeoHookean.cpp:45-55 and siblings)
  S6[0] = S[0];
  S6[1] = S[4];
  S6[2] = S[8];
  S6[3] = S[7];
  S6[4] = S[6];
  S6[5] = S[3];
}

/* CompressibleNeoHookean.cpp:14-58 */
void neo_hookean(const double *p, const double *F, double J, double *S6) {
  double mu = p[1], lambda = p[2];
  double C[9], Ci[9], Cdet;
  gemm3(true, false, 1.0, F, F, C);
  inverse3(C, Ci, &Cdet);
  double logJ = lambda * log(J);
  S6[0] = mu * (1.0 - Ci[0]) + logJ * Ci[0];
  S6[1] = mu * (1.0 - Ci[4]) + logJ * Ci[4];
  S6[2] = mu * (1.0 - Ci[8]) + logJ * Ci[8];
  S6[3] = -mu * Ci[7] + logJ * Ci[7];
  S6[4] = -mu * Ci[6] + logJ * Ci[6];
  S6[5] = -mu * Ci[3] + logJ * Ci[3];
}

/* StVenantKirchhoff.cpp:14-61 */
void st_venant(const double *p, const double *F, double *S6) {
  double mu = p[1], lambda = p[2];
  double E[9];
  gemm3(true, false, 0.5, F, F, E);
  E[0] -= 0.5;
  E[4] -= 0.5;
  E[8] -= 0.5;
  double trE = E[0] + E[4] + E[8];
  double S[9];
  for (int i = 0; i < 9; ++i) S[i] = 2.0 * mu * E[i];
  S[0] += lambda * trE;
  S[4] += lambda * trE;
  S[8] += lambda * trE;
  voigt_from(S, S6);
}

/* LinearElastic.cpp:14-70: S = F^{-1} P, not symmetrised; the upper triangle
 * is what is stored. */
void linear_elastic(const double *p, const double *F, double J, double *S6) {
  double mu = p[1], lambda = p[2];
  double trEps = F[0] + F[4] + F[8] - 3.0;
  double eps[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) eps[j + 3 * i] = F[j + 3 * i] + F[i + 3 * j];
  eps[0] -= 2.0;
  eps[4] -= 2.0;
  eps[8] -= 2.0;
  double P[9];
  for (int i = 0; i < 9; ++i) P[i] = mu * eps[i];
  P[0] += lambda * trEps;
  P[4] += lambda * trEps;
  P[8] += lambda * trEps;
  double fInv[9], S[9];
  inverseF(F, J, fInv);
  gemm3(false, false, 1.0, fInv, P, S);
  voigt_from(S, S6);
}

/* HGOIsotropic.cpp:21-115 (kappa = 1/3) and, with visco != 0,
 * HGOIsotropicViscoelastic.cpp:27-168 (two-term Prony series, global dt). */
void hgo(const double *p, const double *F, double J, bool visco, double dt,
         double *hist, double *S6) {
  const double mu = p[1], lambda = p[2], k1 = p[3], k2 = p[4];
  const double kappa = 1.0 / 3.0;
  const double K = lambda + 2.0 * mu / 3.0;
  const double hydroDiag = 0.5 * K * (J * J - 1.0) / J;
  double Bm[9];
  gemm3(false, true, 1.0, F, F, Bm);
  const double traceB = Bm[0] + Bm[4] + Bm[8];
  const double Jm23 = pow(J, -2.0 / 3.0);
  const double I1 = Jm23 * traceB;
  const double E_alpha = kappa * (I1 - 3.0);
  double fiber = 0.0;
  if (E_alpha > 0.0) fiber = 2.0 * k1 * exp(k2 * E_alpha * E_alpha) * E_alpha * kappa;
  const double traceBby3 = traceB / 3.0;
  Bm[0] = Bm[0] - traceBby3;
  Bm[4] = Bm[4] - traceBby3;
  Bm[8] = Bm[8] - traceBby3;
  if (!visco) {
    const double total = Jm23 * (mu + fiber) / J;
    for (int i = 0; i < 9; ++i) Bm[i] = Bm[i] * total;
  } else {
    const double total = (mu + fiber) / J;
    for (int i = 0; i < 9; ++i) Bm[i] = Bm[i] * Jm23 * total;
  }
  Bm[0] += hydroDiag;
  Bm[4] += hydroDiag;
  Bm[8] += hydroDiag;
  double fInv[9], ST[9], S[9];
  inverseF(F, J, fInv);
  gemm3(false, false, 1.0, fInv, Bm, ST);
  gemm3(false, true, J, ST, fInv, S);
  if (visco) {
    const double g1 = p[5], t1 = p[6], g2 = p[7], t2 = p[8];
    double C[9];
    gemm3(true, false, 1.0, F, F, C);
    double SddC = 0.0;
    for (int i = 0; i < 9; ++i) SddC += C[i] * S[i];
    SddC = SddC / 3.0;
    double Sic[9], Sdev[9];
    gemm3(false, true, SddC, fInv, fInv, Sic);
    for (int i = 0; i < 9; ++i) Sdev[i] = S[i] - Sic[i];
    double *H1 = hist, *H2 = hist + 9, *S0 = hist + 18;
    const double rt1 = dt / t1, rt2 = dt / t2;
    const double c11 = exp(-rt1), c12 = exp(-rt2);
    const double c21 = g1 * (1 - c11) / rt1, c22 = g2 * (1 - c12) / rt2;
    for (int i = 0; i < 9; ++i) {
      H1[i] = c11 * H1[i] + c21 * (Sdev[i] - S0[i]);
      H2[i] = c12 * H2[i] + c22 * (Sdev[i] - S0[i]);
      S[i] = S[i] + H1[i] + H2[i];
    }
    for (int i = 0; i < 9; ++i) S0[i] = Sdev[i];
  }
  voigt_from(S, S6);
}

/* Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (converged
 * to round-off): eigenvalues w, eigenvectors in the columns of V
 * (column-major V[3*col + row]). */
void jacobi3(const double *A, double *w, double *V) {
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = 0.5 * (A[i + 3 * j] + A[j + 3 * i]);
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 50; ++sweep) {
    const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
    const double scale = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
    if (off == 0.0 || off < 1e-300 || off <= 1e-18 * scale) break;
    for (int r = 0; r < 3; ++r) {
      const int p = r == 2 ? 1 : 0, q = r == 0 ? 1 : 2;
      if (a[p][q] == 0.0) continue;
      const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
      double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
      if (theta < 0) t = -t;
      const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
      double b[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) b[i][j] = a[i][j];
      for (int k = 0; k < 3; ++k) {  // a <- R^T a R
        b[k][p] = c * a[k][p] - sn * a[k][q];
        b[k][q] = sn * a[k][p] + c * a[k][q];
      }
      for (int k = 0; k < 3; ++k) {
        a[p][k] = c * b[p][k] - sn * b[q][k];
        a[q][k] = sn * b[p][k] + c * b[q][k];
        if (k != p && k != q) a[k][p] = a[p][k], a[k][q] = a[q][k];
      }
      a[p][q] = a[q][p] = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double vp = v[k][p], vq = v[k][q];
        v[k][p] = c * vp - sn * vq;
        v[k][q] = sn * vp + c * vq;
      }
    }
  }
  for (int i = 0; i < 3; ++i) {
    w[i] = a[i][i];
    for (int k = 0; k < 3; ++k) V[3 * i + k] = v[k][i];
  }
}

/* Kaliske-Rothert two-term Prony update of HGOIsotropicViscoelastic.cpp:
 * 94-159 on a total PK2 S (column-major), applied in place. */
void prony_update(const double *p, const double *F, double dt, double *hist, double *S) {
  const double g1 = p[5], t1 = p[6], g2 = p[7], t2 = p[8];
  double C[9], fInv[9], J = 0.0;
  gemm3(true, false, 1.0, F, F, C);
  {
    double det;
    inverse3(F, fInv, &det);
    J = det;
  }
  (void)J;
  double SddC = 0.0;
  for (int i = 0; i < 9; ++i) SddC += C[i] * S[i];
  SddC = SddC / 3.0;
  double Sic[9], Sdev[9];
  gemm3(false, true, SddC, fInv, fInv, Sic);
  for (int i = 0; i < 9; ++i) Sdev[i] = S[i] - Sic[i];
  double *H1 = hist, *H2 = hist + 9, *S0 = hist + 18;
  const double rt1 = dt / t1, rt2 = dt / t2;
  const double c11 = exp(-rt1), c12 = exp(-rt2);
  const double c21 = g1 * (1 - c11) / rt1, c22 = g2 * (1 - c12) / rt2;
  for (int i = 0; i < 9; ++i) {
    H1[i] = c11 * H1[i] + c21 * (Sdev[i] - S0[i]);
    H2[i] = c12 * H2[i] + c22 * (Sdev[i] - S0[i]);
    S[i] = S[i] + H1[i] + H2[i];
  }
  for (int i = 0; i < 9; ++i) S0[i] = Sdev[i];
}

/* Pull-back S = F^-1 tau F^-T of a Kirchhoff stress (column-major). */
void pull_back(const double *F, const double *tau, double *S) {
  double fInv[9], T[9], det;
  inverse3(F, fInv, &det);
  gemm3(false, false, 1.0, fInv, tau, T);
  gemm3(false, true, 1.0, T, fInv, S);
}

/* 1-term Ogden (no reference counterpart): W = (2 mu / alpha^2) sum
 * (lbar_a^alpha - 1) + K/4 (J^2 - 1 - 2 ln J), K = lambda + 2 mu / 3 as in
 * HGOIsotropic.cpp:44; tau = sum_a (2 mu / alpha)(lbar_a^alpha - mean) n_a
 * n_a + K/2 (J^2 - 1) I; Prony (material 7) as HGOIsotropicViscoelastic. */
void ogden(const double *p, const double *F, double J, bool visco, double dt, double *hist, double *S6) {
  const double mu = p[1], lambda = p[2], alpha = p[3];
  const double K = lambda + 2.0 * mu / 3.0;
  double B[9], w[3], V[9];
  gemm3(false, true, 1.0, F, F, B);
  jacobi3(B, w, V);
  double lb[3], mean = 0.0;
  for (int a = 0; a < 3; ++a) {
    lb[a] = pow(pow(J, -1.0 / 3.0) * sqrt(w[a]), alpha);
    mean += lb[a] / 3.0;
  }
  double tau[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int a = 0; a < 3; ++a) {
    const double ta = 2.0 * mu / alpha * (lb[a] - mean);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) tau[i + 3 * j] += ta * V[3 * a + i] * V[3 * a + j];
  }
  const double pv = 0.5 * K * (J * J - 1.0);
  tau[0] += pv;
  tau[4] += pv;
  tau[8] += pv;
  double S[9];
  pull_back(F, tau, S);
  if (visco) prony_update(p, F, dt, hist, S);
  voigt_from(S, S6);
}

/* Anisotropic HGO with fibre dispersion (no reference counterpart):
 * W = mu/2 (I1bar - 3) + K/4 (J^2 - 1 - 2 ln J) + k1/(2 k2)(exp(k2 E^2) - 1),
 * E = kappa (I1bar - 3) + (1 - 3 kappa)(I4bar - 1) counted for E > 0 only,
 * I4bar = J^-2/3 |F a|^2; kappa = 1/3 is HGOIsotropic.cpp:21-115. */
void hgo_aniso(const double *p, const double *fib, const double *F, double J, double *S6) {
  const double mu = p[1], lambda = p[2], k1 = p[3], k2 = p[4], kappa = p[5];
  const double K = lambda + 2.0 * mu / 3.0;
  double B[9];
  gemm3(false, true, 1.0, F, F, B);
  double fa[3];
  for (int i = 0; i < 3; ++i) fa[i] = F[i] * fib[0] + F[i + 3] * fib[1] + F[i + 6] * fib[2];
  const double Jm23 = pow(J, -2.0 / 3.0);
  const double I1b = Jm23 * (B[0] + B[4] + B[8]);
  const double I4b = Jm23 * (fa[0] * fa[0] + fa[1] * fa[1] + fa[2] * fa[2]);
  const double E = kappa * (I1b - 3.0) + (1.0 - 3.0 * kappa) * (I4b - 1.0);
  const double psi = E > 0.0 ? k1 * E * exp(k2 * E * E) : 0.0;
  double tau[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      tau[i + 3 * j] = Jm23 * (mu + 2.0 * psi * kappa) * B[i + 3 * j] +
                       Jm23 * 2.0 * psi * (1.0 - 3.0 * kappa) * fa[i] * fa[j];
  const double tr = (tau[0] + tau[4] + tau[8]) / 3.0;
  const double pv = 0.5 * K * (J * J - 1.0);
  tau[0] += pv - tr;
  tau[4] += pv - tr;
  tau[8] += pv - tr;
  double S[9];
  pull_back(F, tau, S);
  voigt_from(S, S6);
}

/* tripleProduct, math.cpp:52-56 */
double triple(const double *s, const double *a, const double *b) {
  return s[2] * (a[0] * b[1] - a[1] * b[0]) + s[0] * (a[1] * b[2] - a[2] * b[1]) -
         s[1] * (a[0] * b[2] - a[2] * b[0]);
}

/* normOfCrossProduct, math.cpp:30-43 (ndim == 3) */
double norm_cross(const double *a, const double *b) {
  double z = a[0] * b[1] - a[1] * b[0];
  double x = a[1] * b[2] - a[2] * b[1];
  double y = -a[0] * b[2] + a[2] * b[0];
  return sqrt(x * x + y * y + z * z);
}

void motion_eval(const double *mo, const double *X, double t, double *u,
                 double *v, double *a) {
  int type = (int)mo[0];
  for (int d = 0; d < 3; ++d) u[d] = v[d] = a[d] = 0.0;
  if (type == 1) {
    for (int d = 0; d < 3; ++d) {
      u[d] = t * mo[1];
      v[d] = mo[1];
    }
  } else if (type == 2) {
    double amp = mo[1], T = mo[2];
    double s = t < T ? t / T : 1.0;
    for (int d = 0; d < 3; ++d) {
      u[d] = amp * s;
      v[d] = t < T ? amp / T : 0.0;
    }
  } else if (type == 3) {
    const double n[3] = {mo[1], mo[2], mo[3]};
    const double c[3] = {mo[4], mo[5], mo[6]};
    const double A = mo[7], T = mo[8];
    const double pi = 3.14159265358979323846;
    double alpha, omega, theta;
    if (t <= T) {
      alpha = A * sin(pi * t / T);
      omega = A * T / pi * (1.0 - cos(pi * t / T));
      theta = A * T / pi * (t - T / pi * sin(pi * t / T));
    } else {
      alpha = 0.0;
      omega = 2.0 * A * T / pi;
      theta = A * T * T / pi + omega * (t - T);
    }
    double r0[3] = {X[0] - c[0], X[1] - c[1], X[2] - c[2]};
    // Rodrigues: r = r0 cos + (n x r0) sin + n (n.r0)(1-cos)
    double ct = cos(theta), st = sin(theta);
    double nxr0[3] = {n[1] * r0[2] - n[2] * r0[1], n[2] * r0[0] - n[0] * r0[2],
                      n[0] * r0[1] - n[1] * r0[0]};
    double ndr0 = n[0] * r0[0] + n[1] * r0[1] + n[2] * r0[2];
    double r[3];
    for (int d = 0; d < 3; ++d)
      r[d] = r0[d] * ct + nxr0[d] * st + n[d] * ndr0 * (1.0 - ct);
    double nxr[3] = {n[1] * r[2] - n[2] * r[1], n[2] * r[0] - n[0] * r[2],
                     n[0] * r[1] - n[1] * r[0]};
    double nxnxr[3] = {n[1] * nxr[2] - n[2] * nxr[1], n[2] * nxr[0] - n[0] * nxr[2],
                       n[0] * nxr[1] - n[1] * nxr[0]};
    for (int d = 0; d < 3; ++d) {
      u[d] = r[d] - r0[d];
      v[d] = omega * nxr[d];
      a[d] = alpha * nxr[d] + omega * omega * nxnxr[d];
    }
  }
}

struct Model {
  int nn = 0, ne = 0, nparts = 0, threads = 1;
  std::vector<double> X;
  std::vector<int> conn, pid, matid;
  std::vector<double> props;
  std::vector<double> dNdX, detJ;  // ShapeFunctions output
  std::vector<double> hist;        // 8*27 per element (material 5 only)
  bool has_visco = false;
  std::vector<double> fibres;  // 3 per element (material 8), unit length
  // boundary conditions
  std::vector<int> dof_motion;
  std::vector<double> motions;
  std::vector<double> u0, v0;
  // node -> (element*8 + slot), sorted by element: deterministic gather
  std::vector<int> csr_ptr, csr_idx;
  // injury criteria (ex5.cpp:1253-1430): state of the included elements
  bool injury = false;
  std::vector<int> inj_elem;                    // elementIDInjury
  std::vector<double> inj_ps_old, inj_psxsr;    // PS_Old, PSxSRArray
  std::vector<int> inj_gt15, inj_gt30, inj_psr120, inj_psxsr28;
  std::vector<double> inj_volume;               // reference-configuration volumes
  std::vector<char> inj_in_mps95, inj_in_mpsxsr95;  // latest element lists
  double maxStrain = 0, maxT = 0, minStrain = 0, minT = 0, maxShear = 0, maxShearT = 0;
  double maxPSxSR = 0, maxTimePSxSR = 0, maxMPS95 = 0, maxTimeMPS95 = 0;
  double maxMPSxSR95 = 0, maxTimeMPSxSR95 = 0;
  int maxElem = -1, minElem = -1, shearElem = -1, maxElemPSxSR = -1;
};

void build_csr(Model &m) {
  m.csr_ptr.assign(m.nn + 1, 0);
  for (int e = 0; e < m.ne; ++e)
    for (int a = 0; a < 8; ++a) m.csr_ptr[m.conn[8 * e + a] + 1]++;
  for (int n = 0; n < m.nn; ++n) m.csr_ptr[n + 1] += m.csr_ptr[n];
  m.csr_idx.assign(8 * (size_t)m.ne, 0);
  std::vector<int> fill(m.csr_ptr.begin(), m.csr_ptr.end() - 1);
  for (int e = 0; e < m.ne; ++e)
    for (int a = 0; a < 8; ++a) m.csr_idx[fill[m.conn[8 * e + a]]++] = 8 * e + a;
}

void element_coords(const Model &m, int e, const double *u, double *xe) {
  for (int a = 0; a < 8; ++a) {
    int n = m.conn[8 * e + a];
    for (int d = 0; d < 3; ++d)
      xe[3 * a + d] = m.X[3 * n + d] + (u ? u[3 * n + d] : 0.0);
  }
}

/* GetForce_3D.cpp:15-46 for one element: CalculateDeformationGradient,
 * DeterminateF, StressUpdate, InternalForceUpdate over the 8 Gauss points. */
int element_force(const Model &m, int e, const double *u, double dt,
                  double *hist_e, double *f) {
  const double *dN = &m.dNdX[192 * (size_t)e];
  const double *dJ = &m.detJ[8 * (size_t)e];
  int part = m.pid[e];
  int mat = m.matid[part];
  const double *p = &m.props[9 * part];
  double x[24];
  element_coords(m, e, u, x);
  for (int k = 0; k < 24; ++k) f[k] = 0.0;
  for (int gp = 0; gp < 8; ++gp) {
    // CalculateDeformationGradient.cpp:15-25, F[3j+i] = sum_k x_ki dN_k/dX_j
    double F[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int k = 0; k < 8; ++k) s = s + x[3 * k + i] * dN[gp * 24 + 3 * k + j];
        F[3 * j + i] = s;
      }
    double J = oracle_det3(F);
    double S6[6] = {0, 0, 0, 0, 0, 0};
    int rc = oracle_stress_fibre(mat, p, m.fibres.empty() ? nullptr : &m.fibres[3 * (size_t)e], F, J, dt,
                                 hist_e ? hist_e + 27 * gp : nullptr, S6);
    if (rc) return rc;
    // InternalForceUpdate.cpp:9-26 with StrainDisplacementMatrix.cpp:27-76
    const double pre = 1.0 * dJ[gp];
    for (int k = 0; k < 8; ++k) {
      double dx = dN[gp * 24 + 3 * k], dy = dN[gp * 24 + 3 * k + 1],
             dz = dN[gp * 24 + 3 * k + 2];
      double F11 = F[0], F21 = F[1], F31 = F[2], F12 = F[3], F22 = F[4], F32 = F[5],
             F13 = F[6], F23 = F[7], F33 = F[8];
      double Bc[3][6] = {
          {dx * F11, dy * F12, dz * F13, dy * F13 + dz * F12, dx * F13 + dz * F11,
           dx * F12 + dy * F11},
          {dx * F21, dy * F22, dz * F23, dy * F23 + dz * F22, dx * F23 + dz * F21,
           dx * F22 + dy * F21},
          {dx * F31, dy * F32, dz * F33, dy * F33 + dz * F32, dx * F33 + dz * F31,
           dx * F32 + dy * F31}};
      for (int l = 0; l < 3; ++l) {
        double s = 0.0;
        for (int r = 0; r < 6; ++r) s += Bc[l][r] * S6[r];
        f[3 * k + l] += pre * s;
      }
    }
  }
  return 0;
}

}  // namespace

extern "C" {

void oracle_gauss_c3d8(double chi[24], double w[8]) {
  // GaussQuadrature3D.cpp:19-58: (-,-,+),(+,-,+),(+,+,+),(-,+,+) then zeta = -
  const double s[8][3] = {{-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1},
                          {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}};
  for (int g = 0; g < 8; ++g) {
    for (int d = 0; d < 3; ++d) chi[3 * g + d] = s[g][d] * kGauss;
    w[g] = 1.0;
  }
}

void oracle_shape_c3d8(const double xe[24], double dNdX[192], double detJ[8]) {
  double chi[24], w[8];
  oracle_gauss_c3d8(chi, w);
  for (int gp = 0; gp < 8; ++gp) {
    const double c = chi[3 * gp], e = chi[3 * gp + 1], z = chi[3 * gp + 2];
    double ds[24];  // ShapeFunction_C3D8.cpp:35-62, dN_a/dxi_k at ds[3a+k]
    ds[0] = -((e - 1) * (z - 1)) / 8;
    ds[3] = ((e - 1) * (z - 1)) / 8;
    ds[6] = -((e + 1) * (z - 1)) / 8;
    ds[9] = ((e + 1) * (z - 1)) / 8;
    ds[12] = ((e - 1) * (z + 1)) / 8;
    ds[15] = -((e - 1) * (z + 1)) / 8;
    ds[18] = ((e + 1) * (z + 1)) / 8;
    ds[21] = -((e + 1) * (z + 1)) / 8;
    ds[1] = -((c - 1) * (z - 1)) / 8;
    ds[4] = ((c + 1) * (z - 1)) / 8;
    ds[7] = -((c + 1) * (z - 1)) / 8;
    ds[10] = ((c - 1) * (z - 1)) / 8;
    ds[13] = ((c - 1) * (z + 1)) / 8;
    ds[16] = -((c + 1) * (z + 1)) / 8;
    ds[19] = ((c + 1) * (z + 1)) / 8;
    ds[22] = -((c - 1) * (z + 1)) / 8;
    ds[2] = -((c - 1) * (e - 1)) / 8;
    ds[5] = ((c + 1) * (e - 1)) / 8;
    ds[8] = -((c + 1) * (e + 1)) / 8;
    ds[11] = ((c - 1) * (e + 1)) / 8;
    ds[14] = ((c - 1) * (e - 1)) / 8;
    ds[17] = -((c + 1) * (e - 1)) / 8;
    ds[20] = ((c + 1) * (e + 1)) / 8;
    ds[23] = -((c - 1) * (e + 1)) / 8;
    // ShapeFunction_C3D8.cpp:78-93: xs[k + 3j] = dx_j/dxi_k from differences
    const double *x = xe;
    double xs[9];
    for (int j = 0; j < 3; ++j) {
      xs[0 + 3 * j] = (x[3 * 1 + j] - x[3 * 0 + j]) * ds[3 * 1 + 0] +
                      (x[3 * 2 + j] - x[3 * 3 + j]) * ds[3 * 2 + 0] +
                      (x[3 * 5 + j] - x[3 * 4 + j]) * ds[3 * 5 + 0] +
                      (x[3 * 6 + j] - x[3 * 7 + j]) * ds[3 * 6 + 0];
      xs[1 + 3 * j] = (x[3 * 2 + j] - x[3 * 1 + j]) * ds[3 * 2 + 1] +
                      (x[3 * 3 + j] - x[3 * 0 + j]) * ds[3 * 3 + 1] +
                      (x[3 * 6 + j] - x[3 * 5 + j]) * ds[3 * 6 + 1] +
                      (x[3 * 7 + j] - x[3 * 4 + j]) * ds[3 * 7 + 1];
      xs[2 + 3 * j] = (x[3 * 4 + j] - x[3 * 0 + j]) * ds[3 * 4 + 2] +
                      (x[3 * 5 + j] - x[3 * 1 + j]) * ds[3 * 5 + 2] +
                      (x[3 * 6 + j] - x[3 * 2 + j]) * ds[3 * 6 + 2] +
                      (x[3 * 7 + j] - x[3 * 3 + j]) * ds[3 * 7 + 2];
    }
    double Ji[9], det;
    inverse3(xs, Ji, &det);
    detJ[gp] = det;
    // ShapeFunction_C3D8.cpp:108-123
    for (int a = 0; a < 8; ++a) {
      const double *b = &ds[3 * a];
      double c1 = b[0] * Ji[0] + b[1] * Ji[3] + b[2] * Ji[6];
      double c2 = b[0] * Ji[1] + b[1] * Ji[4] + b[2] * Ji[7];
      double c3 = b[0] * Ji[2] + b[1] * Ji[5] + b[2] * Ji[8];
      dNdX[gp * 24 + 3 * a + 0] = c1;
      dNdX[gp * 24 + 3 * a + 1] = c2;
      dNdX[gp * 24 + 3 * a + 2] = c3;
    }
  }
}

/* DeterminateF.cpp:38-57 */
double oracle_det3(const double *F) {
  double da = F[0], db = F[3], dc = F[6], dd = F[1], de = F[4], df = F[7],
         dg = F[2], dh = F[5], di = F[8];
  return da * (de * di - df * dh) - db * (dd * di - df * dg) + dc * (dd * dh - de * dg);
}

/* StressUpdate.cpp:5-29 */
int oracle_stress(int material, const double props[9], const double F[9], double J,
                  double dt, double *hist, double S6[6]) {
  return oracle_stress_fibre(material, props, nullptr, F, J, dt, hist, S6);
}

int oracle_stress_fibre(int material, const double props[9], const double fibre[3], const double F[9],
                        double J, double dt, double *hist, double S6[6]) {
  static const double kDefaultFibre[3] = {1.0, 0.0, 0.0};
  switch (material) {
    case 6:
      ogden(props, F, J, false, dt, nullptr, S6);
      return 0;
    case 7:
      ogden(props, F, J, true, dt, hist, S6);
      return 0;
    case 8:
      hgo_aniso(props, fibre ? fibre : kDefaultFibre, F, J, S6);
      return 0;
    case 0:
      return 0;  // rigid: stress untouched (zero)
    case 1:
      neo_hookean(props, F, J, S6);
      return 0;
    case 2:
      st_venant(props, F, S6);
      return 0;
    case 3:
      linear_elastic(props, F, J, S6);
      return 0;
    case 4:
      hgo(props, F, J, false, dt, nullptr, S6);
      return 0;
    case 5:
      hgo(props, F, J, true, dt, hist, S6);
      return 0;
    default:
      return 1;  // StressUpdate.cpp:25-26, TerminateFemTech(1)
  }
}

/* volumeHexahedron, Geometry.cpp:3-27 */
double oracle_hex_volume(const double *L) {
  double q0[3], q1[3], q2[3], q3[3], q4[3], q5[3];
  for (int i = 0; i < 3; ++i) {
    q0[i] = L[i] - L[3 + i] + L[6 + i] - L[9 + i] + L[12 + i] - L[15 + i] +
            L[18 + i] - L[21 + i];
    q1[i] = L[i] - L[3 + i] - L[6 + i] + L[9 + i] - L[12 + i] + L[15 + i] +
            L[18 + i] - L[21 + i];
    q2[i] = -L[i] + L[3 + i] + L[6 + i] - L[9 + i] - L[12 + i] + L[15 + i] +
            L[18 + i] - L[21 + i];
    q3[i] = L[i] + L[3 + i] - L[6 + i] - L[9 + i] - L[12 + i] - L[15 + i] +
            L[18 + i] + L[21 + i];
    q4[i] = -L[i] - L[3 + i] + L[6 + i] + L[9 + i] - L[12 + i] - L[15 + i] +
            L[18 + i] + L[21 + i];
    q5[i] = -L[i] - L[3 + i] - L[6 + i] - L[9 + i] + L[12 + i] + L[15 + i] +
            L[18 + i] + L[21 + i];
  }
  return (triple(q0, q4, q3) + triple(q2, q0, q1) + triple(q1, q3, q5)) / 192.0 +
         triple(q2, q4, q5) / 64.0;
}

/* areaHexahedronFace, Geometry.cpp:29-64 (parallelogram test without fabs) */
double oracle_face_area(const double *L, const int *index) {
  const double *p1 = &L[3 * index[0]], *p2 = &L[3 * index[1]], *p3 = &L[3 * index[2]],
               *p4 = &L[3 * index[3]];
  double tol = 1e-6, cD[3], c1[3], c2[3];
  for (int i = 0; i < 3; ++i) {
    cD[i] = 0.25 * (p1[i] - p2[i] + p3[i] - p4[i]);
    c1[i] = 0.25 * (-p1[i] + p2[i] + p3[i] - p4[i]);
    c2[i] = 0.25 * (-p1[i] - p2[i] + p3[i] + p4[i]);
  }
  if ((cD[0] < tol) && (cD[1] < tol) && (cD[2] < tol)) return 4.0 * norm_cross(c1, c2);
  double t = sqrt(3.0) / 3.0;
  const double q[2] = {-t, t};
  double area = 0.0;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      double v1[3], v2[3];
      for (int k = 0; k < 3; ++k) {
        v1[k] = q[j] * cD[k] + c1[k];
        v2[k] = q[i] * cD[k] + c2[k];
      }
      area += norm_cross(v1, v2);
    }
  return area;
}

/* CalculateCharacteristicLength_C3D8.cpp:3-30 */
double oracle_char_length(const double *xe) {
  static const int idx[24] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 3, 7, 4,
                              1, 2, 6, 5, 0, 1, 5, 4, 3, 2, 6, 7};
  double cl = oracle_hex_volume(xe);
  double amax = 0.0;
  for (int i = 0; i < 6; ++i) {
    double fa = oracle_face_area(xe, &idx[4 * i]);
    if (fa > amax) amax = fa;
  }
  return cl / amax;
}

/* MassElementMatrix + row-sum lumping, Mass3D.cpp:5-67,134-144 */
void oracle_element_mass(const double xe[24], double rho, double m8[8]) {
  double dN[192], dJ[8], chi[24], w[8];
  oracle_shape_c3d8(xe, dN, dJ);
  oracle_gauss_c3d8(chi, w);
  // Me[(a,d),(b,e)] = rho * sum_gp (N_a N_b delta_de) w detJ, accumulated per
  // gp, scaled by rho, then row-summed over the 24 columns in column order.
  double Me[24][24];
  for (int r = 0; r < 24; ++r)
    for (int c = 0; c < 24; ++c) Me[r][c] = 0.0;
  for (int gp = 0; gp < 8; ++gp) {
    const double c = chi[3 * gp], e = chi[3 * gp + 1], z = chi[3 * gp + 2];
    double N[8] = {((1 - c) * (1 - e) * (1 - z)) / 8, ((1 + c) * (1 - e) * (1 - z)) / 8,
                   ((1 + c) * (1 + e) * (1 - z)) / 8, ((1 - c) * (1 + e) * (1 - z)) / 8,
                   ((1 - c) * (1 - e) * (1 + z)) / 8, ((1 + c) * (1 - e) * (1 + z)) / 8,
                   ((1 + c) * (1 + e) * (1 + z)) / 8, ((1 - c) * (1 + e) * (1 + z)) / 8};
    const double pre = w[gp] * dJ[gp];
    for (int r = 0; r < 24; ++r)
      for (int cc = 0; cc < 24; ++cc) {
        double v = (r % 3 == cc % 3) ? N[r / 3] * N[cc / 3] : 0.0;
        Me[r][cc] += v * pre;
      }
  }
  for (int r = 0; r < 24; ++r)
    for (int cc = 0; cc < 24; ++cc) Me[r][cc] *= rho;
  double row[24];
  for (int j = 0; j < 24; ++j) row[j] = Me[j][0];
  for (int i = 1; i < 24; ++i)
    for (int j = 0; j < 24; ++j) row[j] += Me[j][i];
  for (int a = 0; a < 8; ++a) m8[a] = row[3 * a];
}

void *oracle_create(int n_nodes, int n_elems, const double *X, const int *conn,
                    const int *pid, int n_parts, const int *material_id,
                    const double *props) {
  Model *m = new Model;
  m->nn = n_nodes;
  m->ne = n_elems;
  m->nparts = n_parts;
  m->X.assign(X, X + 3 * (size_t)n_nodes);
  m->conn.assign(conn, conn + 8 * (size_t)n_elems);
  m->pid.assign(pid, pid + n_elems);
  m->matid.assign(material_id, material_id + n_parts);
  m->props.assign(props, props + 9 * (size_t)n_parts);
  for (int p = 0; p < n_parts; ++p)
    if (material_id[p] == 5 || material_id[p] == 7) m->has_visco = true;
  m->hist.assign(m->has_visco ? (size_t)n_elems * 8 * 27 : 0, 0.0);
  m->dof_motion.assign(3 * (size_t)n_nodes, -1);
  m->u0.assign(3 * (size_t)n_nodes, 0.0);
  m->v0.assign(3 * (size_t)n_nodes, 0.0);
  build_csr(*m);
  // ShapeFunctions.cpp:207-228
  m->dNdX.assign(192 * (size_t)n_elems, 0.0);
  m->detJ.assign(8 * (size_t)n_elems, 0.0);
  for (int e = 0; e < n_elems; ++e) {
    double xe[24];
    element_coords(*m, e, nullptr, xe);
    oracle_shape_c3d8(xe, &m->dNdX[192 * (size_t)e], &m->detJ[8 * (size_t)e]);
  }
  return m;
}

void oracle_destroy(void *m) { delete (Model *)m; }

void oracle_set_threads(void *mp, int threads) {
  ((Model *)mp)->threads = threads < 1 ? 1 : threads;
}

void oracle_shape_functions(void *mp, double *dNdX, double *detJ) {
  Model &m = *(Model *)mp;
  if (dNdX) memcpy(dNdX, m.dNdX.data(), m.dNdX.size() * sizeof(double));
  if (detJ) memcpy(detJ, m.detJ.data(), m.detJ.size() * sizeof(double));
}

/* AssembleLumpedMass, Mass3D.cpp:127-157 (scatter in element order) */
void oracle_lumped_mass(void *mp, double *mass) {
  Model &m = *(Model *)mp;
  for (size_t i = 0; i < 3 * (size_t)m.nn; ++i) mass[i] = 0.0;
  for (int e = 0; e < m.ne; ++e) {
    double xe[24], m8[8];
    element_coords(m, e, nullptr, xe);
    oracle_element_mass(xe, m.props[9 * m.pid[e]], m8);
    for (int a = 0; a < 8; ++a)
      for (int d = 0; d < 3; ++d) mass[3 * m.conn[8 * e + a] + d] += m8[a];
  }
}

int oracle_element_forces(void *mp, const double *u, double dt, int update_history,
                          double *fe) {
  Model &m = *(Model *)mp;
  int err = 0;
#pragma omp parallel for schedule(static) num_threads(m.threads) reduction(max : err)
  for (int e = 0; e < m.ne; ++e) {
    double *h = nullptr;
    double tmp[8 * 27];
    if (m.has_visco) {
      h = &m.hist[(size_t)e * 216];
      if (!update_history) {
        memcpy(tmp, h, sizeof(tmp));
        h = tmp;
      }
    }
    int rc = element_force(m, e, u, dt, h, fe + 24 * (size_t)e);
    if (rc > err) err = rc;
  }
  return err;
}

/* GetForce_3D.cpp:5-53 on one rank: per-element forces, then the node sums in
 * element order (the CSR lists each node's incidences by increasing element,
 * which is the order of the reference's serial scatter-add). */
int oracle_get_force(void *mp, const double *u, double dt, double *fi) {
  Model &m = *(Model *)mp;
  std::vector<double> fe(24 * (size_t)m.ne);
  int rc = oracle_element_forces(mp, u, dt, 1, fe.data());
  if (rc) return rc;
#pragma omp parallel for schedule(static) num_threads(m.threads)
  for (int n = 0; n < m.nn; ++n) {
    double s[3] = {0.0, 0.0, 0.0};
    for (int k = m.csr_ptr[n]; k < m.csr_ptr[n + 1]; ++k) {
      int idx = m.csr_idx[k];
      for (int d = 0; d < 3; ++d) s[d] += fe[(size_t)idx * 3 + d];
    }
    for (int d = 0; d < 3; ++d) fi[3 * n + d] = s[d];
  }
  return 0;
}

/* StableTimeStep.cpp:4-40 with CalculateTimeStep.cpp:7-21 */
double oracle_stable_dt(void *mp, const double *u, const int *boundary) {
  Model &m = *(Model *)mp;
  double dtMin = 1e20;  // GlobalVariables.h:17 huge
#pragma omp parallel for schedule(static) num_threads(m.threads) reduction(min : dtMin)
  for (int e = 0; e < m.ne; ++e) {
    bool notRigid = false;
    for (int a = 0; a < 8; ++a) {
      int idx = 3 * m.conn[8 * e + a];
      if (!(boundary[idx] & boundary[idx + 1] & boundary[idx + 2])) {
        notRigid = true;
        break;
      }
    }
    if (!notRigid) continue;
    double xe[24];
    element_coords(m, e, u, xe);
    double le = oracle_char_length(xe);
    const double *p = &m.props[9 * m.pid[e]];
    double mu = p[1], lambda = p[2], rho = p[0];
    double nu = 0.5 * lambda / (lambda + mu);
    double ce = sqrt(lambda * (1.0 / nu - 1.0) / rho);
    double dte = le / ce;
    if (dte < dtMin) dtMin = dte;
  }
  return dtMin;
}

void oracle_get_history(void *mp, double *hist) {
  Model &m = *(Model *)mp;
  if (!m.hist.empty()) memcpy(hist, m.hist.data(), m.hist.size() * sizeof(double));
}
void oracle_set_history(void *mp, const double *hist) {
  Model &m = *(Model *)mp;
  if (!m.hist.empty()) memcpy(m.hist.data(), hist, m.hist.size() * sizeof(double));
}

void oracle_set_bc(void *mp, const int *dof_motion, int n_motions, const double *motions) {
  Model &m = *(Model *)mp;
  m.dof_motion.assign(dof_motion, dof_motion + 3 * (size_t)m.nn);
  m.motions.assign(motions, motions + ORACLE_MOTION_STRIDE * (size_t)n_motions);
}

void oracle_set_fibres(void *mp, const double *dir) {
  Model &m = *(Model *)mp;
  m.fibres.assign(3 * (size_t)m.ne, 0.0);
  for (int e = 0; e < m.ne; ++e) {
    const double *a = dir + 3 * (size_t)e;
    const double n = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    for (int d = 0; d < 3; ++d) m.fibres[3 * (size_t)e + d] = a[d] / n;
  }
}

void oracle_set_initial(void *mp, const double *u0, const double *v0) {
  Model &m = *(Model *)mp;
  if (u0) m.u0.assign(u0, u0 + 3 * (size_t)m.nn);
  if (v0) m.v0.assign(v0, v0 + 3 * (size_t)m.nn);
}

namespace {
/* Eavg of one element (CalculateStrain.cpp:77-97) and its principal summary
 * (CalculateStrain.cpp:8-75): max principal >= 0, min principal <= 0, max shear. */
void element_principal(const Model &m, int e, const double *u, double E[9], double P[3]) {
  const double kPi = 4.0 * atan(1.0);
  const double *dN = &m.dNdX[192 * (size_t)e];
  double x[24];
  element_coords(m, e, u, x);
  for (int i = 0; i < 9; ++i) E[i] = 0.0;
  const double pre = 0.5 / 8.0;
  for (int gp = 0; gp < 8; ++gp) {
    double F[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int k = 0; k < 8; ++k) s = s + x[3 * k + i] * dN[gp * 24 + 3 * k + j];
        F[3 * j + i] = s;
      }
    // dgemm('T','N', alpha = 1/16, beta = 1): E += pre F^T F
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += F[k + 3 * i] * F[k + 3 * j];
        E[i + 3 * j] += pre * s;
      }
  }
  E[0] -= 0.5;
  E[4] -= 0.5;
  E[8] -= 0.5;
  double a = E[0], b = E[1], c = E[2], d = E[4], ee = E[5], f = E[8];
  double p1 = b * b + c * c + ee * ee, e1, e2, e3;
  if (p1 == 0) {
    e1 = a;
    e2 = d;
    e3 = f;
  } else {
    double I1 = a + d + f;
    double I2 = a * (d + f) + d * f - b * b - c * c - ee * ee;
    double I3 = a * d * f + 2.0 * b * c * ee - b * b * f - c * c * d - ee * ee * a;
    double Q = (3.0 * I2 - I1 * I1) / 9.0;
    double R = (2.0 * I1 * I1 * I1 - 9.0 * I1 * I2 + 27.0 * I3) / 54.0;
    double theta = acos(R / sqrt(-Q * Q * Q));
    double sq = 2.0 * sqrt(-Q);
    I1 = I1 / 3.0;
    e1 = sq * cos(theta / 3.0) + I1;
    e2 = sq * cos((theta + 2.0 * kPi) / 3.0) + I1;
    e3 = sq * cos((theta + 4.0 * kPi) / 3.0) + I1;
  }
  double mx = fmax(e3, fmax(e2, e1)), mn = fmin(e3, fmin(e2, e1));
  P[0] = mx > 0.0 ? mx : 0.0;
  P[1] = mn < 0.0 ? mn : 0.0;
  P[2] = 0.5 * (mx - mn);
}

/* compute95thPercentileValue on one rank (math.cpp:227-332 falls through to
 * the brute force of math.cpp:160-196): the element at index
 * (int)(0.95 n) - 1 of the ascending order (index 0 when that is negative). */
double percentile95(const std::vector<double> &v) {
  std::vector<double> w(v);
  long k = (long)(w.size() * 0.95) - 1;
  if (k < 0) k = 0;
  std::nth_element(w.begin(), w.begin() + k, w.end());
  return w[k];
}
}  // namespace

/* InitInjuryCriteria (ex5.cpp:1253-1308): the included elements are those of
 * parts with include_part[p] != 0 (all parts when NULL); the reference leaves
 * PS_Old uninitialised, here it starts at 0. */
void oracle_injury_init(void *mp, const int *include_part) {
  Model &m = *(Model *)mp;
  m.injury = true;
  m.inj_elem.clear();
  for (int e = 0; e < m.ne; ++e)
    if (!include_part || include_part[m.pid[e]]) m.inj_elem.push_back(e);
  const size_t n = m.inj_elem.size();
  m.inj_ps_old.assign(n, 0.0);
  m.inj_psxsr.assign(n, 0.0);
  m.inj_gt15.assign(n, 0);
  m.inj_gt30.assign(n, 0);
  m.inj_psr120.assign(n, 0);
  m.inj_psxsr28.assign(n, 0);
  m.inj_in_mps95.assign(m.ne, 0);
  m.inj_in_mpsxsr95.assign(m.ne, 0);
  m.inj_volume.assign(m.ne, 0.0);
  for (int e = 0; e < m.ne; ++e) {
    double xe[24];
    element_coords(m, e, nullptr, xe);
    m.inj_volume[e] = oracle_hex_volume(xe);  // calculateVolume, CalculateCentroidAndVolume.cpp:26-37
  }
  m.maxStrain = m.maxT = m.minStrain = m.minT = m.maxShear = m.maxShearT = 0.0;
  m.maxPSxSR = m.maxTimePSxSR = m.maxMPS95 = m.maxTimeMPS95 = 0.0;
  m.maxMPSxSR95 = m.maxTimeMPSxSR95 = 0.0;
  m.maxElem = m.minElem = m.shearElem = m.maxElemPSxSR = -1;
}

/* CalculateInjuryCriterions (ex5.cpp:1311-1430) at displacement u, time and
 * step size dt. */
void oracle_injury_update(void *mp, const double *u, double time, double dt) {
  Model &m = *(Model *)mp;
  const size_t n = m.inj_elem.size();
  for (size_t j = 0; j < n; ++j) {
    const int i = m.inj_elem[j];
    double E[9], P[3];
    element_principal(m, i, u, E, P);
    if (m.maxStrain < P[0]) {
      m.maxStrain = P[0];
      m.maxElem = i;
      m.maxT = time;
    }
    if (m.minStrain > P[1]) {
      m.minStrain = P[1];
      m.minElem = i;
      m.minT = time;
    }
    if (m.maxShear < P[2]) {
      m.maxShear = P[2];
      m.shearElem = i;
      m.maxShearT = time;
    }
    if (!m.inj_gt15[j] && P[0] > 0.15) m.inj_gt15[j] = 1;
    if (!m.inj_gt30[j] && P[0] > 0.30) m.inj_gt30[j] = 1;
    const double PSR = (P[0] - m.inj_ps_old[j]) / dt;
    const double PSxSR = P[0] * PSR;
    if (m.maxPSxSR < PSxSR) {
      m.maxPSxSR = PSxSR;
      m.maxElemPSxSR = i;
      m.maxTimePSxSR = time;
    }
    if (!m.inj_psr120[j] && PSR > 120.0) m.inj_psr120[j] = 1;
    if (!m.inj_psxsr28[j] && PSxSR > 28.0) m.inj_psxsr28[j] = 1;
    m.inj_ps_old[j] = P[0];
    m.inj_psxsr[j] = PSxSR;
  }
  if (n == 0) return;
  const double mps95 = percentile95(m.inj_ps_old);
  if (mps95 > m.maxMPS95) {
    m.maxMPS95 = mps95;
    m.maxTimeMPS95 = time;
    std::fill(m.inj_in_mps95.begin(), m.inj_in_mps95.end(), 0);
    for (size_t j = 0; j < n; ++j)
      if (m.inj_ps_old[j] >= m.maxMPS95) m.inj_in_mps95[m.inj_elem[j]] = 1;
  }
  const double x95 = percentile95(m.inj_psxsr);
  if (x95 > m.maxMPSxSR95) {
    m.maxMPSxSR95 = x95;
    m.maxTimeMPSxSR95 = time;
    std::fill(m.inj_in_mpsxsr95.begin(), m.inj_in_mpsxsr95.end(), 0);
    for (size_t j = 0; j < n; ++j)
      if (m.inj_psxsr[j] >= m.maxMPSxSR95) m.inj_in_mpsxsr95[m.inj_elem[j]] = 1;
  }
}

/* Injury outputs (ex5.cpp:1040-1240): out[0..20] = maxStrain, its element,
 * time; minStrain, element, time; maxShear, element, time; maxPSxSR,
 * element, time; MPS-95 and its time; MPSxSR-95 and its time; CSDM-15,
 * CSDM-30, MPSR-120, MPSxSR-28 (volume fractions of the included volume);
 * included volume.  flags (per element, may be NULL): bit 0 included, 1 MPS
 * > 0.15, 2 MPS > 0.30, 3 PSR > 120, 4 PSxSR > 28, 5 in the MPS-95 list,
 * 6 in the MPSxSR-95 list. */
void oracle_injury_results(void *mp, double *out, int *flags) {
  Model &m = *(Model *)mp;
  double vt = 0, v15 = 0, v30 = 0, v120 = 0, v28 = 0;
  if (flags) memset(flags, 0, sizeof(int) * m.ne);
  for (size_t j = 0; j < m.inj_elem.size(); ++j) {
    const int e = m.inj_elem[j];
    const double eV = m.inj_volume[e];
    vt += eV;
    if (m.inj_gt15[j]) {
      v15 += eV;
      if (m.inj_gt30[j]) v30 += eV;
    }
    if (m.inj_psr120[j]) v120 += eV;
    if (m.inj_psxsr28[j]) v28 += eV;
    if (flags)
      flags[e] = 1 | (m.inj_gt15[j] << 1) | (m.inj_gt30[j] << 2) | (m.inj_psr120[j] << 3) |
                 (m.inj_psxsr28[j] << 4) | (m.inj_in_mps95[e] << 5) | (m.inj_in_mpsxsr95[e] << 6);
  }
  const double r[21] = {m.maxStrain, (double)m.maxElem, m.maxT, m.minStrain, (double)m.minElem, m.minT,
                        m.maxShear, (double)m.shearElem, m.maxShearT, m.maxPSxSR, (double)m.maxElemPSxSR,
                        m.maxTimePSxSR, m.maxMPS95, m.maxTimeMPS95, m.maxMPSxSR95, m.maxTimeMPSxSR95,
                        v15 / vt, v30 / vt, v120 / vt, v28 / vt, vt};
  memcpy(out, r, sizeof(r));
}

/* CalculateStrain.cpp:77-97 and CalculateMaximumPrincipalStrain
 * (CalculateStrain.cpp:8-75), with F from CalculateDeformationGradient. */
void oracle_element_strain(void *mp, const double *u, double *eavg, double *principal) {
  Model &m = *(Model *)mp;
  for (int e = 0; e < m.ne; ++e) {
    double E[9], P[3];
    element_principal(m, e, u, E, P);
    if (eavg) memcpy(eavg + 9 * (size_t)e, E, sizeof(E));
    if (principal) memcpy(principal + 3 * (size_t)e, P, sizeof(P));
  }
}

void oracle_motion_eval(const double *motion, const double X[3], double t, double u[3],
                        double v[3], double a[3]) {
  motion_eval(motion, X, t, u, v, a);
}

int oracle_run(void *mp, int max_steps, double tmax, double dt_factor, double failure_dt,
               double *u_out, double *v_out, double *a_out, double *fi_out,
               double *hist_time, double *hist_dt, double *hist_energy, int n_snap,
               const int *snap_steps, double *snap_u) {
  Model &m = *(Model *)mp;
  const size_t nd = 3 * (size_t)m.nn;
  std::vector<double> u(m.u0), v(m.v0), a(nd, 0.0), vh(nd, 0.0), fi(nd, 0.0),
      mass(nd, 0.0), u_prev(nd), a_prev(nd), fi_prev(nd);
  std::vector<int> boundary(nd, 0);
  oracle_lumped_mass(mp, mass.data());
  double Time = 0.0, Wint = 0.0, Wext = 0.0;

  // ApplyBoundaryConditions, ex9.cpp:187-248, with motions given as data
  auto apply_bc = [&](double t) {
    for (int n = 0; n < m.nn; ++n) {
      for (int d = 0; d < 3; ++d) {
        int k = m.dof_motion[3 * n + d];
        if (k < 0) continue;
        double uu[3], vv[3], aa[3];
        motion_eval(&m.motions[ORACLE_MOTION_STRIDE * k], &m.X[3 * n], t, uu, vv, aa);
        boundary[3 * n + d] = 1;
        u[3 * n + d] = uu[d];
        v[3 * n + d] = vv[d];
        a[3 * n + d] = aa[d];
      }
    }
  };
  auto accelerations = [&]() {  // CalculateAcclerations.cpp:7-11, fe == 0
    for (size_t i = 0; i < nd; ++i)
      if (!boundary[i]) a[i] = (0.0 - fi[i]) / mass[i];
  };
  auto snap = [&](int step) {
    for (int s = 0; s < n_snap; ++s)
      if (snap_steps[s] == step && snap_u) memcpy(snap_u + s * nd, u.data(), nd * sizeof(double));
  };

  apply_bc(Time);
  double dt0 = oracle_stable_dt(mp, u.data(), boundary.data());
  if (dt0 < failure_dt) return -19;  // StableTimeStep.cpp:35-38
  double dt = dt_factor * dt0;
  int rc = oracle_get_force(mp, u.data(), dt, fi.data());
  if (rc) return -rc;
  accelerations();
  if (hist_time) hist_time[0] = Time;
  if (hist_dt) hist_dt[0] = dt;
  if (hist_energy) hist_energy[0] = hist_energy[1] = hist_energy[2] = hist_energy[3] = 0.0;
  snap(0);

  int step = 0;
  while (Time < tmax && step < max_steps) {
    double t_n = Time;
    double t_np1 = Time + dt;
    Time = t_np1;
    double dt_nphalf = dt;
    double t_nphalf = 0.5 * (t_np1 + t_n);
    for (size_t i = 0; i < nd; ++i)
      vh[i] = boundary[i] ? v[i] : v[i] + (t_nphalf - t_n) * a[i];
    u_prev = u;
    a_prev = a;
    fi_prev = fi;
    for (size_t i = 0; i < nd; ++i)
      if (!boundary[i]) u[i] = u[i] + dt_nphalf * vh[i];
    apply_bc(Time);
    rc = oracle_get_force(mp, u.data(), dt, fi.data());
    if (rc) return -rc;
    accelerations();
    for (size_t i = 0; i < nd; ++i)
      if (!boundary[i]) v[i] = vh[i] + (t_np1 - t_nphalf) * a[i];
    // CheckEnergy.cpp:15-66 (one rank: every node included)
    double sWint = 0.0, sWext = 0.0, WKE = 0.0;
    for (int n = 0; n < m.nn; ++n)
      for (int j = 0; j < 3; ++j) {
        size_t i = 3 * (size_t)n + j;
        double dd = u[i] - u_prev[i];
        WKE += mass[i] * v[i] * v[i];
        if (boundary[i]) {
          double reaction = fi_prev[i] + fi[i] + mass[i] * (a[i] + a_prev[i]);
          sWext += dd * reaction;
        }
        sWint += dd * (fi_prev[i] + fi[i]);
        sWext += dd * (0.0 + 0.0);
      }
    WKE *= 0.5;
    sWint *= 0.5;
    sWext *= 0.5;
    Wint += sWint;
    Wext += sWext;
    double total = fabs(WKE + Wint - Wext);
    if (m.injury) oracle_injury_update(mp, u.data(), Time, dt);  // ex5.cpp:240
    ++step;
    double dtn = oracle_stable_dt(mp, u.data(), boundary.data());
    if (dtn < failure_dt) return -19;
    dt = dt_factor * dtn;
    if (hist_time) hist_time[step] = Time;
    if (hist_dt) hist_dt[step] = dt;
    if (hist_energy) {
      hist_energy[4 * step + 0] = Wint;
      hist_energy[4 * step + 1] = Wext;
      hist_energy[4 * step + 2] = WKE;
      hist_energy[4 * step + 3] = total;
    }
    snap(step);
  }
  if (u_out) memcpy(u_out, u.data(), nd * sizeof(double));
  if (v_out) memcpy(v_out, v.data(), nd * sizeof(double));
  if (a_out) memcpy(a_out, a.data(), nd * sizeof(double));
  if (fi_out) memcpy(fi_out, fi.data(), nd * sizeof(double));
  return step;
}

}  // extern "C"
