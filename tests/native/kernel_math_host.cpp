// Host build of the engine's per-element math (femtech_amd/csrc/ft_device.h)
// for CPU tests: the exact templated routine the element kernel runs, called
// on one element at a time and compared with the oracle.  Test code only.
#include "../../femtech_amd/csrc/ft_device.h"

#include <string.h>

namespace {
template <typename R>
struct HistArr {
  R *h;
  bool valid() const { return h != nullptr; }
  R load(int gp, int k) const { return h[gp * 18 + k]; }
  void store(int gp, int k, R v) { h[gp * 18 + k] = v; }
};

template <typename R>
int run(const double *X24, const double *U24, int mat, const double *p9, double dt, double *hist144,
        int update, double *f24, double *dte, const double *fib3) {
  ft::PartDev pp;
  memset(&pp, 0, sizeof(pp));
  pp.mat = mat;
  pp.rho = p9[0];
  pp.mu = p9[1];
  pp.lambda = p9[2];
  pp.k1 = p9[3];
  pp.k2 = p9[4];
  pp.g1 = p9[5];
  pp.t1 = p9[6];
  pp.g2 = p9[7];
  pp.t2 = p9[8];
  pp.alpha = p9[3];
  pp.kappa = p9[5];
  pp.visco = mat == 5 || mat == 7;
  double nu = 0.5 * pp.lambda / (pp.lambda + pp.mu);
  pp.ce = sqrt(pp.lambda * (1.0 / nu - 1.0) / pp.rho);
  pp.inv_ce = 1.0 / pp.ce;
  // the element kernel's gather: FP64 nodal values in cube-position order
  R xv[7][3], uv[7][3];
  for (int d = 0; d < 3; ++d) {
    double vx[8], vu[8];
    for (int p = 0; p < 8; ++p) {
      vx[p] = X24[3 * ft::p2n(p) + d];
      vu[p] = U24[3 * ft::p2n(p) + d];
    }
    ft::field_coeffs(vx, xv, d);
    ft::field_coeffs(vu, uv, d);
  }
  R hl[144];
  for (int k = 0; k < 144; ++k) hl[k] = hist144 ? R(hist144[k]) : R(0);
  HistArr<R> h{hist144 ? hl : nullptr};
  R out[24];
  double my = 1e300;
  int flags = 0;
  R fib[3] = {R(1), R(0), R(0)};
  if (fib3) {
    const double n = sqrt(fib3[0] * fib3[0] + fib3[1] * fib3[1] + fib3[2] * fib3[2]);
    for (int d = 0; d < 3; ++d) fib[d] = R(fib3[d] / n);
  }
  // the per-material instance the element kernel launches for this material
  switch (mat) {
#define FT_KM_CASE(m)                                                                     \
  case m:                                                                                 \
    ft::element_eval<R, true, true, m>(xv, uv, pp, false, dt, h, update != 0, out, my, flags, fib); \
    break;
    FT_KM_CASE(0)
    FT_KM_CASE(1)
    FT_KM_CASE(2)
    FT_KM_CASE(3)
    FT_KM_CASE(4)
    FT_KM_CASE(5)
    FT_KM_CASE(6)
    FT_KM_CASE(7)
    FT_KM_CASE(8)
#undef FT_KM_CASE
    default:
      ft::element_eval<R, true, true, -1>(xv, uv, pp, false, dt, h, update != 0, out, my, flags);
  }
  for (int k = 0; k < 24; ++k) f24[k] = out[k];
  if (hist144)
    for (int k = 0; k < 144; ++k) hist144[k] = hl[k];
  *dte = my;
  return flags;
}
}  // namespace

extern "C" int km_element(int fp32, const double *X24, const double *U24, int mat, const double *p9,
                          double dt, double *hist144, int update, double *f24, double *dte,
                          const double *fib3) {
  return fp32 ? run<float>(X24, U24, mat, p9, dt, hist144, update, f24, dte, fib3)
              : run<double>(X24, U24, mat, p9, dt, hist144, update, f24, dte, fib3);
}

// The Prony history layout of the element kernel (64-element blocks).
extern "C" long long km_hist_index(int c, int e) { return (long long)ft::hist_index(c, e); }
extern "C" long long km_hist_len(int ne) { return (long long)ft::hist_len(ne); }
