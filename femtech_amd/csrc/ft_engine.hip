// ft_engine.hip -- MI355X kernels and the C ABI of the explicit C3D8 engine.
//
// Kernels (one GPU holds the whole mesh, so the reference's ParMETIS
// partition and halo exchange collapse to a single partition):
//   k_element  fused C3D8 element kernel: gradients, F, PK2 stress, 24 nodal
//              forces written once per element, and the element's stable time
//              step folded into a block minimum + one integer atomicMin
//              (GetForce_3D.cpp:15-46, StableTimeStep.cpp:11-33).
//   k_nodal    node-to-element CSR gather of the element forces in element
//              order (deterministic, no float atomics), a = f/m, velocity and
//              displacement updates, boundary motions, and the FP64 energy
//              balance (ex9.cpp:116-152, CalculateAcclerations.cpp:7-11,
//              CheckEnergy.cpp:15-66).
//   k_mass     one-off lumped mass (Mass3D.cpp:127-157) via the same gather.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/femtech_amd.h"
#include "ft_device.h"
#include "ft_mlp.h"

using namespace ft;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

}  // namespace

// Shared with the host-side mesh tools (ft_mesh.cpp).
extern "C" __attribute__((visibility("hidden"))) void ft_internal_set_error(const char *msg) {
  g_err = msg;
}

namespace {

#define HIPCHK(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(FT_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                             \
  } while (0)

constexpr int kBlock = 256;
constexpr uint64_t kInfBits = 0x7FF0000000000000ull;
constexpr int kSkipBit = 1 << 30;
// corner pairs (FT_LAYOUT_CORNER_PAIRS): the element adds its left lane's
// xi+ corner forces to its own xi- corners / leaves its xi+ corners to its
// right lane
constexpr int kPairLeft = 1 << 28;
constexpr int kPairRight = 1 << 27;
// corner pairs without a right partner: the xi+ corners go to the node's
// free xi- slots (no element has the node at that corner: a boundary node)
constexpr int kPlusToMinus = 1 << 26;
// corner pairs, an element with neither (a partner-less element whose
// neighbour along xi lies in another material segment, or across a
// wavefront boundary of a Prony segment): its xi+ corner j (1, 2, 5, 6 ->
// j = 0..3) goes to the free xi- slot when bit kExtraFree + j is set, else
// to its extras entry, which k_pair_extras adds to the owned xi- slot
// before the node sum
constexpr int kExtra = 1 << 25;
constexpr int kExtraFree = 20;
// element tile flags: the tile's wavefronts own 63 elements each behind a
// shadow lane (corner pairs, laws without history: a shadow of a Prony
// element would read the history its owner is rewriting in the same
// launch); the element before the tile belongs to its segment and instance
// (its first wavefront's shadow lane computes it)
constexpr int kTileShadows = 4;
constexpr int kTileShadow = 2;
constexpr int kPidMask = (1 << 20) - 1;
constexpr int kMaxMotions = 64;  // distinct motions per instance (held in LDS by the nodal kernel)
constexpr int kEaccStride = 8;  // doubles per nodal block in the energy accumulators
// launch classes of the timing: element, nodal/gather, other
constexpr int kTimingClasses = 3;

// Blocks are dispatched round-robin over the 8 XCDs; remap so that each XCD
// walks one contiguous range of elements or nodes (its own L2 then sees the
// shared nodes of neighbouring elements).  gridDim.x is a multiple of 8.
__device__ inline int xcd_block() {
  int per = gridDim.x >> 3;
  return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

int grid_for(long n) {
  long nb = (n + kBlock - 1) / kBlock;
  nb = (nb + 7) & ~7L;
  return (int)std::max(nb, 8L);
}

// Blocks for n tiles: a multiple of 8 (the XCD remap), the excess exits.
int grid_for_tiles(long n) {
  long nb = (n + 7) & ~7L;
  return (int)std::max(nb, 8L);
}

struct StepSlot {
  double t;        // t_k
  double t_prev;   // t_{k-1}
  double dt;       // dt_k, the size of step k
  unsigned long long dtmin_bits;  // min_e Le/c at u_k (order-preserving bits)
  double t_done;   // time of the last completed step
  double dt_done;  // dt computed at its end (the next step's size)
  double dt_guide; // last min_e Le/c: elements near it get an FP64 dt (FP32 kernel)
  int step;        // k: step in progress (done == 0) or last completed (done == 1)
  int done;        // 1 when the run has stopped at the end of step k
  int fresh;       // 1 if the last nodal launch completed a step
  int failed;      // 1 when it stopped because no stable step was left
  // FP32 run loop: best unrefined block minimum of the step (dt_candidate),
  // 1 once k_motion_states has merged it into dtmin_bits, and the FP32 bits
  // of the smallest block minimum beyond the skip margin (far_bits)
  unsigned long long cand_bits;
  int settled;
  unsigned far_bits;
};

struct Ctl {
  StepSlot slot[2];
  int reserved;
  int hist_count;
  int pad[2];
};


__device__ inline double bits_to_double(unsigned long long b) { return __longlong_as_double((long long)b); }

// Key of an element time step for the unsigned atomicMin: the bits of a
// positive double order like the values; a step <= 0 (an inverted element:
// negative volume) maps to +0, below every positive step, so the minimum
// collapses as in the reference (dt < FailureTimeStep, StableTimeStep.cpp:
// 33-38); NaN keeps the minimum.
__device__ inline unsigned long long dt_key(double d) {
  if (d > 0.0) return (unsigned long long)__double_as_longlong(d);
  return d <= 0.0 ? 0ull : kInfBits;
}
// Candidate key of a block minimum that was not refined (FP32 run loop):
// the FP32 bits of the positive step (ordered like the values) above the
// element tile's index, so the unsigned minimum gives the smallest step and
// the first tile on ties.
constexpr unsigned long long kNoCandidate = ~0ull;
// FP32 run loop: elements whose time-step lower bound (char_length_lower)
// exceeds the last minimum by this factor skip the exact geometry.
constexpr double kSkipMargin = 1.10;
__device__ inline unsigned long long dt_candidate(double d, int tile) {
  return ((unsigned long long)__float_as_uint((float)d) << 32) | (unsigned)tile;
}

template <typename T>
__device__ inline T block_min(T v, T *sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  T r = sh[0];
  for (int i = 1; i < (int)(blockDim.x >> 6); ++i) r = fmin(r, sh[i]);
  return r;
}

// DPP move of both halves of a double (control CTRL, rows ROW_MASK): a lane
// whose source lies outside its row, or whose row is masked off, reads zero.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the wavefront (every lane active) in a fixed tree, complete in
// lane 63: prefix sums within each 16-lane row (row_shr 1, 2, 4, 8), row 0's
// total added to row 1 and row 2's to row 3 (row_bcast:15), rows 0-1's
// added to rows 2-3 (row_bcast:31).  Vector moves only, no LDS round trips.
__device__ __forceinline__ double wave_sum_lane63(double v) {
  v += dpp_f64<0x111, 0xf>(v);
  v += dpp_f64<0x112, 0xf>(v);
  v += dpp_f64<0x114, 0xf>(v);
  v += dpp_f64<0x118, 0xf>(v);
  v += dpp_f64<0x142, 0xa>(v);
  v += dpp_f64<0x143, 0xc>(v);
  return v;
}

// Fixed-order block sums of N <= 64 values with one barrier: each
// wavefront's total (wave_sum_lane63 with DPP, else the butterfly of
// block_sum through LDS lane permutes), then the wavefront totals added in
// wavefront order by lane i of the first wavefront, which returns the block
// sum of value i (the other lanes return 0).
template <int N, bool DPP>
__device__ inline double block_sum_spread(double (&v)[N], double (*sh)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if constexpr (DPP)
      v[i] = wave_sum_lane63(v[i]);
    else
      for (int o = 32; o > 0; o >>= 1) v[i] += __shfl_xor(v[i], o);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 63)
#pragma unroll
    for (int i = 0; i < N; ++i) sh[w][i] = v[i];
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x < N)
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) r += sh[k][threadIdx.x];
  return r;
}

// Fixed-order block sum (same tree every launch -> bitwise reproducible).
__device__ inline double block_sum(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  return r;
}

// ---------------------------------------------------------------------------
// Boundary motions (the reference drivers' ApplyBoundaryConditions, as data)
// ---------------------------------------------------------------------------
// A motion at one instant: the scalar part (time functions) is evaluated once
// per block and per time into LDS; the per-node part is cheap and inline.
struct MotionT {
  double s0, s1, s2, s3;  // LINEAR/RAMP: u, v scalars; ROT: cos, sin, omega, alpha
  double R[9];            // ROT: rotation matrix (row-major) of the angle at this time
};

__host__ __device__ inline void motion_time(const double *mo, double t, MotionT &m) {
  const int type = (int)mo[0];
  m.s0 = m.s1 = m.s2 = m.s3 = 0.0;
  for (int k = 0; k < 9; ++k) m.R[k] = 0.0;
  if (type == FT_MOTION_LINEAR) {  // ex9.cpp:191-243: AppliedDisp = Time*dMax/tMax
    m.s0 = t * mo[1];
    m.s1 = mo[1];
  } else if (type == FT_MOTION_RAMP_HOLD) {
    const double amp = mo[1], T = mo[2];
    const double s = t < T ? t / T : 1.0;
    m.s0 = amp * s;
    m.s1 = t < T ? amp / T : 0.0;
  } else if (type == FT_MOTION_ROT_HALFSINE) {
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
    m.s0 = cos(theta);
    m.s1 = sin(theta);
    m.s2 = omega;
    // Rodrigues: R = cos I + sin [n]x + (1 - cos) n n^T
    const double n[3] = {mo[1], mo[2], mo[3]};
    const double c = m.s0, sn = m.s1;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) m.R[3 * i + j] = (i == j ? c : 0.0) + (1.0 - c) * n[i] * n[j];
    m.R[1] -= sn * n[2];
    m.R[2] += sn * n[1];
    m.R[3] += sn * n[2];
    m.R[5] -= sn * n[0];
    m.R[6] -= sn * n[1];
    m.R[7] += sn * n[0];
    m.s3 = alpha;
  }
}

// Component d of (u, v, a) of motion `mo` at the node with reference
// position X, given the motion's scalar state m.
__host__ __device__ inline void motion_apply(const double *mo, const MotionT &m, const double *X,
                                             int d, double &u, double &v, double &a) {
  const int type = (int)mo[0];
  u = v = a = 0.0;
  if (type == FT_MOTION_LINEAR || type == FT_MOTION_RAMP_HOLD) {
    u = m.s0;
    v = m.s1;
  } else if (type == FT_MOTION_ROT_HALFSINE) {
    // r = R r0 (R from motion_time); v = w n x r; a = alpha n x r + w^2 n x (n x r)
    const double n0 = mo[1], n1 = mo[2], n2 = mo[3];
    const double r0[3] = {X[0] - mo[4], X[1] - mo[5], X[2] - mo[6]};
    double r[3];
    for (int k = 0; k < 3; ++k) r[k] = m.R[3 * k] * r0[0] + m.R[3 * k + 1] * r0[1] + m.R[3 * k + 2] * r0[2];
    const double nxr[3] = {n1 * r[2] - n2 * r[1], n2 * r[0] - n0 * r[2], n0 * r[1] - n1 * r[0]};
    const double n[3] = {n0, n1, n2};
    const int d1 = d == 2 ? 0 : d + 1, d2 = d == 0 ? 2 : d - 1;
    u = r[d] - r0[d];
    v = m.s2 * nxr[d];
    a = m.s3 * nxr[d] + m.s2 * m.s2 * (n[d1] * nxr[d2] - n[d2] * nxr[d1]);
  }
}

// Whole-vector evaluation (host set-up and read-out).
__host__ __device__ inline void motion_eval(const double *mo, const double *X, double t, double u[3],
                                            double v[3], double a[3]) {
  MotionT m;
  motion_time(mo, t, m);
  for (int d = 0; d < 3; ++d) motion_apply(mo, m, X, d, u[d], v[d], a[d]);
}

// ---------------------------------------------------------------------------
// Element kernel
// ---------------------------------------------------------------------------
// Element of an array at a 32-bit byte offset from a (uniform) base.
// The value of the lane to the left in the wavefront (DPP wave_shr:1; lane
// 0 reads zero).
__device__ __forceinline__ float lane_from_left(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ double lane_from_left(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x138, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <typename T>
__device__ __forceinline__ T &at_byte(T *base, unsigned off) {
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off);
}
template <typename T>
__device__ __forceinline__ const T &at_byte(const T *base, unsigned off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}

// Prony history in HBM: 64-element blocks (hist_index, hist_len of ft_device.h).

template <typename R>
struct HistSoA {
  R *base;
  int ne, e;
  __device__ bool valid() const { return base != nullptr; }
  __device__ R load(int gp, int k) const { return base[hist_index(gp * 18 + k, e)]; }
  __device__ void store(int gp, int k, R v) { base[hist_index(gp * 18 + k, e)] = v; }
};

// A tile is the work of one block: a range of at most kBlock consecutive
// elements (or nodes) of one instance.  int4 {begin, end, instance, flags};
// node tiles carry kTileFirst on the first tile of their instance and
// kTileMotion when one of their nodes is constrained (only those tiles stage
// the motion states).
constexpr int kTileFirst = 1;
constexpr int kTileMotion = 2;

template <typename R>
struct ElemArgs {
  int nn;                // nodes (stride of the component-major nodal arrays)
  int ne;                // all elements (history stride)
  const int4 *tiles;     // this launch's tiles (one material segment, or all)
  int n_tiles;
  const int *conn;       // 8 per element
  const int *einfo;      // part id | skip-dt flag
  const double *X;       // component-major [3][nn]
  const double *u;       // component-major [3][nn]
  const PartDev *parts;
  R *fe;                 // element forces (nullable when only dt is wanted)
  int corner;            // fe[slot][node][xyz] instead of fe[slot][element][xyz]
  const unsigned *eslot; // node slots: rank of the element at local node k in bits 4k..4k+3 (nullptr: slot k)
  R *fe_ovf;             // node slots: forces of ranks 8-15, [rank - 8][overflow node][xyz]
  const int *ovf;        // per node: (overflow node index << 4) | ranks beyond 8, 0 for none
  int n_ovf;             // overflow nodes
  R *hist;               // 144 per element, in 64-element blocks (hist_index)
  const R *fib;          // fibre directions [3][ne] (material 8)
  Ctl *ctl;              // run loop: per-instance control, slot `parity` gives dt_k and done
  int parity;
  int refine;            // FP32 kernel in the run loop: FP64 dt of near-minimum elements
  int pair;              // corner pairs (kPairLeft / kPairRight in einfo)
  R *xf;                 // corner pairs: extras values, 4 x 3 per kExtra element
  const int *xidx;       // corner pairs: per element, its first extras entry (kExtra elements)
  double *cand_geom;       // FP32 run loop: kCandStride per element tile (candidate geometry)
  // per-Gauss-point fields (k_element_gp): 72 / 8 / 48 doubles per element in
  // the caller's element order (perm: internal -> caller)
  double *gp_F, *gp_detF, *gp_pk2;
  const int *perm;
  double dt_fixed;       // used outside the run loop (ctl == nullptr)
  unsigned long long *dtbits;  // outside the run loop: per-instance minimum (nullable)
  int *status;
  int want_force;
  int want_dt;
  int update_hist;
  // constitutive surrogate (materials 7, 8): network outputs of this
  // segment's Gauss points [8][n_seg][8], its first element, the
  // normalisation block, audit switch and share, per-tile tallies
  const float *surr_y;
  size_t surr_stride;
  int surr_e0;
  const float *surr_norm;
  int surr_audit;
  unsigned surr_permille;
  double *surr_tally;
  int tile0;
};

// Occupancy target (waves per SIMD; 4 = at most 128 VGPRs) of the default
// instances: the FP32 laws written in closed form are held to 4; the Prony
// laws (5, 7), Ogden's eigen-decomposition (6), the fibre law (8) and the
// FP64 instances are not held.  WAVES > 0 overrides it (the FP32 run-loop
// instances are compiled for WAVES = 1..5 and the tuner picks one per
// material segment).
#ifndef FT_ELEM_WAVES_FP32
#define FT_ELEM_WAVES_FP32 4
#endif
template <typename R, int MAT, int WAVES>
constexpr int elem_waves() {
  return WAVES > 0 ? WAVES : ((sizeof(R) == 4 && MAT <= 4 && MAT != -1) ? FT_ELEM_WAVES_FP32 : 1);
}
constexpr int kMaxTuneWaves = 5;

// Elements audited against the analytic law (surrogate runs): a fixed
// pseudo-random share of permille / 1000.
__device__ inline bool audited(int e, unsigned permille) {
  return ((unsigned)e * 2654435761u >> 8) % 1000u < permille;
}

// FP64 normalised Walsh-Hadamard coefficients c[m][d] (m = xi, eta, xi.eta,
// zeta, xi.zeta, eta.zeta, xi.eta.zeta) of the deformed position x = X + u
// of element es, written by threads 0..2 (one component each) to c.
__device__ inline void deformed_coeffs(const int *conn, const double *X, const double *u, int nn, int es,
                                       double *c) {
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    const double *Xc = X + (size_t)d * nn, *Uc = u + (size_t)d * nn;
    const int n0 = conn[8 * (size_t)es];
    double v[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const int n = conn[8 * (size_t)es + p2n(p)];
      v[p] = p == 0 ? 0.0 : (Xc[n] - Xc[n0]) + (Uc[n] - Uc[n0]);
    }
    wht8(v);
#pragma unroll
    for (int m = 0; m < 7; ++m) c[3 * m + d] = v[m + 1] * 0.125;
  }
}

// Le / ce (StableTimeStep.cpp:22-27, CalculateCharacteristicLength_C3D8.cpp:
// 3-30) in FP64 from coefficients sh_c (filled by deformed_coeffs, LDS),
// returned to every thread.  Called by every thread of the block; the six
// face areas and the volume go to threads 0..6, so it adds no registers to
// the calling kernel.
__device__ inline double refined_dt(const double (*sh_c)[3], double ce) {
  __shared__ double sh_a[7];
  __syncthreads();
  if (threadIdx.x < 7) sh_a[threadIdx.x] = hex_part(sh_c, threadIdx.x);
  __syncthreads();
  double amax = 0.0;
  for (int f = 0; f < 6; ++f)
    if (sh_a[f] > amax) amax = sh_a[f];
  const double dte = sh_a[6] / amax / ce;
  __syncthreads();  // sh_a may be rewritten by the next call
  return dte;
}

// Geometry of a block's unrefined minimum (FP32 run loop), one slot per
// element tile: the 21 coefficients of deformed_coeffs and the wave speed.
constexpr int kCandStride = 22;

template <typename R, bool FORCE, bool DT, int MAT, bool SURR = false, bool GPOUT = false>
__device__ __forceinline__ void element_tile(const ElemArgs<R> &a, const int L) {
  __shared__ double sh_min[kBlock / 64];
  if (L >= a.n_tiles) return;  // grid padding (uniform over the block)
  const int4 tile = a.tiles[L];
  const int end = tile.y;
  // Corner pairs: each wavefront owns 63 consecutive elements (lanes 1-63)
  // and its lane 0 recomputes the element before them (a shadow: it writes
  // nothing, but hands its xi+ corner forces to lane 1), so every pair of
  // face neighbours along xi meets in one wavefront.
  const int lane = threadIdx.x & 63;
  int e = tile.x + threadIdx.x;
  bool owner = true, active = e < end;
  if (a.pair && (tile.w & kTileShadows)) {  // uniform
    e = tile.x + (int)(threadIdx.x >> 6) * 63 + lane - 1;
    owner = lane != 0;
    active = owner ? e < end : (e + 1 < end && ((threadIdx.x >> 6) != 0 || (tile.w & kTileShadow)));
  }
  double dt_step = a.dt_fixed;
  double dt_guide = 1e300;  // no previous minimum: refine every element
  unsigned long long *dtmin = a.dtbits ? a.dtbits + tile.z : nullptr;
  if (a.ctl) {
    const StepSlot *slot = &a.ctl[tile.z].slot[a.parity];
    if (slot->done) return;  // uniform over the block
    dt_step = slot->dt;
    dt_guide = slot->dt_guide;
    dtmin = &a.ctl[tile.z].slot[a.parity].dtmin_bits;
  }
  double my_dt = 1e300;
  int flags = 0;
  SurrTally tally = {{0, 0, 0, 0, 0, 0}}, shadow_tally = tally;
  if (active) {
    const int4 c0 = reinterpret_cast<const int4 *>(a.conn)[2 * e];
    const int4 c1 = reinterpret_cast<const int4 *>(a.conn)[2 * e + 1];
    const int nd[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    const int info = a.einfo[e];
    const PartDev pp = a.parts[info & kPidMask];
    // Monomial coefficients of the reference coordinates and displacements
    // (field_coeffs).  Uniform per-component bases and one 32-bit byte offset
    // per node (the loads take a scalar base and a vector offset: no
    // per-load 64-bit address arithmetic; the mesh size limit keeps 24 nn
    // below 2^31).
    const double *Xd[3] = {a.X, a.X + a.nn, a.X + 2 * (size_t)a.nn};
    const double *Ud[3] = {a.u, a.u + a.nn, a.u + 2 * (size_t)a.nn};
    unsigned off[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) off[p] = 8u * (unsigned)nd[p2n(p)];
    R cx[7][3], cu[7][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double vx[8], vu[8];
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        vx[p] = at_byte(Xd[d], off[p]);
        vu[p] = at_byte(Ud[d], off[p]);
      }
      field_coeffs(vx, cx, d);
      field_coeffs(vu, cu, d);
    }
    HistSoA<R> hist{a.hist, a.ne, e};
    R outv[24];
    R fib[3] = {R(1), R(0), R(0)};
    if (MAT == 8 && a.fib) {
      fib[0] = a.fib[e];
      fib[1] = a.fib[(size_t)a.ne + e];
      fib[2] = a.fib[2 * (size_t)a.ne + e];
    }
    if constexpr (SURR) {
      const SurrIn sin{a.surr_y + (size_t)(e - a.surr_e0) * kSurrOut, a.surr_stride, a.surr_norm,
                       a.surr_audit != 0 && audited(e, a.surr_permille)};
      element_eval<R, FORCE, DT, MAT, HistSoA<R>, true>(cx, cu, pp, (info & kSkipBit) != 0, dt_step, hist,
                                                       a.update_hist != 0 && owner, outv, my_dt, flags, fib, &sin,
                                                       owner ? &tally : &shadow_tally);
    } else if constexpr (GPOUT) {
      const size_t ce = (size_t)a.perm[e];
      const GpSink gs{a.gp_F + 72 * ce, a.gp_detF + 8 * ce, a.gp_pk2 + 48 * ce};
      element_eval<R, FORCE, DT, MAT, HistSoA<R>, false, true>(cx, cu, pp, (info & kSkipBit) != 0, dt_step, hist,
                                                              a.update_hist != 0, outv, my_dt, flags, fib,
                                                              nullptr, nullptr, &gs);
    } else {
      // FP32 run loop: exact lengths only within kSkipMargin of the last minimum
      const double exact_below = sizeof(R) == 4 && a.refine ? dt_guide * kSkipMargin : 1e300;
      element_eval<R, FORCE, DT, MAT>(cx, cu, pp, (info & kSkipBit) != 0, dt_step, hist,
                                 a.update_hist != 0 && owner, outv, my_dt, flags, fib, nullptr, nullptr, nullptr,
                                 exact_below, a.ctl == nullptr);
    }
    if (FORCE) {
      // Slot-major layouts: fe[slot][element][xyz] or fe[corner][node][xyz].
      // Each of the 8 stores of a wavefront covers one contiguous run (for
      // the corner layout: neighbouring elements of a block have
      // neighbouring corner-k nodes), and so does each load of the node sum.
      if (a.eslot) {
        // node slots: the slot varies per lane (the element's rank at node
        // k); ranks 8-15 of the few nodes with more than 8 elements go to the
        // overflow slots
        const unsigned sl = a.eslot[e];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const unsigned r = (sl >> (4 * k)) & 15u;
          R *dst = r < 8u ? a.fe + 3 * ((size_t)r * a.nn + (size_t)nd[k])
                          : a.fe_ovf + 3 * ((size_t)(r - 8u) * a.n_ovf + (size_t)(a.ovf[nd[k]] >> 4));
          dst[0] = outv[3 * k];
          dst[1] = outv[3 * k + 1];
          dst[2] = outv[3 * k + 2];
        }
      } else {
        // corner pairs: the xi+ corners go to the right lane, to the node's
        // free xi- slot, or to the element's extras entry
        bool right = false, to_minus = false, extra = false;
        if (a.pair) {  // uniform (corner slots)
          // C3D8 corners on the face xi = -1 (0, 3, 4, 7) against the same
          // corners of the left neighbour's face xi = +1 (1, 2, 5, 6); the
          // lane to the left is active whenever this one has a left partner
          // (a shadow lane 0, or an owner lane)
          const bool left = (info & kPairLeft) != 0;
          right = (info & kPairRight) != 0;
          to_minus = (info & kPlusToMinus) != 0;
          extra = (info & kExtra) != 0;
          constexpr int km[4] = {0, 3, 4, 7}, kp[4] = {1, 2, 5, 6};
#pragma unroll
          for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              const R w = lane_from_left(outv[3 * kp[j] + d]);
              if (left) outv[3 * km[j] + d] += w;
            }
        }
        if (owner) {
          const int xb = extra ? a.xidx[e] : 0;
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const bool plus = ((k + 1) & 2) != 0;  // k = 1, 2, 5, 6
            if (plus && right) continue;
            const int j = (k >> 1 & 2) | (k == 2 || k == 6);  // 1, 2, 5, 6 -> 0, 1, 2, 3
            if (plus && extra && !((info >> (kExtraFree + j)) & 1)) {
              R *x = a.xf + 3 * ((size_t)xb + j);
              x[0] = outv[3 * k];
              x[1] = outv[3 * k + 1];
              x[2] = outv[3 * k + 2];
              continue;
            }
            // uniform slot bases (k, and for the free xi- slot k ^ 1)
            R *dst = a.fe + 3 * (size_t)k * (a.corner ? a.nn : a.ne);
            if (plus && (to_minus || extra)) dst = a.fe + 3 * (size_t)(k ^ 1) * a.nn;
            const unsigned o = 3u * (unsigned)sizeof(R) * (unsigned)(a.corner ? nd[k] : e);
            at_byte(dst, o) = outv[3 * k];
            at_byte(dst, o + (unsigned)sizeof(R)) = outv[3 * k + 1];
            at_byte(dst, o + 2u * (unsigned)sizeof(R)) = outv[3 * k + 2];
          }
        }
      }
    }
  }
  if (DT) {
    const double bmin = block_min(my_dt, sh_min);
    // FP32 run loop: the step size sets the time grid and is sensitive to
    // round-off, so it always comes from FP64 geometry.  A block whose
    // minimum lies within 0.1 % of the last step's global minimum refines
    // its minimising element (smallest id on ties) in FP64 at once; every
    // other block posts (FP32 minimum, tile) to the instance's candidate key
    // and the FP64 coefficients of its minimising element to its tile's
    // slot, and the nodal kernel refines the best candidate where it could
    // undercut the refined minimum (the minimum moved up by more than the
    // window in one step).
    const bool exact = sizeof(R) == 4 && a.refine && a.ctl;
    const bool refine_blk = sizeof(R) == 4 && a.refine && bmin < 1e300 && bmin <= dt_guide * 1.001;
    // A block whose minimum is beyond the skip margin holds no element that
    // can set the step unless the minimum rises past the margin in one step;
    // it posts its value only, for the conservative rule of merge_candidate.
    const bool far = exact && bmin < 1e300 && bmin > 0.0 && bmin > dt_guide * kSkipMargin;
    if (far) {
      if (threadIdx.x == 0) atomicMin(&a.ctl[tile.z].slot[a.parity].far_bits, __float_as_uint((float)bmin));
    } else if (refine_blk || (exact && bmin < 1e300)) {
      __shared__ int sh_arg;
      __shared__ double sh_c[7][3];
      if (threadIdx.x == 0) sh_arg = INT32_MAX;
      __syncthreads();
      if (active && my_dt == bmin) atomicMin(&sh_arg, e);
      __syncthreads();
      const int es = sh_arg;
      const double ce = a.parts[a.einfo[es] & kPidMask].ce;
      if (refine_blk) {
        deformed_coeffs(a.conn, a.X, a.u, a.nn, es, &sh_c[0][0]);
        const double dte = refined_dt(sh_c, ce);
        if (threadIdx.x == 0) atomicMin(dtmin, dt_key(dte));
      } else if (bmin > 0.0) {
        // the tile's candidate: its FP32 minimum (ordered key) and geometry;
        // the nodal kernel refines the instance's best one when it could
        // undercut the refined minimum
        const int gt = a.tile0 + L;
        double *slot = a.cand_geom + (size_t)kCandStride * gt;
        deformed_coeffs(a.conn, a.X, a.u, a.nn, es, slot);
        if (threadIdx.x == 0) {
          slot[21] = ce;
          atomicMin(&a.ctl[tile.z].slot[a.parity].cand_bits, dt_candidate(bmin, gt));
        }
      } else if (threadIdx.x == 0) {
        atomicMin(dtmin, dt_key(bmin));  // an inverted element: the step collapses
      }
    } else if (threadIdx.x == 0 && bmin < 1e300 && dtmin) {
      atomicMin(dtmin, dt_key(bmin));
    }
  }
  if constexpr (SURR) {
    // per-tile tallies (each tile is one block's: no atomics)
    __shared__ double sh_t[kBlock / 64];
    for (int k = 0; k < 6; ++k) {
      const double v = block_sum(tally.v[k], sh_t);
      if (threadIdx.x == 0) a.surr_tally[(size_t)(a.tile0 + L) * 6 + k] += v;
    }
  }
  if (flags) atomicOr(a.status, flags);
}

template <typename R, bool FORCE, bool DT, int MAT, int WAVES>
__global__ __launch_bounds__(kBlock, (elem_waves<R, MAT, WAVES>())) void k_element(ElemArgs<R> a) {
  element_tile<R, FORCE, DT, MAT>(a, xcd_block());
}

// Forces plus FemTech's per-Gauss-point F, det F and PK2 (every law through
// its PK2 form; the compatibility layer's GetForce_3D).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_element_gp(ElemArgs<R> a) {
  element_tile<R, true, false, -1, false, true>(a, xcd_block());
}

// Run-loop element kernel of a surrogate segment (FP32, force + dt).
template <int MAT>
__global__ __launch_bounds__(kBlock) void k_element_surr(ElemArgs<float> a) {
  element_tile<float, true, true, MAT, true>(a, xcd_block());
}

// grad_X u at Gauss point gp from the normalised coefficients of x and u
// (the Jacobians of element_eval).
template <typename R>
__device__ inline void gp_grad_u(const R cx[7][3], const R cu[7][3], int gp, R H[3][3]) {
  const R g = R(kGp), g2 = R(kGp * kGp);
  const R sx = ((gp & 3) == 1 || (gp & 3) == 2) ? R(1) : R(-1);
  const R sy = (gp & 2) ? R(1) : R(-1);
  const R sz = gp < 4 ? R(1) : R(-1);
  R J0[3][3], Ju[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    J0[d][0] = cx[0][d] + (sy * g) * cx[2][d] + (sz * g) * cx[4][d] + (sy * sz * g2) * cx[6][d];
    J0[d][1] = cx[1][d] + (sx * g) * cx[2][d] + (sz * g) * cx[5][d] + (sx * sz * g2) * cx[6][d];
    J0[d][2] = cx[3][d] + (sx * g) * cx[4][d] + (sy * g) * cx[5][d] + (sx * sy * g2) * cx[6][d];
    Ju[d][0] = cu[0][d] + (sy * g) * cu[2][d] + (sz * g) * cu[4][d] + (sy * sz * g2) * cu[6][d];
    Ju[d][1] = cu[1][d] + (sx * g) * cu[2][d] + (sz * g) * cu[5][d] + (sx * sz * g2) * cu[6][d];
    Ju[d][2] = cu[3][d] + (sx * g) * cu[4][d] + (sy * g) * cu[5][d] + (sx * sy * g2) * cu[6][d];
  }
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
  const R idet0 = R(1) / det0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) H[i][j] = (Ju[i][0] * A[0][j] + Ju[i][1] * A[1][j] + Ju[i][2] * A[2][j]) * idet0;
}

// Coefficients of x - x_0 and u - u_0 of element e (node-0 relative, FP64
// differences rounded to R) and its fibre.
template <typename R>
__device__ inline void element_coeffs(int nn, int ne, const int *conn, const double *X, const double *u,
                                      const float *fibf, const double *fibd, int e, R cx[7][3], R cu[7][3],
                                      R fib[3]) {
  const int *ce = conn + 8 * (size_t)e;
  const int n0 = ce[0];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double *Xc = X + (size_t)d * nn, *Uc = u + (size_t)d * nn;
    R xv[8], uv[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const int n = ce[p2n(p)];
      xv[p] = p == 0 ? R(0) : R(Xc[n] - Xc[n0]);
      uv[p] = p == 0 ? R(0) : R(Uc[n] - Uc[n0]);
    }
    wht8(xv);
    wht8(uv);
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      cx[m][d] = xv[m + 1] * R(0.125);
      cu[m][d] = uv[m + 1] * R(0.125);
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
    fib[d] = fibd ? R(fibd[(size_t)d * ne + e]) : (fibf ? R(fibf[(size_t)d * ne + e]) : (d == 0 ? R(1) : R(0)));
}

// Training samples of the surrogate at the current displacement: raw
// features and the analytic S_iso in the material frame (FP64 arithmetic) of
// every Gauss point of elements [e0, e1) of material MAT, Gauss-point major
// from `out0` (out index out0 + gp (e1 - e0) + e - e0).
template <int MAT>
__global__ __launch_bounds__(kBlock) void k_surr_sample(int nn, int ne, const int *conn, const double *X,
                                                        const double *u, const float *fibf, const double *fibd,
                                                        const PartDev *parts, const int *einfo, int e0, int e1,
                                                        long out0, float *xs, float *ts) {
  const int e = e0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= e1) return;
  double cx[7][3], cu[7][3], fib[3];
  element_coeffs<double>(nn, ne, conn, X, u, fibf, fibd, e, cx, cu, fib);
  const PartDev pp = parts[einfo[e] & kPidMask];
  double Q[3][3];
  material_frame<double>(MAT, fib, Q);
  for (int gp = 0; gp < 8; ++gp) {
    double H[3][3];
    gp_grad_u<double>(cx, cu, gp, H);
    const double Jm1 = det_ip1_m1(H);
    double f[kSurrFeat];
    surr_features<double>(H, Jm1, Q, MAT, f);
    Sym<double> Siso;
    double p;
    iso_pk2<double, MAT>(pp, H, Jm1, fib, Siso, p);
    const Sym<double> Sm = rotate_sym(Siso, Q, true);
    const size_t o = (size_t)out0 + (size_t)gp * (e1 - e0) + (e - e0);
#pragma unroll
    for (int k = 0; k < kSurrFeat; ++k) xs[o * kSurrFeat + k] = (float)f[k];
#pragma unroll
    for (int k = 0; k < kSurrOut; ++k) ts[o * kSurrOut + k] = k < 6 ? (float)Sm.v[k] : 0.0f;
  }
}

// Network inputs of a surrogate segment's Gauss points at the run loop's
// displacement (FP32, as the element kernel forms them), normalised, BF16,
// Gauss-point major.
template <int MAT>
__global__ __launch_bounds__(kBlock) void k_surr_features(int nn, int ne, const int *conn, const double *X,
                                                          const double *u, const float *fibf, int e0, int e1,
                                                          const float *norm, __bf16 *feat, const Ctl *ctl,
                                                          int parity) {
  const int e = e0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= e1) return;
  if (ctl[0].slot[parity].done) return;
  float cx[7][3], cu[7][3], fib[3];
  element_coeffs<float>(nn, ne, conn, X, u, fibf, nullptr, e, cx, cu, fib);
  float Q[3][3];
  material_frame<float>(MAT, fib, Q);
#pragma unroll 1
  for (int gp = 0; gp < 8; ++gp) {
    float H[3][3];
    gp_grad_u<float>(cx, cu, gp, H);
    const float Jm1 = det_ip1_m1(H);
    float f[kSurrFeat];
    surr_features<float>(H, Jm1, Q, MAT, f);
    __bf16 *dst = feat + ((size_t)gp * (e1 - e0) + (e - e0)) * kSurrFeat;
#pragma unroll
    for (int k = 0; k < kSurrFeat; ++k) dst[k] = (__bf16)((f[k] - norm[k]) / norm[16 + k]);
  }
}

// ---------------------------------------------------------------------------
// Node kernels
// ---------------------------------------------------------------------------
// Node-to-element incidences in ELL form: entry j of node n at ell[j * nn + n]
// (slot * ne + element, or -1 past the node's valence), in the original
// element order, so the sum below adds the element forces in the order of the
// reference's scatter-add (GetForce_3D.cpp:39-44).  W is the ELL width when
// it is the common 8 (fully unrolled: 8 independent loads in flight).
// Node slots beyond 8 (nodes with more than 8 elements): their forces, in
// rank order after the first 8, so the sum keeps the element order.
template <typename R>
struct Overflow {
  const R *fe;     // [rank - 8][overflow node][xyz]
  const int *idx;  // per node: (overflow node index << 4) | ranks beyond 8; nullptr: none in the mesh
  int n;           // overflow nodes
};

template <typename R, int W>
__device__ inline void gather3(const int *ell, int nn, int width, const R *fe, int n, double f[3],
                               const Overflow<R> &ov = Overflow<R>{nullptr, nullptr, 0}) {
  f[0] = f[1] = f[2] = 0.0;
  if (W == -2) {
    // Corner pairs, every element paired with its right neighbour or
    // writing its xi+ corners to free xi- slots: the four xi- slots hold
    // the node's whole force.
    R v[4][3];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = (0x7430 >> (4 * j)) & 15;  // 0, 3, 4, 7
      const R *src = fe + 3 * (size_t)k * nn;  // uniform
      const unsigned o = 3u * (unsigned)n;
      v[j][0] = src[o];
      v[j][1] = src[o + 1];
      v[j][2] = src[o + 2];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      f[0] += (double)v[j][0];
      f[1] += (double)v[j][1];
      f[2] += (double)v[j][2];
    }
  } else if (W < 0) {
    // Corner-slot layout fe[corner][node][xyz]: the element that has node n
    // as its local corner k wrote slot k (zero where no element does; with
    // corner pairs, zero also where the element's right partner wrote the
    // pair's sum into its own slot).
    R v[8][3];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const R *src = fe + 3 * (size_t)k * nn;  // uniform
      const unsigned o = 3u * (unsigned)n;
      v[k][0] = src[o];
      v[k][1] = src[o + 1];
      v[k][2] = src[o + 2];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      f[0] += (double)v[k][0];
      f[1] += (double)v[k][1];
      f[2] += (double)v[k][2];
    }
    if (ov.idx) {  // uniform: the mesh has nodes with more than 8 elements
      const int o = ov.idx[n];
      for (int r = 0; r < (o & 15); ++r) {
        const R *src = ov.fe + 3 * ((size_t)r * ov.n + (size_t)(o >> 4));
        f[0] += (double)src[0];
        f[1] += (double)src[1];
        f[2] += (double)src[2];
      }
    }
  } else if (W > 0) {
    constexpr int WA = W > 0 ? W : 1;
    int id[WA];
#pragma unroll
    for (int j = 0; j < W; ++j) id[j] = ell[(size_t)j * nn + n];
    R v[WA][3];
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const size_t o = 3 * (size_t)(id[j] < 0 ? 0 : id[j]);
      v[j][0] = id[j] < 0 ? R(0) : fe[o];
      v[j][1] = id[j] < 0 ? R(0) : fe[o + 1];
      v[j][2] = id[j] < 0 ? R(0) : fe[o + 2];
    }
#pragma unroll
    for (int j = 0; j < W; ++j) {
      if (id[j] < 0) continue;
      f[0] += (double)v[j][0];
      f[1] += (double)v[j][1];
      f[2] += (double)v[j][2];
    }
  } else {
    for (int j = 0; j < width; ++j) {
      const int id = ell[(size_t)j * nn + n];
      if (id < 0) break;
      const size_t o = 3 * (size_t)id;
      f[0] += (double)fe[o];
      f[1] += (double)fe[o + 1];
      f[2] += (double)fe[o + 2];
    }
  }
}

template <typename R, int W>
__global__ __launch_bounds__(kBlock) void k_gather(int nn, const int *ell, int width, const R *fe,
                                                   double *fi, Overflow<R> ov) {
  const int n = xcd_block() * kBlock + threadIdx.x;
  if (n >= nn) return;
  double f[3];
  gather3<R, W>(ell, nn, width, fe, n, f, ov);
  fi[n] = f[0];
  fi[(size_t)nn + n] = f[1];
  fi[2 * (size_t)nn + n] = f[2];
}

enum NodalMode { kInitial = 0, kStep = 1, kResume = 2 };

template <typename R>
struct NodalArgs {
  int nn;
  const int *ell;
  int width;
  const R *fe;
  Overflow<R> ov;         // node slots beyond 8
  const double *X;
  double *u, *vs, *fi;
  const double *mass;  // 1 per node
  const int *node_code;  // per node: byte d = motion index + 1 of DOF d (0 free)
  const double *motions;  // per-instance copies, instance i at motion mrange[2 i]
  int n_motions;          // total (stride of mstate)
  const int *mrange;      // per instance: first motion, count (<= kMaxMotions)
  const MotionT *mstate;  // [3][n_motions], from k_motion_states
  const int4 *tiles;      // node tiles
  int n_tiles;
  Ctl *ctl;               // per instance; ctl[0] also holds the status word
  int parity;
  int mode;
  int energy;
  double *eacc;  // kEaccStride per tile: W_int, W_ext, KE, and the fi_{k-1}
                 // halves of the next step's W_int and W_ext
  const double *tmax;     // per instance
  const int *cap;         // per instance: last step this call may complete
  double factor, failure_dt;
  int *status;            // engine status word (FT_STATUS_* bits)
  int inline_motions;     // motion states evaluated by the tile itself (mstate unused)
  int *running;           // nullable: the first tile of each instance that goes on adds 1
  int mcap;               // largest motion count of any instance (>= 1): dynamic LDS rows
  const double *cand_geom;  // FP32 engine: candidate geometry per element tile (element kernel)
};

// Dynamic LDS of a node tile of a run with rotations: the motion states of
// its instance at the three instants, [3][mcap] MotionT, then the motion
// parameters, [mcap][8] (runs without rotations use a fixed-size table).
template <bool ROT>
size_t nodal_lds_bytes(int mcap) {
  return ROT ? 3 * (size_t)mcap * sizeof(MotionT) + (size_t)mcap * 8 * sizeof(double) : 0;
}

// Time of motion state w (0: t_{k-1}, 1: t_k, 2: t_{k+1} = t_k + factor *
// min Le/c) for the nodal update of step k, rounded step by step (the same
// bits wherever it is evaluated).
__device__ inline double motion_state_time(const StepSlot &cur, double factor, int w, double dtmin) {
  if (w == 0) return cur.t_prev;
  if (w == 1) return cur.t;
  return __dadd_rn(cur.t, __dmul_rn(factor, dtmin));
}

// Whether the step's best unrefined candidate (FP32 run loop) could undercut
// the refined minimum dtmin.
__device__ inline bool candidate_pending(const StepSlot &cur, double dtmin) {
  return cur.cand_bits != kNoCandidate &&
         (double)__uint_as_float((unsigned)(cur.cand_bits >> 32)) <= dtmin * 1.001;
}

// The candidate's FP64 step (the geometry its tile left), by one thread:
// the same parts as refined_dt, evaluated in turn.
__device__ inline double candidate_dt(const double *cg) {
  double c[7][3];
  for (int i = 0; i < 21; ++i) c[i / 3][i % 3] = cg[i];
  double amax = 0.0;
  for (int f = 0; f < 6; ++f) {
    const double af = hex_part(c, f);
    if (af > amax) amax = af;
  }
  return hex_part(c, 6) / amax / cg[21];
}

// The step's minimum with the best unrefined candidate's FP64 step dte.
__device__ inline double merge_candidate(double dtmin, double dte) {
  if (dte <= 0.0) return 0.0;  // inverted in FP64: the step collapses
  return dte < dtmin ? dte : dtmin;
}

// If the step's minimum d rose past the skip margin over the last minimum,
// elements that reported only a lower bound (their own, or their block's
// beyond the margin) could undercut it: the smallest such value, a lower
// bound of every one of their steps, is taken instead (conservative; *bounded
// set).  When no element reported anything finite (non-finite forces
// everywhere), d stays infinite and the step collapses.
__device__ inline double bound_rule(const StepSlot &cur, double d, bool *bounded) {
  if (!(d > cur.dt_guide * kSkipMargin)) return d;
  double cb = cur.cand_bits != kNoCandidate ? (double)__uint_as_float((unsigned)(cur.cand_bits >> 32)) : 1e300;
  if (cur.far_bits < 0x7f800000u) cb = fmin(cb, (double)__uint_as_float(cur.far_bits));
  if (cb < 1e300 && cb < d) {
    *bounded = true;
    return cb;
  }
  return d;
}

// End of step k (gather, a = f/m, v = v_half + dt/2 a, energy) fused with the
// start of step k+1 (v_half, u += dt v_half, boundary motions at t_{k+1}),
// ex9.cpp:107-175.  vs holds v_half(k) for free DOFs and the prescribed v
// for constrained ones while a run is in progress, v(k) once it stops.
// ROT: some motion of the run is a rotation; its parameters are then staged
// in LDS (a rotation DOF evaluates them three times per step), which costs
// the other runs a little occupancy.
template <typename R, int W, bool ROT = true>
__device__ __forceinline__ void nodal_tile(const NodalArgs<R> &a, const int L) {
  __shared__ double sh[kBlock / 64];
  if (L >= a.n_tiles) return;  // grid padding
  const int4 tile = a.tiles[L];
  const int inst = tile.z;
  Ctl *ic = a.ctl + inst;
  const StepSlot cur = ic->slot[a.parity];
  // the instance's end time, step cap and motion range, read before the
  // kernel's first store: uniform loads that nothing in the launch writes
  // go through the scalar cache, together with the slot, instead of one
  // vector round trip after another
  const double tmax = a.tmax[inst];
  const int cap = a.cap[inst];
  const int mb = a.mrange[2 * inst], mc = a.mrange[2 * inst + 1];
  if (cur.done && a.mode != kResume) {
    // keep both ping-pong slots stopped so later launches stay no-ops
    if ((tile.w & kTileFirst) && threadIdx.x == 0) {
      StepSlot nx = cur;
      nx.fresh = 0;
      ic->slot[a.parity ^ 1] = nx;
    }
    return;
  }
  const int n = tile.x + threadIdx.x;
  const double t_k = cur.t, t_km1 = cur.t_prev, dt_k = cur.dt;
  double dtmin = bits_to_double(cur.dtmin_bits);
  if (sizeof(R) == 4 && a.cand_geom && !cur.settled) {
    // FP32 engine, unless k_motion_states has settled the step already: the
    // step's best unrefined element could undercut the refined minimum (it
    // moved up by more than the element kernel's window): its FP64 step from
    // the geometry its tile left (uniform over the block; every block of the
    // instance computes the same value)
    if (candidate_pending(cur, dtmin)) {
      __shared__ double sh_c[7][3];
      const double *cg = a.cand_geom + (size_t)kCandStride * (unsigned)(cur.cand_bits & 0xffffffffu);
      if (threadIdx.x < 21) (&sh_c[0][0])[threadIdx.x] = cg[threadIdx.x];
      dtmin = merge_candidate(dtmin, refined_dt(sh_c, cg[21]));
    }
    bool bounded = false;
    dtmin = bound_rule(cur, dtmin, &bounded);
    if (bounded && (tile.w & kTileFirst) && threadIdx.x == 0) atomicOr(a.status, FT_STATUS_DT_BOUND);
  }
  const double dt_next = a.factor * dtmin;
  // no stable step left: below the failure step (an inverted element's is
  // negative and counts as 0), or none finite at all (non-finite forces)
  const bool collapse = !(dtmin >= a.failure_dt) || !(dtmin < 1e308);
  const bool go = (t_k < tmax) && (cur.step < cap) && !collapse;
  // this instance's motion states at t_{k-1}, t_k, t_{k+1} (from
  // k_motion_states, or evaluated here with inline_motions) and, when
  // rotations are present, the parameters motion_apply reads (type,
  // axis, centre), in LDS sized for the run's largest motion count
  // (runs without rotations keep the fixed-size table: its constant strides
  // cost no registers)
  extern __shared__ __align__(16) unsigned char ft_nodal_lds[];
  __shared__ MotionT ms_fixed[3][ROT ? 1 : kMaxMotions];
  MotionT *ms_dyn = reinterpret_cast<MotionT *>(ft_nodal_lds);
  double(*mp)[8] = reinterpret_cast<double(*)[8]>(ms_dyn + 3 * a.mcap);
  // motion state w (0: t_{k-1}, 1: t_k, 2: t_{k+1}) of the instance's motion j
  auto ms = [&](int w, int j) -> MotionT & {
    if constexpr (ROT)
      return ms_dyn[w * a.mcap + j];
    else
      return ms_fixed[w][j];
  };
  if (tile.w & kTileMotion) {  // uniform: a tile without constrained nodes reads no motion state
    for (int i = threadIdx.x; i < 3 * mc; i += blockDim.x) {
      if (a.inline_motions)
        motion_time(a.motions + (size_t)FT_MOTION_STRIDE * (mb + i % mc),
                    motion_state_time(cur, a.factor, i / mc, dtmin), ms(i / mc, i % mc));
      else
        ms(i / mc, i % mc) = a.mstate[(size_t)(i / mc) * a.n_motions + mb + i % mc];
    }
    if constexpr (ROT) {
      for (int i = threadIdx.x; i < 7 * mc; i += blockDim.x)
        mp[i / 7][i % 7] = a.motions[(size_t)FT_MOTION_STRIDE * (mb + i / 7) + i % 7];
    }
    __syncthreads();
  }
  const double *motions = a.motions + (size_t)FT_MOTION_STRIDE * mb;
  // Energy (CheckEnergy.cpp:15-52): the terms of step k pair each DOF's
  // displacement increment du_k with fi_{k-1} + fi_k.  The fi_{k-1} halves
  // (pInt, pExt) are added by the launch that knows fi_{k-1} and already
  // computes du_k, so fi is not carried from step to step; it is stored only
  // when the run stops.
  double sWint = 0.0, sWext = 0.0, KE = 0.0, pInt = 0.0, pExt = 0.0;
  if (n < tile.y) {
    double f[3];
    if (a.mode == kResume) {
      f[0] = a.fi[n];
      f[1] = a.fi[(size_t)a.nn + n];
      f[2] = a.fi[2 * (size_t)a.nn + n];
    } else {
      gather3<R, W>(a.ell, a.nn, a.width, a.fe, n, f, a.ov);
      // non-finite element forces (the element kernels of the run loop
      // leave this check to the node sums)
      if (!(isfinite(f[0]) && isfinite(f[1]) && isfinite(f[2]))) atomicOr(a.status, FT_STATUS_NONFINITE);
    }
    const double m = a.mass[n];
    const int code = a.node_code[n];  // byte d: 0 free, else motion index + 1
    const double t_khalf = 0.5 * (t_k + t_km1);
    const double t_k1 = t_k + dt_next;
    const double t_k1half = 0.5 * (t_k1 + t_k);
    double Xn[3] = {0.0, 0.0, 0.0};
    if (code) {  // reference position, read for constrained nodes only
      Xn[0] = a.X[n];
      Xn[1] = a.X[(size_t)a.nn + n];
      Xn[2] = a.X[2 * (size_t)a.nn + n];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int mi = ((code >> (8 * d)) & 0xff) - 1;
      const size_t id = (size_t)d * a.nn + n;
      const double uo = a.u[id];
      const double vso = a.vs[id];
      double acc, v;
      double ub = 0.0, vb = 0.0, ab = 0.0;
      const double *mo = ROT ? mp[mi < 0 ? 0 : mi] : motions + FT_MOTION_STRIDE * (mi < 0 ? 0 : mi);
      if (mi < 0) {
        acc = (0.0 - f[d]) / m;
        v = (a.mode == kStep) ? vso + (t_k - t_khalf) * acc : vso;
      } else {
        motion_apply(mo, ms(1, mi), Xn, d, ub, vb, ab);
        acc = ab;
        v = vb;
      }
      if (a.mode == kStep && a.energy) {
        // this step's halves, with du = dt v_half for free DOFs and the
        // boundary reaction fi + m a at constrained ones (CheckEnergy.cpp:34-46)
        double dd;
        if (mi < 0) {
          dd = dt_k * vso;
        } else {
          double up, vp, ap;
          motion_apply(mo, ms(0, mi), Xn, d, up, vp, ap);
          dd = ub - up;
          sWext += dd * (f[d] + m * acc);
        }
        KE += m * v * v;
        sWint += dd * f[d];
      }
      if (go) {
        if (mi < 0) {
          const double vh = v + (t_k1half - t_k) * acc;
          a.vs[id] = vh;
          a.u[id] = uo + dt_next * vh;
          pInt += (dt_next * vh) * f[d];
        } else {
          double un, vn, an;
          motion_apply(mo, ms(2, mi), Xn, d, un, vn, an);
          a.u[id] = un;
          a.vs[id] = vn;
          const double dd1 = un - ub;
          pInt += dd1 * f[d];
          pExt += dd1 * (f[d] + m * acc);
        }
      } else {
        a.vs[id] = v;
        if (a.mode != kResume) a.fi[id] = f[d];
      }
    }
  }
  if (a.energy && (a.mode == kStep || go)) {  // uniform over the grid
    // accumulators 0-4 of this block (W_int, W_ext, KE, and the next step's
    // halves pInt, pExt), one per lane 0-4:
    //   acc0 += (acc3 + W_int,k) / 2, acc1 += (acc4 + W_ext,k) / 2,
    //   acc2 = KE_k / 2, acc3 = pInt, acc4 = pExt
    __shared__ double sh5[kBlock / 64][5];
    double t5[5] = {sWint, sWext, KE, pInt, pExt};
    // DPP sums for the four-slot gather of corner pairs; the node-slot and
    // ELL gathers keep the butterfly, so that their runs stay equal bit for
    // bit, energies included (1M block, node update with DPP against the
    // butterfly: corner pairs 31.4 against 34.3 us, ELL 66.8 against 67.1
    // us, node slots 40.8 against 37.0 us)
    const double r = block_sum_spread<5, W == -2>(t5, sh5);
    double acc_old = 0.0;
    if (threadIdx.x < 5) acc_old = a.eacc[kEaccStride * (size_t)L + threadIdx.x];
    const double half = __shfl(acc_old, (int)(threadIdx.x + 3) & 63);  // lanes 0, 1: acc3, acc4
    if (threadIdx.x < 5) {
      const int i = threadIdx.x;
      double *acc = a.eacc + kEaccStride * (size_t)L;
      if (a.mode == kStep && i < 2) acc[i] = acc_old + 0.5 * (half + r);
      if (a.mode == kStep && i == 2) acc[2] = 0.5 * r;
      if (go && i >= 3) acc[i] = r;
    }
  }
  if ((tile.w & kTileFirst) && threadIdx.x == 0) {
    StepSlot nx;
    if (go) {
      nx.t = t_k + dt_next;
      nx.t_prev = t_k;
      nx.dt = dt_next;
      nx.dtmin_bits = kInfBits;
      nx.cand_bits = kNoCandidate;
      nx.settled = 0;
      nx.far_bits = ~0u;
      nx.step = cur.step + 1;
      nx.done = 0;
      nx.failed = 0;
    } else {
      nx = cur;
      nx.done = 1;
      if (collapse && cur.step < cap && t_k < tmax) {
        atomicOr(a.status, FT_STATUS_DT_COLLAPSE);
        nx.failed = 1;
      }
    }
    nx.t_done = t_k;
    nx.dt_done = dt_next;
    nx.dt_guide = dtmin;
    nx.fresh = a.mode == kStep;
    ic->slot[a.parity ^ 1] = nx;
    if (go && a.running) atomicAdd(a.running, 1);
  }
}

template <typename R, int W, bool ROT>
// 7 waves per SIMD: the node update is bound by memory latency, and 71
// VGPRs (no spills) keep 7 of the 8 wave slots in flight (5 at the
// compiler's unconstrained 85: 5 % slower on the 1M block).  The rotation
// path needs a few more registers: 6 waves.
__global__ __launch_bounds__(kBlock, ROT ? 6 : 7) void k_nodal(NodalArgs<R> a) {
  nodal_tile<R, W, ROT>(a, xcd_block());
}

// Scalar state of every motion at t_{k-1}, t_k and t_{k+1} = t_k + dt_{k+1}
// for the nodal kernel of step k (one small block; the trigonometry of the
// rotation motions is kept out of the per-node kernel).
// Motion states of every motion at t_{k-1}, t_k and t_{k+1}.  In the FP32
// engine t_{k+1} needs the step's final minimum: each thread settles a
// pending candidate itself (the same value), and the first state of each
// instance's first motion writes it back to the slot and marks the slot
// settled, so the nodal kernel takes it as final.
__global__ void k_motion_states(Ctl *ctl, int parity, double factor, const double *motions,
                                const int *motion_inst, int nm, MotionT *out, const double *cand_geom,
                                int *status) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 3 * nm; i += gridDim.x * blockDim.x) {
    const int w = i / nm, k = i % nm;
    const int inst = motion_inst[k];
    const StepSlot cur = ctl[inst].slot[parity];
    double dtmin = bits_to_double(cur.dtmin_bits);
    if (cand_geom) {
      if (candidate_pending(cur, dtmin))
        dtmin = merge_candidate(
            dtmin, candidate_dt(cand_geom + (size_t)kCandStride * (unsigned)(cur.cand_bits & 0xffffffffu)));
      bool bounded = false;
      dtmin = bound_rule(cur, dtmin, &bounded);
      if (w == 0 && (k == 0 || motion_inst[k - 1] != inst)) {
        ctl[inst].slot[parity].dtmin_bits = dt_key(dtmin);
        ctl[inst].slot[parity].settled = 1;  // read only by the nodal kernel
        if (bounded) atomicOr(status, FT_STATUS_DT_BOUND);
      }
    }
    motion_time(motions + FT_MOTION_STRIDE * k, motion_state_time(cur, factor, w, dtmin), out[i]);
  }
}

// Per-step history row (time, dt, W_int, W_ext, KE, |KE + W_int - W_ext|):
// sums the per-block energy accumulators in block order.
__device__ inline void history_row(const double *eacc, int nblk, Ctl *ctl, int parity, double *rows, int cap) {
  __shared__ double sh[kBlock / 64];
  double w = 0, x = 0, k = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    w += eacc[kEaccStride * b];
    x += eacc[kEaccStride * b + 1];
    k += eacc[kEaccStride * b + 2];
  }
  w = block_sum(w, sh);
  x = block_sum(x, sh);
  k = block_sum(k, sh);
  const StepSlot s = ctl->slot[parity];
  if (threadIdx.x == 0 && s.fresh) {
    const int i = ctl->hist_count;
    if (i < cap) {
      double *row = rows + 6 * (size_t)i;
      row[0] = s.t_done;
      row[1] = s.dt_done;
      row[2] = w;
      row[3] = x;
      row[4] = k;
      row[5] = fabs(k + w - x);
    }
    ctl->hist_count = i + 1;
  }
}

__global__ void k_history(const double *eacc, int nblk, Ctl *ctl, int parity, double *rows, int cap) {
  history_row(eacc, nblk, ctl, parity, rows, cap);
}

// ---------------------------------------------------------------------------
// Persistent run loop: one cooperative launch runs up to `steps` steps of the
// loop (element phase, node phase, history row) for every instance, the
// phases separated by grid-wide barriers.  All blocks are co-resident (the
// launch is cooperative), every block passes the same barriers, and the loop
// ends for all of them together: after a step in which no instance went on.
// ---------------------------------------------------------------------------
constexpr int kStatusBarrier = 1 << 20;  // a barrier wait hit its bound

// Corner pairs: the partner-less elements' extras added to the xi- slots
// their nodes' owners wrote this step (each slot takes at most one extra:
// the order is fixed).
template <typename R>
__device__ inline void pair_extras(int i, const int2 *xadd, const R *xf, R *fe) {
  const int2 t = xadd[i];
  R *dst = fe + t.x;
  const R *src = xf + 3 * (size_t)t.y;
  dst[0] += src[0];
  dst[1] += src[1];
  dst[2] += src[2];
}

template <typename R>
struct PersistArgs {
  ElemArgs<R> ea;        // all element tiles
  NodalArgs<R> na;       // all node tiles, kStep
  const int2 *xadd;      // corner-pair extras (n_xadd entries) added between the phases
  int n_xadd;
  int steps;
  int parity0;
  int n_inst;
  unsigned long long spins;    // bound of a barrier wait, in polls
  unsigned long long *arrive;  // barrier arrivals (zero at launch)
  int *running;                // 3 counters, the first zero at launch
  const double *eacc;          // history row (hist_cap > 0)
  int nblk;
  double *rows;
  int hist_cap;
};

// Grid-wide barrier number `b` (1, 2, ...) of the launch.  Every wave waits
// for its own memory operations to complete (its stores are then in its
// XCD's L2); one thread per block makes the block's writes visible at device
// scope (release), arrives on a 64-bit counter that only grows, polls it
// without cache maintenance until b * gridDim.x blocks have arrived, and
// then acquires once (the block's waves share the CU's L1 and the XCD's L2).
// The wait is bounded: past `max_spins` polls, or once another block has
// given up, the block raises kStatusBarrier and goes on, so no wave waits
// for ever (the run's results are then void and the host reports the error).
__device__ inline void grid_barrier(unsigned long long *arrive, unsigned long long b, int *status,
                                    unsigned long long max_spins) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long target = b * (unsigned long long)gridDim.x;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(arrive, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long spins = 0;
    while (__hip_atomic_load(arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(16);
      if ((++spins & 255) == 0 &&
          (spins > max_spins ||
           (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kStatusBarrier))) {
        atomicOr(status, kStatusBarrier);
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Tile i of this block in a phase of n tiles: each XCD (blocks b with the
// same b mod 8) takes a contiguous eighth of the tiles, as xcd_block does for
// the one-tile-per-block launches.
__device__ inline int persist_tile(int n, int i) {
  const int per_xcd = gridDim.x >> 3;
  const int chunk = (n + 7) >> 3;
  const int x = blockIdx.x & 7;
  const int L = x * chunk + (blockIdx.x >> 3) + i * per_xcd;
  return L < min(n, (x + 1) * chunk) ? L : -1;
}

// Occupancy target of the persistent kernel (waves per SIMD): the FP32
// neo-Hookean body is held to 168 VGPRs (3 waves per SIMD, no spills).
template <typename R, int MAT>
constexpr int persist_waves() {
  return (sizeof(R) == 4 && MAT == 1) ? 3 : 1;
}

// Element tile L with the kernel body of its material (tiles never mix
// materials; the switch is uniform over the block).
template <typename R>
__device__ __forceinline__ void element_tile_any(const ElemArgs<R> &a, int L) {
  const int4 tile = a.tiles[L];
  switch (a.parts[a.einfo[tile.x] & kPidMask].mat) {
    case 0: element_tile<R, true, true, 0>(a, L); break;
    case 1: element_tile<R, true, true, 1>(a, L); break;
    case 2: element_tile<R, true, true, 2>(a, L); break;
    case 3: element_tile<R, true, true, 3>(a, L); break;
    case 4: element_tile<R, true, true, 4>(a, L); break;
    case 5: element_tile<R, true, true, 5>(a, L); break;
    case 6: element_tile<R, true, true, 6>(a, L); break;
    case 7: element_tile<R, true, true, 7>(a, L); break;
    case 8: element_tile<R, true, true, 8>(a, L); break;
    default: element_tile<R, true, true, -1>(a, L); break;  // raises the unknown-material bit
  }
}

// MAT = 1: a mesh of neo-Hookean elements only; MAT = -1: each tile runs its
// material's body (element_tile_any).
template <typename R, int MAT, int W>
__global__ __launch_bounds__(kBlock, (persist_waves<R, MAT>())) void k_persistent(PersistArgs<R> p) {
  ElemArgs<R> ea = p.ea;
  NodalArgs<R> na = p.na;
  int parity = p.parity0;
  unsigned long long b = 0;
  for (int k = 0; k < p.steps; ++k) {
    ea.parity = parity;
    // a block's consecutive tiles reuse the same LDS: each starts after a
    // block barrier
    for (int i = 0;; ++i) {
      const int L = persist_tile(ea.n_tiles, i);
      if (L < 0) break;
      __syncthreads();
      if constexpr (MAT >= 0)
        element_tile<R, true, true, MAT>(ea, L);
      else
        element_tile_any<R>(ea, L);
    }
    grid_barrier(p.arrive, ++b, na.status, p.spins);
    if (p.n_xadd > 0) {  // uniform
      for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < p.n_xadd; i += gridDim.x * blockDim.x)
        pair_extras(i, p.xadd, ea.xf, ea.fe);
      grid_barrier(p.arrive, ++b, na.status, p.spins);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) p.running[(k + 1) % 3] = 0;
    na.parity = parity;
    na.running = p.running + k % 3;
    for (int i = 0;; ++i) {
      const int L = persist_tile(na.n_tiles, i);
      if (L < 0) break;
      __syncthreads();
      nodal_tile<R, W>(na, L);
    }
    grid_barrier(p.arrive, ++b, na.status, p.spins);
    parity ^= 1;
    // the history row of this step (before any block reaches the next node
    // phase, which rewrites the accumulators)
    if (p.hist_cap > 0 && blockIdx.x == 0) history_row(p.eacc, p.nblk, na.ctl, parity, p.rows, p.hist_cap);
    const int going = __hip_atomic_load(p.running + k % 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int abort = __hip_atomic_load(na.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kStatusBarrier;
    if (going == 0 || abort) {
      // every instance has stopped: both ping-pong slots of each get its
      // final state (as the no-op steps of the launch-per-step loop leave
      // them), so the caller's parity bookkeeping holds for any `steps`
      if (going == 0 && blockIdx.x == 0) {
        for (int inst = threadIdx.x; inst < p.n_inst; inst += blockDim.x) {
          StepSlot nx = na.ctl[inst].slot[parity];
          nx.fresh = 0;
          na.ctl[inst].slot[parity ^ 1] = nx;
        }
      }
      break;
    }
  }
}

// Element-averaged Green-Lagrange strain (CalculateStrain.cpp:77-97:
// Eavg = (1/8) sum_gp (F^T F - I) / 2, column-major) and the principal-strain
// summary of CalculateMaximumPrincipalStrain (CalculateStrain.cpp:8-75: max
// principal clamped at >= 0, min principal clamped at <= 0, max shear), FP64.
// One element's averaged Green-Lagrange strain E (row-major 3x3) and
// principal summary P = (max principal >= 0, min principal <= 0, max shear)
// in FP64, from H = grad_X u at the 8 Gauss points: E = (1/8) sum (H + H^T +
// H^T H) / 2 equals the reference's (1/8) sum (F^T F - I) / 2 without the
// cancellation.
__device__ inline void element_strain64(int nn, const int *nd, const double *X, const double *u, double E[3][3],
                                        double P[3]) {
  double cx[7][3], cu[7][3];
  for (int d = 0; d < 3; ++d) {
    double xv[8], uv[8];
    for (int p = 0; p < 8; ++p) {
      const int n = nd[p2n(p)];
      xv[p] = X[(size_t)d * nn + n] - X[(size_t)d * nn + nd[0]];
      uv[p] = u[(size_t)d * nn + n] - u[(size_t)d * nn + nd[0]];
    }
    wht8(xv);
    wht8(uv);
    for (int m = 0; m < 7; ++m) {
      cx[m][d] = xv[m + 1] * 0.125;
      cu[m][d] = uv[m + 1] * 0.125;
    }
  }
  const double g = kGp, g2 = kGp * kGp;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) E[i][j] = 0.0;
  for (int gp = 0; gp < 8; ++gp) {
    const double sx = ((gp & 3) == 1 || (gp & 3) == 2) ? 1.0 : -1.0;
    const double sy = (gp & 2) ? 1.0 : -1.0;
    const double sz = gp < 4 ? 1.0 : -1.0;
    double J0[3][3], Ju[3][3];
    for (int d = 0; d < 3; ++d) {
      J0[d][0] = cx[0][d] + (sy * g) * cx[2][d] + (sz * g) * cx[4][d] + (sy * sz * g2) * cx[6][d];
      J0[d][1] = cx[1][d] + (sx * g) * cx[2][d] + (sz * g) * cx[5][d] + (sx * sz * g2) * cx[6][d];
      J0[d][2] = cx[3][d] + (sx * g) * cx[4][d] + (sy * g) * cx[5][d] + (sx * sy * g2) * cx[6][d];
      Ju[d][0] = cu[0][d] + (sy * g) * cu[2][d] + (sz * g) * cu[4][d] + (sy * sz * g2) * cu[6][d];
      Ju[d][1] = cu[1][d] + (sx * g) * cu[2][d] + (sz * g) * cu[5][d] + (sx * sz * g2) * cu[6][d];
      Ju[d][2] = cu[3][d] + (sx * g) * cu[4][d] + (sy * g) * cu[5][d] + (sx * sy * g2) * cu[6][d];
    }
    double A[3][3];
    A[0][0] = J0[1][1] * J0[2][2] - J0[1][2] * J0[2][1];
    A[0][1] = J0[0][2] * J0[2][1] - J0[0][1] * J0[2][2];
    A[0][2] = J0[0][1] * J0[1][2] - J0[0][2] * J0[1][1];
    A[1][0] = J0[1][2] * J0[2][0] - J0[1][0] * J0[2][2];
    A[1][1] = J0[0][0] * J0[2][2] - J0[0][2] * J0[2][0];
    A[1][2] = J0[0][2] * J0[1][0] - J0[0][0] * J0[1][2];
    A[2][0] = J0[1][0] * J0[2][1] - J0[1][1] * J0[2][0];
    A[2][1] = J0[0][1] * J0[2][0] - J0[0][0] * J0[2][1];
    A[2][2] = J0[0][0] * J0[1][1] - J0[0][1] * J0[1][0];
    const double det0 = J0[0][0] * A[0][0] + J0[0][1] * A[1][0] + J0[0][2] * A[2][0];
    double H[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) H[i][j] = (Ju[i][0] * A[0][j] + Ju[i][1] * A[1][j] + Ju[i][2] * A[2][j]) / det0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        E[i][j] += 0.0625 * (H[i][j] + H[j][i] + H[0][i] * H[0][j] + H[1][i] * H[1][j] + H[2][i] * H[2][j]);
  }
  // principal values: trigonometric solution of the characteristic cubic
  const double a = E[0][0], b = E[1][0], c = E[2][0], d = E[1][1], ee = E[2][1], f = E[2][2];
  const double p1 = b * b + c * c + ee * ee;
  double e1, e2, e3;
  if (p1 == 0) {
    e1 = a;
    e2 = d;
    e3 = f;
  } else {
    const double pi = 4.0 * atan(1.0);
    double I1 = a + d + f;
    const double I2 = a * (d + f) + d * f - b * b - c * c - ee * ee;
    const double I3 = a * d * f + 2.0 * b * c * ee - b * b * f - c * c * d - ee * ee * a;
    const double Q = (3.0 * I2 - I1 * I1) / 9.0;
    const double R = (2.0 * I1 * I1 * I1 - 9.0 * I1 * I2 + 27.0 * I3) / 54.0;
    // The reference takes acos of the raw ratio; round-off near a repeated
    // eigenvalue can push it outside [-1, 1] (or make Q >= 0) and the
    // reference then returns NaN.  Clamping gives the limit value instead.
    const double den = sqrt(fmax(-Q * Q * Q, 0.0));
    const double ratio = den > 0.0 ? fmin(fmax(R / den, -1.0), 1.0) : 1.0;
    const double theta = acos(ratio);
    const double sq = 2.0 * sqrt(fmax(-Q, 0.0));
    I1 = I1 / 3.0;
    e1 = sq * cos(theta / 3.0) + I1;
    e2 = sq * cos((theta + 2.0 * pi) / 3.0) + I1;
    e3 = sq * cos((theta + 4.0 * pi) / 3.0) + I1;
  }
  const double mx = fmax(e3, fmax(e2, e1)), mn = fmin(e3, fmin(e2, e1));
  P[0] = mx > 0.0 ? mx : 0.0;
  P[1] = mn < 0.0 ? mn : 0.0;
  P[2] = 0.5 * (mx - mn);
}

// Element-averaged Green-Lagrange strain (CalculateStrain.cpp:77-97,
// column-major) and the principal-strain summary of
// CalculateMaximumPrincipalStrain (CalculateStrain.cpp:8-75), FP64.
__global__ __launch_bounds__(kBlock) void k_strain(int nn, int ne, const int *conn, const double *X,
                                                   const double *u, double *eavg, double *princ) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int nd[8];
  for (int k = 0; k < 8; ++k) nd[k] = conn[8 * (size_t)e + k];
  double E[3][3], P[3];
  element_strain64(nn, nd, X, u, E, P);
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) eavg[9 * (size_t)e + i + 3 * j] = E[i][j];
  for (int k = 0; k < 3; ++k) princ[3 * (size_t)e + k] = P[k];
}

// ---------------------------------------------------------------------------
// Injury criteria of FemTech's ex5 (CalculateInjuryCriterions, ex5.cpp:
// 1311-1430), evaluated after every step at the step's displacement.
// ---------------------------------------------------------------------------
// Per element flags (internal element order).
constexpr int kInjIncluded = 1, kInjGt15 = 2, kInjGt30 = 4, kInjPsr120 = 8, kInjPsxsr28 = 16;
constexpr int kInjMps95 = 32, kInjMpsxsr95 = 64;
// Running scalars: (value, original element id, time) for max principal,
// min principal, max shear and max PS x SR; (value, time) for MPS-95 and
// MPSxSR-95; then two "list needs refresh" flags.
constexpr int kInjScal = 20;
struct InjuryArgs {
  int nn, ne;
  const int *conn;
  const double *X, *u;
  const int *perm;   // internal -> original element id
  int *flags;
  double *ps_old, *psxsr;
  double *part;      // 8 doubles per block
  double *scal;      // kInjScal
  const StepSlot *slot;  // time / dt source (nullable: use time, dt)
  double time, dt;
};

struct ArgBest {
  double v;
  int id;
};
__device__ inline ArgBest better_max(ArgBest a, ArgBest b) {
  return (b.v > a.v || (b.v == a.v && b.id < a.id)) ? b : a;
}

__device__ inline ArgBest block_argbest(ArgBest x, bool is_max) {
  __shared__ double sv[kBlock];
  __shared__ int si[kBlock];
  if (!is_max) x.v = -x.v;
  sv[threadIdx.x] = x.v;
  si[threadIdx.x] = x.id;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      ArgBest o = better_max({sv[threadIdx.x], si[threadIdx.x]}, {sv[threadIdx.x + w], si[threadIdx.x + w]});
      sv[threadIdx.x] = o.v;
      si[threadIdx.x] = o.id;
    }
    __syncthreads();
  }
  ArgBest r{sv[0], si[0]};
  __syncthreads();
  if (!is_max) r.v = -r.v;
  return r;
}

__global__ __launch_bounds__(kBlock) void k_injury_elem(InjuryArgs a) {
  if (a.slot && a.slot->done) return;  // uniform
  const double dt = a.slot ? a.slot->dt : a.dt;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const double ninf = -INFINITY;
  ArgBest bmax{ninf, INT32_MAX}, bmin{INFINITY, INT32_MAX}, bsh{ninf, INT32_MAX}, bx{ninf, INT32_MAX};
  if (e < a.ne && (a.flags[e] & kInjIncluded)) {
    int nd[8];
    for (int k = 0; k < 8; ++k) nd[k] = a.conn[8 * (size_t)e + k];
    double E[3][3], P[3];
    element_strain64(a.nn, nd, a.X, a.u, E, P);
    const int id = a.perm[e];
    int fl = a.flags[e];
    if (P[0] > 0.15) fl |= kInjGt15;
    if (P[0] > 0.30) fl |= kInjGt30;
    const double psr = (P[0] - a.ps_old[e]) / dt;
    const double psxsr = P[0] * psr;
    if (psr > 120.0) fl |= kInjPsr120;
    if (psxsr > 28.0) fl |= kInjPsxsr28;
    a.flags[e] = fl;
    a.ps_old[e] = P[0];
    a.psxsr[e] = psxsr;
    bmax = {P[0], id};
    bmin = {P[1], id};
    bsh = {P[2], id};
    bx = {psxsr, id};
  }
  bmax = block_argbest(bmax, true);
  bmin = block_argbest(bmin, false);
  bsh = block_argbest(bsh, true);
  bx = block_argbest(bx, true);
  if (threadIdx.x == 0) {
    double *p = a.part + 8 * (size_t)blockIdx.x;
    p[0] = bmax.v, p[1] = bmax.id, p[2] = bmin.v, p[3] = bmin.id;
    p[4] = bsh.v, p[5] = bsh.id, p[6] = bx.v, p[7] = bx.id;
  }
}

// The step's extrema over all blocks (block order, smallest id on ties), then
// the running maxima with strict comparisons (ex5.cpp:1318-1366).
__global__ __launch_bounds__(kBlock) void k_injury_reduce(InjuryArgs a, int nblk) {
  if (a.slot && a.slot->done) return;
  const double ninf = -INFINITY;
  ArgBest r[4] = {{ninf, INT32_MAX}, {ninf, INT32_MAX}, {ninf, INT32_MAX}, {ninf, INT32_MAX}};
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    const double *p = a.part + 8 * (size_t)b;
    for (int k = 0; k < 4; ++k) {
      const double v = k == 1 ? -p[2 * k] : p[2 * k];  // the minimum as a maximum of -v
      r[k] = better_max(r[k], {v, (int)p[2 * k + 1]});
    }
  }
  for (int k = 0; k < 4; ++k) r[k] = block_argbest(r[k], true);
  if (threadIdx.x == 0) {
    const double t = a.slot ? a.slot->t : a.time;
    double *s = a.scal;
    if (r[0].id != INT32_MAX && s[0] < r[0].v) s[0] = r[0].v, s[1] = r[0].id, s[2] = t;
    if (r[1].id != INT32_MAX && s[3] > -r[1].v) s[3] = -r[1].v, s[4] = r[1].id, s[5] = t;
    if (r[2].id != INT32_MAX && s[6] < r[2].v) s[6] = r[2].v, s[7] = r[2].id, s[8] = t;
    if (r[3].id != INT32_MAX && s[9] < r[3].v) s[9] = r[3].v, s[10] = r[3].id, s[11] = t;
  }
}

// 95th percentile (math.cpp:160-196 on one rank: the element at index
// (int)(0.95 n) - 1 of the ascending order) of ps_old and of psxsr by a
// radix select over order-preserving 64-bit keys, 8 bits per pass, both
// arrays in the same launches.  sel[m]: {prefix, remaining rank}.
__device__ inline unsigned long long order_key(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double key_value(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  return __longlong_as_double((long long)b);
}

__global__ void k_select_init(unsigned long long *sel, unsigned *hist, long n_incl, const StepSlot *slot) {
  if (slot && slot->done) return;
  long k = (long)(n_incl * 0.95) - 1;
  if (k < 0) k = 0;
  if (threadIdx.x < 2) {
    sel[2 * threadIdx.x] = 0;
    sel[2 * threadIdx.x + 1] = (unsigned long long)k;
  }
  for (int i = threadIdx.x; i < 512; i += blockDim.x) hist[i] = 0;
}

__global__ __launch_bounds__(kBlock) void k_select_hist(int ne, const int *flags, const double *v0,
                                                        const double *v1, const unsigned long long *sel,
                                                        int pass, unsigned *hist, const StepSlot *slot) {
  if (slot && slot->done) return;
  __shared__ unsigned h[2][256];
  for (int i = threadIdx.x; i < 512; i += blockDim.x) h[i >> 8][i & 255] = 0;
  __syncthreads();
  const int shift = 56 - 8 * pass;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
    if (!(flags[e] & kInjIncluded)) continue;
    const unsigned long long k0 = order_key(v0[e]), k1 = order_key(v1[e]);
    if (pass == 0 || (k0 >> (shift + 8)) == sel[0]) atomicAdd(&h[0][(k0 >> shift) & 255], 1u);
    if (pass == 0 || (k1 >> (shift + 8)) == sel[2]) atomicAdd(&h[1][(k1 >> shift) & 255], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 512; i += blockDim.x)
    if (h[i >> 8][i & 255]) atomicAdd(&hist[i], h[i >> 8][i & 255]);
}

// Picks the digit holding the remaining rank, extends the prefix, clears the
// histogram for the next pass.
__global__ void k_select_pick(unsigned long long *sel, unsigned *hist, const StepSlot *slot) {
  if (slot && slot->done) return;
  if (threadIdx.x < 2) {
    const int m = threadIdx.x;
    unsigned long long k = sel[2 * m + 1], cum = 0;
    int b = 0;
    for (; b < 255; ++b) {
      if (k < cum + hist[256 * m + b]) break;
      cum += hist[256 * m + b];
    }
    sel[2 * m] = (sel[2 * m] << 8) | (unsigned long long)b;
    sel[2 * m + 1] = k - cum;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 512; i += blockDim.x) hist[i] = 0;
}

// MPS-95 / MPSxSR-95 running maxima (ex5.cpp:1370-1429).
__global__ void k_injury_p95(const unsigned long long *sel, double *scal, const StepSlot *slot, double time) {
  if (slot && slot->done) return;
  if (threadIdx.x < 2) {
    const int m = threadIdx.x;
    const double v = key_value(sel[2 * m]);
    const double t = slot ? slot->t : time;
    double *s = scal + 12 + 2 * m;
    scal[16 + m] = 0.0;
    if (v > s[0]) {
      s[0] = v;
      s[1] = t;
      scal[16 + m] = 1.0;
    }
  }
}

// Element lists of a new MPS-95 / MPSxSR-95 maximum: value >= the maximum.
__global__ __launch_bounds__(kBlock) void k_injury_members(int ne, int *flags, const double *ps_old,
                                                           const double *psxsr, const double *scal,
                                                           const StepSlot *slot) {
  if (slot && slot->done) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  int fl = flags[e];
  if (scal[16] != 0.0) fl = (fl & ~kInjMps95) | ((fl & kInjIncluded) && ps_old[e] >= scal[12] ? kInjMps95 : 0);
  if (scal[17] != 0.0)
    fl = (fl & ~kInjMpsxsr95) | ((fl & kInjIncluded) && psxsr[e] >= scal[14] ? kInjMpsxsr95 : 0);
  flags[e] = fl;
}

// Reference-configuration element volumes (calculateVolume,
// CalculateCentroidAndVolume.cpp:26-37) for the CSDM fractions.
__global__ __launch_bounds__(kBlock) void k_element_volume(int nn, int ne, const int *conn, const double *X,
                                                           double *vol) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double c[7][3];
  for (int d = 0; d < 3; ++d) {
    double v[8];
    const int n0 = conn[8 * (size_t)e];
    for (int p = 0; p < 8; ++p) v[p] = X[(size_t)d * nn + conn[8 * (size_t)e + p2n(p)]] - X[(size_t)d * nn + n0];
    wht8(v);
    for (int m = 0; m < 7; ++m) c[m][d] = v[m + 1] * 0.125;
  }
  vol[e] = hex_part(c, 6);
}

// Volume sums for CSDM-15, CSDM-30, MPSR-120, MPSxSR-28 (ex5.cpp:1049-1066):
// per-block partials [included, gt15, gt15 and gt30, psr120, psxsr28].
__global__ __launch_bounds__(kBlock) void k_injury_volumes(int ne, const int *flags, const double *vol,
                                                           double *part) {
  __shared__ double sh[kBlock / 64];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double s[5] = {0, 0, 0, 0, 0};
  if (e < ne && (flags[e] & kInjIncluded)) {
    const int fl = flags[e];
    const double v = vol[e];
    s[0] = v;
    if (fl & kInjGt15) {
      s[1] = v;
      if (fl & kInjGt30) s[2] = v;
    }
    if (fl & kInjPsr120) s[3] = v;
    if (fl & kInjPsxsr28) s[4] = v;
  }
  for (int k = 0; k < 5; ++k) {
    const double r = block_sum(s[k], sh);
    if (threadIdx.x == 0) part[5 * (size_t)blockIdx.x + k] = r;
  }
}

// CalculateAccelerations (CalculateAcclerations.cpp:4-12) on node-major host
// arrays: a = f_net / m on DOFs whose boundary flag is 0.
__global__ __launch_bounds__(kBlock) void k_accel(int ndof, const double *f_net, const int *boundary,
                                                  const double *mass, double *a) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ndof; i += gridDim.x * blockDim.x)
    if (!boundary[i]) a[i] = f_net[i] / mass[i / 3];
}

// CheckEnergy's per-step sums (CheckEnergy.cpp:15-52, one rank) on node-major
// arrays: [0] sum dd (fi_prev + fi), [1] sum dd (fe_prev + fe) plus the
// boundary reactions dd (fi_prev + fi + m (a + a_prev)), [2] sum m v^2.
// Per-block partials in a fixed order; k_sum3 adds them in a fixed order.
constexpr int kEnergyBlocks = 512;
struct EnergyIn {
  const double *u, *u_prev, *v, *a, *a_prev, *fi, *fi_prev, *fe, *fe_prev, *mass;
  const int *boundary;
};
__global__ __launch_bounds__(kBlock) void k_energy_terms(int ndof, EnergyIn in, double *part) {
  double s[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ndof; i += gridDim.x * blockDim.x) {
    const double m = in.mass[i / 3], dd = in.u[i] - in.u_prev[i], fsum = in.fi_prev[i] + in.fi[i];
    s[2] += m * in.v[i] * in.v[i];
    if (in.boundary[i]) s[1] += dd * (fsum + m * (in.a[i] + in.a_prev[i]));
    s[0] += dd * fsum;
    s[1] += dd * (in.fe_prev[i] + in.fe[i]);
  }
  __shared__ double sh[3][kBlock];
  for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}
// Sums of nv-tuples over nblk blocks, in block order (one thread per value).
__global__ void k_sum_tuples(int nblk, int nv, const double *part, double *out) {
  if ((int)threadIdx.x < nv) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(size_t)nv * b + threadIdx.x];
    out[threadIdx.x] = s;
  }
}
__global__ void k_sum3(int nblk, const double *part, double *out) {
  if (threadIdx.x < 3) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[3 * b + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// Scratch of ft_energy_terms_noalloc: the kEnergyBlocks partial triples of
// k_energy_terms and the three totals of k_sum3 (12.3 KB), one copy per GPU,
// held until the process ends.  Host code never passes `scratch` itself to a
// kernel (in host code the name is only a stand-in for the symbol): the
// device address of each GPU's copy comes from hipGetSymbolAddress, taken
// once per device at engine creation (scratch_address) and kept per device.
constexpr size_t kScratchDoubles = 3 * (size_t)kEnergyBlocks + 3;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {nullptr};

// Device address of `scratch` on `device` (which must be current).
int scratch_address(int device, double **out) {
  if (device < 0 || device >= kMaxDevices) return fail(FT_ERR_ARG, "device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FT_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = (double *)p;
  }
  *out = g_scratch_ptr[device];
  return 0;
}

// Plain streaming copy (calibration of the HBM traffic counters).
__global__ __launch_bounds__(kBlock) void k_copy_probe(const float4 *src, float4 *dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Lumped mass: per-element row sums (Mass3D.cpp:5-67,134-144) ...
__global__ __launch_bounds__(kBlock) void k_mass_elem(int nn, int ne, const int *conn, const int *einfo,
                                                      const double *X, const PartDev *parts,
                                                      double *me) {
  const int e = xcd_block() * kBlock + threadIdx.x;
  if (e >= ne) return;
  int nd[8];
  for (int k = 0; k < 8; ++k) nd[k] = conn[8 * e + k];
  const double rho = parts[einfo[e] & kPidMask].rho;
  double xv[3][8];
  for (int p = 0; p < 8; ++p)
    for (int d = 0; d < 3; ++d) xv[d][p] = X[(size_t)d * nn + nd[p2n(p)]] - X[(size_t)d * nn + nd[0]];
  for (int d = 0; d < 3; ++d) wht8(xv[d]);
  double c[7][3];
  for (int m = 0; m < 7; ++m)
    for (int d = 0; d < 3; ++d) c[m][d] = xv[d][m + 1] * 0.125;
  double m8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double g = kGp, g2 = kGp * kGp;
  for (int gp = 0; gp < 8; ++gp) {
    const double sx = (gp == 1 || gp == 2 || gp == 5 || gp == 6) ? 1.0 : -1.0;
    const double sy = (gp == 2 || gp == 3 || gp == 6 || gp == 7) ? 1.0 : -1.0;
    const double sz = gp < 4 ? 1.0 : -1.0;
    double J[3][3];
    for (int d = 0; d < 3; ++d) {
      J[d][0] = c[0][d] + (sy * g) * c[2][d] + (sz * g) * c[4][d] + (sy * sz * g2) * c[6][d];
      J[d][1] = c[1][d] + (sx * g) * c[2][d] + (sz * g) * c[5][d] + (sx * sz * g2) * c[6][d];
      J[d][2] = c[3][d] + (sx * g) * c[4][d] + (sy * g) * c[5][d] + (sx * sy * g2) * c[6][d];
    }
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    // row sum of N_a N_b over b is N_a (partition of unity)
    for (int a = 0; a < 8; ++a) {
      const double sa = (a == 1 || a == 2 || a == 5 || a == 6) ? 1.0 : -1.0;
      const double ta = (a == 2 || a == 3 || a == 6 || a == 7) ? 1.0 : -1.0;
      const double ua = a >= 4 ? 1.0 : -1.0;
      const double Na = (1 + sa * sx * g) * (1 + ta * sy * g) * (1 + ua * sz * g) / 8;
      m8[a] += Na * det;
    }
  }
  for (int a = 0; a < 8; ++a) me[(size_t)a * ne + e] = rho * m8[a];  // slot-major
}

// ... gathered per node in element order.
__global__ __launch_bounds__(kBlock) void k_mass_gather(int nn, const int *ell, int width,
                                                        const double *me, double *mass) {
  const int n = xcd_block() * kBlock + threadIdx.x;
  if (n >= nn) return;
  double s = 0.0;
  for (int j = 0; j < width; ++j) {
    const int id = ell[(size_t)j * nn + n];
    if (id < 0) break;
    s += me[id];
  }
  mass[n] = s;
}

// Per-element skip flag of StableTimeStep.cpp:12-20 from per-node masks.
__global__ void k_elem_flags(int ne, const int *conn, const unsigned char *node_mask, int *einfo) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  bool all = true;
  for (int k = 0; k < 8; ++k) all = all && (node_mask[conn[8 * e + k]] == 7);
  int v = einfo[e] & ~kSkipBit;
  einfo[e] = all ? (v | kSkipBit) : v;
}

// Velocity / acceleration read-out for ft_get_state.
__global__ void k_state_out(int nn, const int4 *tiles, int n_tiles, const Ctl *ctl, int parity,
                            const double *fi, const double *mass, const double *X, const int *node_code,
                            const double *motions, const int *mrange, double *a_out) {
  const int L = blockIdx.x;
  if (L >= n_tiles) return;
  const int4 tile = tiles[L];
  const int n = tile.x + threadIdx.x;
  if (n >= tile.y) return;
  const double t = ctl[tile.z].slot[parity].t;
  const double *mo = motions + (size_t)FT_MOTION_STRIDE * mrange[2 * tile.z];
  for (int d = 0; d < 3; ++d) {
    const int mi = ((node_code[n] >> (8 * d)) & 0xff) - 1;
    const size_t id = (size_t)d * nn + n;
    if (mi < 0) {
      a_out[id] = (0.0 - fi[id]) / mass[n];
    } else {
      double ub[3], vb[3], ab[3];
      const double Xn[3] = {X[n], X[(size_t)nn + n], X[2 * (size_t)nn + n]};
      motion_eval(mo + FT_MOTION_STRIDE * mi, Xn, t, ub, vb, ab);
      a_out[id] = ab[d];
    }
  }
}

// Per (column, chunk): sum, sum of squares, min, max of column `blockIdx.x`
// of x[n][w] over rows [chunk len, +len) -> part[(chunk w + col) 4 + k].
__global__ __launch_bounds__(kBlock) void k_col_stats(const float *x, long n, int w, long len, double *part) {
  __shared__ double sh[kBlock / 64];
  const int col = blockIdx.x;
  const long r0 = (long)blockIdx.y * len, r1 = r0 + len < n ? r0 + len : n;
  double s = 0, q = 0, lo = 1e300, hi = -1e300;
  for (long r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    const double v = x[r * w + col];
    s += v;
    q += v * v;
    lo = fmin(lo, v);
    hi = fmax(hi, v);
  }
  s = block_sum(s, sh);
  q = block_sum(q, sh);
  __syncthreads();
  lo = block_min(lo, sh);
  __syncthreads();
  hi = -block_min(-hi, sh);
  if (threadIdx.x == 0) {
    double *o = part + ((size_t)blockIdx.y * w + col) * 4;
    o[0] = s;
    o[1] = q;
    o[2] = lo;
    o[3] = hi;
  }
}

// Network targets of the samples: S_iso / s (s = raw feature 6; 0 where
// s vanishes), in place.
__global__ void k_surr_per_strain(const float *xs, float *ts, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float sn = xs[i * kSurrFeat + 6];
  const float is = sn > (float)kSurrSmin ? 1.0f / sn : 0.0f;
#pragma unroll
  for (int k = 0; k < kSurrOut; ++k) ts[i * kSurrOut + k] *= is;
}

// Normalised training arrays: x -> BF16 (x - mean) / scale, t -> (t - mean) / scale.
__global__ void k_surr_normalise(const float *xs, const float *ts, long n, const float *norm, __bf16 *xb, float *tb) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int k = 0; k < kSurrFeat; ++k) xb[i * kSurrFeat + k] = (__bf16)((xs[i * kSurrFeat + k] - norm[k]) / norm[16 + k]);
#pragma unroll
  for (int k = 0; k < kSurrOut; ++k) tb[i * kSurrOut + k] = (ts[i * kSurrOut + k] - norm[64 + k]) / norm[72 + k];
}

// Every stride-th sample of (xb, tb) into (vx, vt), and its strain
// magnitude (raw feature 6 of xs) into vs; m samples.
__global__ void k_gather_stride(const __bf16 *xb, const float *tb, const float *xs, long stride, long m, __bf16 *vx,
                                float *vt, float *vs) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  vs[i] = xs[i * stride * kSurrFeat + 6];
#pragma unroll
  for (int k = 0; k < kSurrFeat; ++k) vx[i * kSurrFeat + k] = xb[i * stride * kSurrFeat + k];
#pragma unroll
  for (int k = 0; k < kSurrOut; ++k) vt[i * kSurrOut + k] = tb[i * stride * kSurrOut + k];
}

}  // namespace

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
struct ft_engine {
  int nn = 0, ne = 0, nparts = 0, precision = FT_FP64;
  int device = 0;
  hipStream_t stream = nullptr;
  // mesh
  double *X = nullptr;
  int *conn = nullptr, *einfo = nullptr, *ell = nullptr;
  int width = 8;  // ELL width: max node valence, at least 8
  // Node-major force layouts: element forces go straight to
  // fe[slot][node][xyz] and the node sum reads 8 contiguous slots with no
  // incidence list.  Corner slots (slot = local corner k) when no two
  // elements share a node at the same local corner (every structured
  // block); node slots (slot = the element's rank among the node's elements
  // in element order, eslot) when every node has at most 8 elements.
  // Otherwise fe[slot][element][xyz] + the ELL.
  bool corner = false;
  unsigned *eslot = nullptr;        // node slots: per internal element, 4 bits per local node (rank)
  int *ovf = nullptr;               // node slots: per node (overflow index << 4) | ranks beyond 8
  void *fe_ovf = nullptr;           // node slots: forces of ranks 8-15
  int n_ovf = 0;                    // nodes with more than 8 elements
  std::vector<unsigned> h_eslot;    // the same in the caller's element order
  std::vector<int> h_ovf;           // host copy of ovf
  int layout = 0;                   // FT_LAYOUT_* in use
  // Corner pairs (layout FT_LAYOUT_CORNER_PAIRS, set by build_pairs): when
  // element i - 1 (same segment and instance) shares its face xi = +1 with
  // element i, i adds i - 1's four corner forces there to its own and writes
  // the sums, i - 1 writes only its other four (each wavefront's lane 0
  // recomputes the element before the wavefront's own, so pairs meet in
  // registers; segments of the Prony laws have no such shadow lanes and
  // pair only within a wavefront).  An element without a right partner whose xi+ corners'
  // matching xi- slots no element owns (a row end of a block) writes those
  // corners there.  Complete pairing (every element one or the other): the
  // node sum reads the four xi- slots only; otherwise all eight, the xi+
  // slots the loop does not write holding zeros.  pzero, [2][nn]: per node
  // the slots to zero before a pass outside the run loop (which writes every
  // element's own eight corners; the free slots the loop writes) and after
  // it (the xi+ slots the loop leaves unwritten).
  bool paired = false;
  bool pair_complete = false;  // every xi+ corner force reaches a xi- slot: the node sum reads 4 slots
  bool unpaired = false;       // set while a pass outside the run loop writes every corner
  unsigned char *pzero = nullptr;  // per node: slots to zero around passes outside the run loop
  long n_pairs = 0;
  // extras of the partner-less elements (kExtra): values [4 n][3], per
  // element its first entry, and the (fe offset, entry) pairs k_pair_extras adds
  void *xf = nullptr;
  int *xidx = nullptr;
  int2 *xadd = nullptr;
  int n_xadd = 0;
  std::vector<int> h_conn;  // caller's connectivity (element-force read-back)
  PartDev *parts = nullptr;
  std::vector<PartDev> h_parts;
  bool has_visco = false;
  // Elements are stored grouped by material (stable within a group) so each
  // material runs its own specialised element kernel over a contiguous range.
  std::vector<int> perm;  // internal -> original element index
  std::vector<int> inv;   // original -> internal
  // waves: occupancy variant of the FP32 run-loop kernel (0: default);
  // tile: elements per tile = threads per block.  Both set by ft_autotune.
  struct Seg { int mat, begin, end, tile_begin = 0, tile_end = 0, waves = 0, tile = kBlock; };
  std::vector<Seg> segs;
  // Independent instances (batched simulations): node range and element
  // range of each in the caller's numbering; a single mesh is 1 instance.
  int n_inst = 1;
  std::vector<int> inst_node, inst_elem;   // n_inst + 1 offsets each
  int4 *etiles = nullptr, *ntiles = nullptr;
  std::vector<int4> h_etiles;   // host copy (launch order)
  std::vector<int4> h_ntiles;   // host copy
  std::vector<int> h_code;      // host copy of node_code (empty until a boundary is set)
  double *cand_geom = nullptr;  // kCandStride per element tile (FP32 exact step size)
  double *gp_fields = nullptr;  // per-Gauss-point F, det F, PK2 (128 per element), on demand
  int *perm_dev = nullptr;      // internal -> caller element index, on demand
  int n_etiles = 0, n_ntiles = 0;
  std::vector<int> ntile_of_inst;          // n_inst + 1: node tiles per instance
  int *mrange = nullptr;                   // per instance: first motion, count
  int *motion_inst = nullptr;              // per (expanded) motion: instance
  int *inst_cap = nullptr;                 // per instance: step cap of the current call
  double *inst_tmax = nullptr;             // per instance: end time
  std::vector<double> h_tmax;
  bool tmax_per_instance = false;          // end times from ft_set_instance_end_times
  // state
  double *u = nullptr, *vs = nullptr, *fi = nullptr, *mass = nullptr, *tmp3 = nullptr;
  void *fe = nullptr;
  void *hist = nullptr;
  void *fib = nullptr;  // fibre directions, element type, [3][ne] internal order
  int *node_code = nullptr;
  unsigned char *node_mask = nullptr;
  double *motions = nullptr;
  int n_motions = 0;
  bool timed_motions = false;  // any motion other than FIXED
  bool rot_motions = false;    // any rotation motion
  int mcap = 1;                // largest motion count of any instance (at least 1)
  MotionT *mstate = nullptr;
  std::vector<int> h_dof_motion;
  std::vector<double> h_motions;
  Ctl *ctl = nullptr;
  unsigned long long *dtbits = nullptr;
  int *status = nullptr;
  double *eacc = nullptr;
  int nodal_grid = 0;
  bool mass_ready = false;
  // run
  bool run_active = false;
  int parity = 0;
  double tmax = 0, factor = 0.8, failure_dt = 1e-11;
  // history
  int hist_cap = 0, hist_n = 0;
  double *hist_rows = nullptr;
  // injury criteria (optional, per step inside the resident loop)
  bool inj_on = false;
  long inj_n = 0;  // included elements
  int *inj_flags = nullptr, *inj_perm = nullptr;
  double *inj_ps_old = nullptr, *inj_psxsr = nullptr, *inj_vol = nullptr, *inj_part = nullptr;
  double *inj_scal = nullptr, *inj_vpart = nullptr, *inj_vsum = nullptr;
  unsigned long long *inj_sel = nullptr;
  unsigned *inj_hist = nullptr;
  // staging for the entry points that take host arrays every call
  double *aux = nullptr;
  size_t aux_len = 0;
  double *scratch = nullptr;  // this device's copy of `scratch` (scratch_address)
  // Run-loop chunks captured as HIP graphs, one per starting parity;
  // graph_steps == 0 launches every step eagerly.
  int graph_steps = 0;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  // asynchronous displacement snapshots (ft_snapshot_begin / ft_snapshot_end)
  double *snap_dev = nullptr;   // 3 nn, node-major copy of u
  Ctl *snap_ctl_dev = nullptr;  // the control slots with it
  double *snap_host = nullptr;  // pinned
  Ctl *snap_ctl_host = nullptr; // pinned
  int snap_ctl_n = 0;
  hipStream_t copy_stream = nullptr;
  hipEvent_t snap_taken = nullptr, snap_done = nullptr;
  bool snap_pending = false;
  int snap_parity = 0;
  // run loop's early stop: the control slots copied behind the launches
  Ctl *h_ctl = nullptr;                // pinned, n_inst entries
  int h_ctl_n = 0;
  hipEvent_t done_ev = nullptr;
  // Persistent run loop (ft_set_persistent): barrier counter, 3 running
  // counters, blocks of the cooperative launch.
  bool persistent = false;
  unsigned long long *p_arrive = nullptr;
  int *p_running = nullptr;
  // stream of the next launch (the engine's stream)
  hipStream_t lstream = nullptr;
  // Constitutive surrogate of materials 7 and 8 (FP32 run loop): network,
  // normalisation block, per-step inputs and outputs of the surrogate
  // segments' Gauss points, audit schedule and per-tile tallies, training
  // samples.
  ft::MlpNet *mlp = nullptr;
  bool surr_ready = false, surr_on = false;
  float *surr_norm = nullptr;
  std::vector<float> h_surr_norm;
  __bf16 *surr_feat = nullptr;
  float *surr_y = nullptr;
  long surr_gps = 0;
  int surr_audit_every = 100, surr_permille = 10;
  long surr_iter = 0;
  double *surr_tally = nullptr;
  float *ds_x = nullptr, *ds_t = nullptr;
  long ds_n = 0, ds_cap = 0;
  // timing: every `timing_every`-th step of the run loop is timed (sample)
  bool timing = false;
  int timing_every = 1;
  long loop_iter = 0;
  bool sample = true;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[kTimingClasses];
  std::vector<hipEvent_t> ev_pool;
  double ms[kTimingClasses] = {0, 0, 0};
  int launches[kTimingClasses] = {0, 0, 0};
};

namespace {

template <typename T>
int dalloc(T **p, size_t count) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess)
    return fail(FT_ERR_ALLOC, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                hipGetErrorString(e));
  return 0;
}

hipEvent_t take_event(ft_engine *g) {
  if (!g->ev_pool.empty()) {
    hipEvent_t e = g->ev_pool.back();
    g->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct Timed {
  ft_engine *g;
  int cls;
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t st;
  Timed(ft_engine *g_, int c) : g(g_), cls(c), st(g_->lstream) {
    if (g->timing && g->sample) {
      a = take_event(g);
      b = take_event(g);
      if (a) (void)hipEventRecord(a, st);
    }
  }
  ~Timed() {
    if (a && b) {
      (void)hipEventRecord(b, st);
      g->ev[cls].push_back({a, b});
    }
  }
};

int collect_timing(ft_engine *g) {
  HIPCHK(hipStreamSynchronize(g->stream));
  for (int c = 0; c < kTimingClasses; ++c) {
    for (auto &pr : g->ev[c]) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
      g->ms[c] += ms;
      g->launches[c] += 1;
      g->ev_pool.push_back(pr.first);
      g->ev_pool.push_back(pr.second);
    }
    g->ev[c].clear();
  }
  return 0;
}

// Captured graphs hold kernel arguments (pointers, tile ranges, launch
// shapes, run parameters): every change of those drops them.
void drop_graphs(ft_engine *g) {
  for (auto &x : g->gexec)
    if (x) {
      (void)hipGraphExecDestroy(x);
      x = nullptr;
    }
}

// The device keeps nodal arrays component-major ([3][nn]); the C ABI uses the
// reference's node-major layout (x[3 n + d]).
std::vector<double> to_soa(const double *aos, size_t nn) {
  std::vector<double> soa(3 * nn);
  for (size_t n = 0; n < nn; ++n)
    for (int d = 0; d < 3; ++d) soa[d * nn + n] = aos[3 * n + d];
  return soa;
}

int h2d_nodal(ft_engine *g, double *dev, const double *aos) {
  std::vector<double> soa = to_soa(aos, g->nn);
  HIPCHK(hipMemcpyAsync(dev, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

int d2h_nodal(ft_engine *g, double *aos, const double *dev) {
  const size_t nn = g->nn;
  std::vector<double> soa(3 * nn);
  HIPCHK(hipMemcpyAsync(soa.data(), dev, soa.size() * sizeof(double), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  for (size_t n = 0; n < nn; ++n)
    for (int d = 0; d < 3; ++d) aos[3 * n + d] = soa[d * nn + n];
  return 0;
}

// Element launches cover tile ranges [t0, t1) of g->etiles.  In the run loop
// (in_loop) each tile reads its instance's step slot; elsewhere dt_fixed is
// the step size and dtbits (per instance, nullable) collects the minimum.
template <typename R>
ElemArgs<R> elem_args(ft_engine *g, int t0, int t1, bool in_loop, double dt_fixed, unsigned long long *dtbits,
                      int update_hist) {
  ElemArgs<R> a;
  a.nn = g->nn;
  a.ne = g->ne;
  a.tiles = g->etiles + t0;
  a.n_tiles = t1 - t0;
  a.conn = g->conn;
  a.einfo = g->einfo;
  a.X = g->X;
  a.u = g->u;
  a.parts = g->parts;
  a.fe = reinterpret_cast<R *>(g->fe);
  a.corner = g->corner ? 1 : 0;
  a.eslot = g->eslot;
  a.fe_ovf = reinterpret_cast<R *>(g->fe_ovf);
  a.ovf = g->ovf;
  a.n_ovf = g->n_ovf;
  a.hist = reinterpret_cast<R *>(g->hist);
  a.fib = reinterpret_cast<const R *>(g->fib);
  a.ctl = in_loop ? g->ctl : nullptr;
  a.parity = g->parity;
  a.refine = (sizeof(R) == 4 && in_loop) ? 1 : 0;
  a.pair = (g->paired && !g->unpaired) ? 1 : 0;
  a.xf = reinterpret_cast<R *>(g->xf);
  a.xidx = g->xidx;
  a.cand_geom = g->cand_geom;
  a.gp_F = a.gp_detF = a.gp_pk2 = nullptr;
  a.perm = nullptr;
  a.dt_fixed = dt_fixed;
  a.dtbits = in_loop ? nullptr : dtbits;
  a.status = g->status;
  a.want_force = 1;
  a.want_dt = 1;
  a.update_hist = update_hist;
  a.surr_y = nullptr;
  a.surr_stride = 0;
  a.surr_e0 = 0;
  a.surr_norm = nullptr;
  a.surr_audit = 0;
  a.surr_permille = 0;
  a.surr_tally = nullptr;
  a.tile0 = t0;
  return a;
}

template <typename R, bool FORCE, bool DT, int MAT, int WAVES = 0>
int launch_element(ft_engine *g, int t0, int t1, bool in_loop, double dt_fixed, unsigned long long *dtbits,
                   int update_hist, int block = kBlock) {
  if (t1 <= t0) return 0;
  ElemArgs<R> a = elem_args<R>(g, t0, t1, in_loop, dt_fixed, dtbits, update_hist);
  a.want_force = FORCE;
  a.want_dt = DT;
  Timed t(g, 0);
  hipLaunchKernelGGL((k_element<R, FORCE, DT, MAT, WAVES>), dim3(grid_for_tiles(t1 - t0)), dim3(block), 0,
                     g->lstream, a);
  HIPCHK(hipGetLastError());
  return 0;
}

// A material segment's force kernel with the segment's launch shape: its
// tile size is the block size; its occupancy variant applies to the FP32
// force + dt kernel (the run loop's), the others use the default.
template <typename R, bool DT, int MAT>
int force_segment_mat(ft_engine *g, const ft_engine::Seg &sg, int waves, bool in_loop, double dt_fixed,
                      unsigned long long *dtbits, int update_hist) {
  const int t0 = sg.tile_begin, t1 = sg.tile_end, b = sg.tile;
  if constexpr (sizeof(R) == 4 && DT) {
    switch (waves) {
      case 1: return launch_element<R, true, DT, MAT, 1>(g, t0, t1, in_loop, dt_fixed, dtbits, update_hist, b);
      case 2: return launch_element<R, true, DT, MAT, 2>(g, t0, t1, in_loop, dt_fixed, dtbits, update_hist, b);
      case 3: return launch_element<R, true, DT, MAT, 3>(g, t0, t1, in_loop, dt_fixed, dtbits, update_hist, b);
      case 4: return launch_element<R, true, DT, MAT, 4>(g, t0, t1, in_loop, dt_fixed, dtbits, update_hist, b);
      case 5: return launch_element<R, true, DT, MAT, 5>(g, t0, t1, in_loop, dt_fixed, dtbits, update_hist, b);
      default: break;
    }
  }
  return launch_element<R, true, DT, MAT>(g, t0, t1, in_loop, dt_fixed, dtbits, update_hist, b);
}

// Scratch (private) bytes per thread of the FP32 run-loop element kernel of
// a material at an occupancy variant (0 when unknown).
template <int MAT>
size_t elem_scratch_mat(int waves) {
  const void *f = nullptr;
  switch (waves) {
    case 1: f = reinterpret_cast<const void *>(&k_element<float, true, true, MAT, 1>); break;
    case 2: f = reinterpret_cast<const void *>(&k_element<float, true, true, MAT, 2>); break;
    case 3: f = reinterpret_cast<const void *>(&k_element<float, true, true, MAT, 3>); break;
    case 4: f = reinterpret_cast<const void *>(&k_element<float, true, true, MAT, 4>); break;
    case 5: f = reinterpret_cast<const void *>(&k_element<float, true, true, MAT, 5>); break;
    default: f = reinterpret_cast<const void *>(&k_element<float, true, true, MAT, 0>); break;
  }
  hipFuncAttributes fa;
  if (hipFuncGetAttributes(&fa, f) != hipSuccess) return 0;
  return fa.localSizeBytes;
}

// Largest register spill (bytes of scratch per lane) of a variant the
// autotuner still times.
constexpr size_t kTuneSpillMax = 128;

size_t elem_scratch(int mat, int waves) {
  switch (mat) {
    case 0: return elem_scratch_mat<0>(waves);
    case 1: return elem_scratch_mat<1>(waves);
    case 2: return elem_scratch_mat<2>(waves);
    case 3: return elem_scratch_mat<3>(waves);
    case 4: return elem_scratch_mat<4>(waves);
    case 5: return elem_scratch_mat<5>(waves);
    case 6: return elem_scratch_mat<6>(waves);
    case 7: return elem_scratch_mat<7>(waves);
    case 8: return elem_scratch_mat<8>(waves);
  }
  return 0;
}

template <typename R, bool DT>
int force_segment(ft_engine *g, const ft_engine::Seg &sg, int waves, bool in_loop, double dt_fixed,
                  unsigned long long *dtbits, int update_hist) {
  switch (sg.mat) {
    case 0: return force_segment_mat<R, DT, 0>(g, sg, waves, in_loop, dt_fixed, dtbits, update_hist);
    case 1: return force_segment_mat<R, DT, 1>(g, sg, waves, in_loop, dt_fixed, dtbits, update_hist);
    case 2: return force_segment_mat<R, DT, 2>(g, sg, waves, in_loop, dt_fixed, dtbits, update_hist);
    case 3: return force_segment_mat<R, DT, 3>(g, sg, waves, in_loop, dt_fixed, dtbits, update_hist);
    case 4: return force_segment_mat<R, DT, 4>(g, sg, waves, in_loop, dt_fixed, dtbits, update_hist);
    case 5: return force_segment_mat<R, DT, 5>(g, sg, waves, in_loop, dt_fixed, dtbits, update_hist);
    case 6: return force_segment_mat<R, DT, 6>(g, sg, waves, in_loop, dt_fixed, dtbits, update_hist);
    case 7: return force_segment_mat<R, DT, 7>(g, sg, waves, in_loop, dt_fixed, dtbits, update_hist);
    case 8: return force_segment_mat<R, DT, 8>(g, sg, waves, in_loop, dt_fixed, dtbits, update_hist);
  }
  return fail(FT_ERR_MATERIAL, "Unknown material type %d", sg.mat);
}

template <typename R>
int element_pass_t(ft_engine *g, bool force, bool dt, bool in_loop, double dt_fixed,
                   unsigned long long *dtbits, int update_hist) {
  // A stand-alone stable-time-step pass always runs in FP64 (exact minimum).
  if (!force) {
    for (const auto &sg : g->segs) {
      int rc = launch_element<double, false, true, -1>(g, sg.tile_begin, sg.tile_end, false, dt_fixed, dtbits, 0,
                                                       sg.tile);
      if (rc) return rc;
    }
    return 0;
  }
  for (const auto &sg : g->segs) {
    int rc = dt ? force_segment<R, true>(g, sg, sg.waves, in_loop, dt_fixed, dtbits, update_hist)
                : force_segment<R, false>(g, sg, sg.waves, in_loop, dt_fixed, dtbits, update_hist);
    if (rc) return rc;
  }
  return 0;
}

bool is_surr_mat(int mat) { return mat == 7 || mat == 8; }

// The run loop's element pass with the surrogate: network inputs of every
// surrogate segment, one inference launch over all their Gauss points, then
// each segment's element kernel (surrogate stress, or the plain kernel for
// the other laws).
int element_pass_surr(ft_engine *g) {
  const bool audit = g->surr_iter % g->surr_audit_every == 0;
  g->surr_iter += 1;
  long off = 0;
  const float *fibf = reinterpret_cast<const float *>(g->fib);
  {
    Timed t(g, 2);
    for (const auto &sg : g->segs) {
      if (!is_surr_mat(sg.mat)) continue;
      const int n = sg.end - sg.begin;
      __bf16 *feat = g->surr_feat + (size_t)off * kSurrFeat;
      const dim3 grid((n + kBlock - 1) / kBlock);
      if (sg.mat == 7)
        hipLaunchKernelGGL(k_surr_features<7>, grid, dim3(kBlock), 0, g->lstream, g->nn, g->ne, g->conn, g->X, g->u,
                           fibf, sg.begin, sg.end, g->surr_norm, feat, g->ctl, g->parity);
      else
        hipLaunchKernelGGL(k_surr_features<8>, grid, dim3(kBlock), 0, g->lstream, g->nn, g->ne, g->conn, g->X, g->u,
                           fibf, sg.begin, sg.end, g->surr_norm, feat, g->ctl, g->parity);
      HIPCHK(hipGetLastError());
      off += 8L * n;
    }
    int rc = ft::mlp_infer(g->mlp, g->surr_feat, g->surr_gps, g->surr_y, g->lstream);
    if (rc) return rc;
  }
  off = 0;
  for (const auto &sg : g->segs) {
    int rc;
    if (!is_surr_mat(sg.mat)) {
      if ((rc = force_segment<float, true>(g, sg, sg.waves, true, 0.0, nullptr, 1))) return rc;
      continue;
    }
    const int n = sg.end - sg.begin;
    ElemArgs<float> a = elem_args<float>(g, sg.tile_begin, sg.tile_end, true, 0.0, nullptr, 1);
    a.surr_y = g->surr_y + (size_t)off * kSurrOut;
    a.surr_stride = (size_t)n * kSurrOut;
    a.surr_e0 = sg.begin;
    a.surr_norm = g->surr_norm;
    a.surr_audit = audit ? 1 : 0;
    a.surr_permille = (unsigned)g->surr_permille;
    a.surr_tally = g->surr_tally;
    a.tile0 = sg.tile_begin;
    Timed t(g, 0);
    const dim3 grid(grid_for_tiles(sg.tile_end - sg.tile_begin));
    if (sg.mat == 7)
      hipLaunchKernelGGL(k_element_surr<7>, grid, dim3(sg.tile), 0, g->lstream, a);
    else
      hipLaunchKernelGGL(k_element_surr<8>, grid, dim3(sg.tile), 0, g->lstream, a);
    HIPCHK(hipGetLastError());
    off += 8L * n;
  }
  return 0;
}

int element_pass(ft_engine *g, bool force, bool dt, bool in_loop, double dt_fixed, unsigned long long *dtbits,
                 int update_hist) {
  if (g->surr_on && force && dt && in_loop && update_hist) return element_pass_surr(g);
  if (g->precision == FT_FP32) return element_pass_t<float>(g, force, dt, in_loop, dt_fixed, dtbits, update_hist);
  return element_pass_t<double>(g, force, dt, in_loop, dt_fixed, dtbits, update_hist);
}

// Motion states of timed motions: on a mesh of at most kInlineMotionTiles
// node tiles each tile that holds a constrained node evaluates them itself
// (no separate launch in the step: 27^3 cube, FP64, 445 steps 15.5 -> 14.9
// ms); on larger meshes k_motion_states evaluates them once per step (on
// the 2.0M-hex ellipsoid, whose every node tile holds surface nodes, the
// in-tile evaluation cost the node update 10 us).
constexpr int kInlineMotionTiles = 1024;
inline bool inline_motion_states(const ft_engine *g) {
  return g->timed_motions && g->n_ntiles <= kInlineMotionTiles;
}

// Nodal launches read the per-instance step caps and end times from
// g->inst_cap / g->inst_tmax (set by the host before a run call).
// Node tiles [t0, t1) (their energy accumulators at the same positions).
template <typename R>
NodalArgs<R> nodal_args(ft_engine *g, int mode, int t0, int t1) {
  NodalArgs<R> a;
  a.nn = g->nn;
  a.ell = g->ell;
  a.width = g->width;
  a.fe = reinterpret_cast<const R *>(g->fe);
  a.ov = Overflow<R>{reinterpret_cast<const R *>(g->fe_ovf), g->ovf, g->n_ovf};
  a.X = g->X;
  a.u = g->u;
  a.vs = g->vs;
  a.fi = g->fi;
  a.mass = g->mass;
  a.node_code = g->node_code;
  a.motions = g->motions;
  a.n_motions = g->n_motions;
  a.mrange = g->mrange;
  a.mstate = g->mstate;
  a.tiles = g->ntiles + t0;
  a.n_tiles = t1 - t0;
  a.ctl = g->ctl;
  a.parity = g->parity;
  a.mode = mode;
  a.energy = 1;
  a.eacc = g->eacc + kEaccStride * (size_t)t0;
  a.tmax = g->inst_tmax;
  a.cap = g->inst_cap;
  a.factor = g->factor;
  a.failure_dt = g->failure_dt;
  a.status = g->status;
  a.inline_motions = inline_motion_states(g) ? 1 : 0;
  a.running = nullptr;
  a.mcap = g->mcap;
  a.cand_geom = g->precision == FT_FP32 ? g->cand_geom : nullptr;
  return a;
}

template <typename R>
int launch_nodal_t(ft_engine *g, int mode, int t0, int t1) {
  if (t1 <= t0) return 0;
  const NodalArgs<R> a = nodal_args<R>(g, mode, t0, t1);
  Timed t(g, 1);
  const dim3 grid(grid_for_tiles(t1 - t0));
  if (g->rot_motions) {
    const size_t lds = nodal_lds_bytes<true>(g->mcap);
    if (g->pair_complete)
      hipLaunchKernelGGL((k_nodal<R, -2, true>), grid, dim3(kBlock), lds, g->lstream, a);
    else if (g->corner)
      hipLaunchKernelGGL((k_nodal<R, -1, true>), grid, dim3(kBlock), lds, g->lstream, a);
    else if (g->width == 8)
      hipLaunchKernelGGL((k_nodal<R, 8, true>), grid, dim3(kBlock), lds, g->lstream, a);
    else
      hipLaunchKernelGGL((k_nodal<R, 0, true>), grid, dim3(kBlock), lds, g->lstream, a);
  } else {
    const size_t lds = nodal_lds_bytes<false>(g->mcap);
    if (g->pair_complete)
      hipLaunchKernelGGL((k_nodal<R, -2, false>), grid, dim3(kBlock), lds, g->lstream, a);
    else if (g->corner)
      hipLaunchKernelGGL((k_nodal<R, -1, false>), grid, dim3(kBlock), lds, g->lstream, a);
    else if (g->width == 8)
      hipLaunchKernelGGL((k_nodal<R, 8, false>), grid, dim3(kBlock), lds, g->lstream, a);
    else
      hipLaunchKernelGGL((k_nodal<R, 0, false>), grid, dim3(kBlock), lds, g->lstream, a);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int launch_nodal(ft_engine *g, int mode) {
  if (g->timed_motions && !inline_motion_states(g)) {
    Timed t(g, 2);
    const int nb = std::max(1, std::min(64, (3 * g->n_motions + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(k_motion_states, dim3(nb), dim3(kBlock), 0, g->lstream, g->ctl, g->parity, g->factor,
                       g->motions, g->motion_inst, g->n_motions, g->mstate,
                       g->precision == FT_FP32 ? (const double *)g->cand_geom : nullptr, g->status);
    HIPCHK(hipGetLastError());
  }
  const int nt = g->n_ntiles;
  int rc = g->precision == FT_FP32 ? launch_nodal_t<float>(g, mode, 0, nt) : launch_nodal_t<double>(g, mode, 0, nt);
  if (rc) return rc;
  g->parity ^= 1;
  return 0;
}

// One cooperative launch of k_persistent for `steps` steps.  The element
// phase runs the neo-Hookean body alone when the whole mesh is material 1,
// otherwise every material's body behind a per-tile switch (the kernel's
// registers are then those of the largest); the per-segment launch shapes of
// the autotuner do not apply.
template <typename R, int MAT, int W>
int launch_persistent_k(ft_engine *g, int steps) {
  PersistArgs<R> p;
  p.ea = elem_args<R>(g, 0, g->n_etiles, true, 0.0, nullptr, 1);
  p.ea.want_force = 1;
  p.ea.want_dt = 1;
  p.na = nodal_args<R>(g, kStep, 0, g->n_ntiles);
  p.na.inline_motions = g->timed_motions ? 1 : 0;
  p.na.cand_geom = g->precision == FT_FP32 ? g->cand_geom : nullptr;
  p.xadd = g->paired ? g->xadd : nullptr;
  p.n_xadd = g->paired ? g->n_xadd : 0;
  p.steps = steps;
  p.parity0 = g->parity;
  p.n_inst = g->n_inst;
  // A phase takes well under 1 ns per element; a poll takes at least 400 ns
  // (the sleep alone): ne / 10 polls leave a margin of 40x or more.
  p.spins = (1ull << 16) + (unsigned long long)g->ne / 10;
  p.arrive = g->p_arrive;
  p.running = g->p_running;
  p.eacc = g->eacc;
  p.nblk = g->ntile_of_inst[1];
  p.rows = g->hist_rows;
  p.hist_cap = g->hist_cap;
  const void *f = reinterpret_cast<const void *>(&k_persistent<R, MAT, W>);
  int per_cu = 0, n_cu = 0, coop = 0;
  HIPCHK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, g->device));
  if (!coop) return fail(FT_ERR_HIP, "the device does not take cooperative launches");
  const size_t lds = nodal_lds_bytes<true>(g->mcap);  // the persistent node phase stages rotations
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kBlock, lds));
  HIPCHK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, g->device));
  const long need = std::max(g->n_etiles, g->n_ntiles);
  long grid = std::min<long>((long)per_cu * n_cu, (need + 7) & ~7L) & ~7L;
  if (grid < 8) return fail(FT_ERR_HIP, "persistent launch: %d blocks per CU do not fill 8 XCDs", per_cu);
  HIPCHK(hipMemsetAsync(g->p_arrive, 0, sizeof(unsigned long long), g->stream));
  HIPCHK(hipMemsetAsync(g->p_running, 0, 3 * sizeof(int), g->stream));
  void *args[] = {&p};
  HIPCHK(hipLaunchCooperativeKernel(f, dim3((unsigned)grid), dim3(kBlock), args, lds, g->stream));
  return 0;
}

int launch_persistent(ft_engine *g, int steps) {
  if (!g->p_arrive) {
    int rc;
    if ((rc = dalloc(&g->p_arrive, 1)) || (rc = dalloc(&g->p_running, 3)))
      return rc;
  }
  bool all_nh = !g->segs.empty();
  for (const auto &sg : g->segs) all_nh = all_nh && sg.mat == 1;
  if (g->precision == FT_FP32) {
    if (g->pair_complete)
      return all_nh ? launch_persistent_k<float, 1, -2>(g, steps) : launch_persistent_k<float, -1, -2>(g, steps);
    if (g->corner) return all_nh ? launch_persistent_k<float, 1, -1>(g, steps) : launch_persistent_k<float, -1, -1>(g, steps);
    return launch_persistent_k<float, -1, 0>(g, steps);
  }
  if (g->pair_complete) return launch_persistent_k<double, -1, -2>(g, steps);
  if (g->corner) return launch_persistent_k<double, -1, -1>(g, steps);
  return launch_persistent_k<double, -1, 0>(g, steps);
}

InjuryArgs injury_args(ft_engine *g, const StepSlot *slot, double time, double dt) {
  InjuryArgs a;
  a.nn = g->nn;
  a.ne = g->ne;
  a.conn = g->conn;
  a.X = g->X;
  a.u = g->u;
  a.perm = g->inj_perm;
  a.flags = g->inj_flags;
  a.ps_old = g->inj_ps_old;
  a.psxsr = g->inj_psxsr;
  a.part = g->inj_part;
  a.scal = g->inj_scal;
  a.slot = slot;
  a.time = time;
  a.dt = dt;
  return a;
}

// One update of the injury criteria (CalculateInjuryCriterions, ex5.cpp:
// 1311-1430) on the stream: time and dt from `slot` when given.
int injury_launch(ft_engine *g, const StepSlot *slot, double time, double dt) {
  Timed t(g, 2);
  const int nblk = (g->ne + kBlock - 1) / kBlock;
  const InjuryArgs a = injury_args(g, slot, time, dt);
  hipLaunchKernelGGL(k_injury_elem, dim3(nblk), dim3(kBlock), 0, g->stream, a);
  hipLaunchKernelGGL(k_injury_reduce, dim3(1), dim3(kBlock), 0, g->stream, a, nblk);
  hipLaunchKernelGGL(k_select_init, dim3(1), dim3(kBlock), 0, g->stream, g->inj_sel, g->inj_hist, g->inj_n, slot);
  const int sgrid = std::min(nblk, 1024);
  for (int pass = 0; pass < 8; ++pass) {
    hipLaunchKernelGGL(k_select_hist, dim3(sgrid), dim3(kBlock), 0, g->stream, g->ne, g->inj_flags, g->inj_ps_old,
                       g->inj_psxsr, g->inj_sel, pass, g->inj_hist, slot);
    hipLaunchKernelGGL(k_select_pick, dim3(1), dim3(kBlock), 0, g->stream, g->inj_sel, g->inj_hist, slot);
  }
  hipLaunchKernelGGL(k_injury_p95, dim3(1), dim3(64), 0, g->stream, g->inj_sel, g->inj_scal, slot, time);
  hipLaunchKernelGGL(k_injury_members, dim3(nblk), dim3(kBlock), 0, g->stream, g->ne, g->inj_flags, g->inj_ps_old,
                     g->inj_psxsr, g->inj_scal, slot);
  HIPCHK(hipGetLastError());
  return 0;
}

// Corner pairs: zero the slots marked in mask[n].
template <typename R>
__global__ void k_zero_slots(int nn, const unsigned char *mask, R *fe) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const unsigned m = mask[n];
  if (!m) return;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if ((m >> k) & 1u) {
      R *dst = fe + 3 * ((size_t)k * nn + n);
      dst[0] = dst[1] = dst[2] = R(0);
    }
}

// which 0: the free slots the run loop writes; 1: the slots it leaves
// unwritten.
int zero_slots(ft_engine *g, int which) {
  if (!g->paired || !g->pzero) return 0;
  const unsigned char *mask = g->pzero + (size_t)which * g->nn;
  if (g->precision == FT_FP32)
    hipLaunchKernelGGL(k_zero_slots<float>, dim3((g->nn + kBlock - 1) / kBlock), dim3(kBlock), 0, g->stream,
                       g->nn, mask, reinterpret_cast<float *>(g->fe));
  else
    hipLaunchKernelGGL(k_zero_slots<double>, dim3((g->nn + kBlock - 1) / kBlock), dim3(kBlock), 0, g->stream,
                       g->nn, mask, reinterpret_cast<double *>(g->fe));
  HIPCHK(hipGetLastError());
  return 0;
}

// Element passes outside the run loop (forces, per-element forces and
// per-Gauss-point fields asked for by the caller) write all eight corners of
// every element into its own slots, as the plain corner layout does, and the
// force read-back sums all eight slots: the free slots the run loop writes
// are zeroed first, and the slots the run loop leaves unwritten are zeroed
// again when the pass's results have been read.
struct Unpaired {
  ft_engine *g;
  explicit Unpaired(ft_engine *e) : g(e) {
    g->unpaired = true;
    (void)zero_slots(g, 0);
  }
  ~Unpaired() {
    g->unpaired = false;
    (void)zero_slots(g, 1);
  }
};

template <typename R>
__global__ __launch_bounds__(kBlock) void k_pair_extras(int n, const int2 *xadd, const R *xf, R *fe) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pair_extras(i, xadd, xf, fe);
}

int launch_pair_extras(ft_engine *g) {
  if (!g->paired || g->n_xadd == 0) return 0;
  Timed t(g, 2);
  const dim3 grid((g->n_xadd + kBlock - 1) / kBlock);
  if (g->precision == FT_FP32)
    hipLaunchKernelGGL(k_pair_extras<float>, grid, dim3(kBlock), 0, g->lstream, g->n_xadd, g->xadd,
                       reinterpret_cast<const float *>(g->xf), reinterpret_cast<float *>(g->fe));
  else
    hipLaunchKernelGGL(k_pair_extras<double>, grid, dim3(kBlock), 0, g->lstream, g->n_xadd, g->xadd,
                       reinterpret_cast<const double *>(g->xf), reinterpret_cast<double *>(g->fe));
  HIPCHK(hipGetLastError());
  return 0;
}

// einfo's pair bits from bits[e] (1: kPairLeft, 2: kPairRight, 4: kPlusToMinus,
// 8: kExtra; bits 4-7: kExtraFree + 0..3).
__global__ void k_pair_bits(int ne, const unsigned char *bits, int *einfo) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int b = bits[e];
  einfo[e] = (einfo[e] & ~(kPairLeft | kPairRight | kPlusToMinus | kExtra | (15 << kExtraFree))) |
             ((b & 1) ? kPairLeft : 0) | ((b & 2) ? kPairRight : 0) | ((b & 4) ? kPlusToMinus : 0) |
             ((b & 8) ? kExtra : 0) | (((b >> 4) & 15) << kExtraFree);
}

// Corner pairs for the current layout and element tiles (FT_LAYOUT_CORNER_PAIRS;
// every other layout clears them): element i pairs with i - 1 when both lie
// in one material segment and instance and i's corners 0, 3, 4, 7 are
// i - 1's corners 1, 2, 5, 6 (a face shared along xi: consecutive elements of
// a row of a structured block).  Sets the pair bits in einfo, decides
// whether pairing is complete and zeroes the slots the run loop leaves
// unwritten.
int build_pairs(ft_engine *g) {
  const int ne = g->ne, nn = g->nn;
  std::vector<unsigned char> bits(ne, 0), zero(2 * (size_t)nn, 0);
  std::vector<int> xidx(ne, 0);
  std::vector<int2> xadd;
  int n_extras = 0;
  long pairs = 0;
  bool complete = false;
  if (g->layout == FT_LAYOUT_CORNER_PAIRS && g->corner && !g->eslot && !g->h_etiles.empty()) {
    const int *conn = g->h_conn.data();
    std::vector<int> inst_of(ne, 0);
    for (int i = 0; i < g->n_inst; ++i)
      for (int e = g->inst_elem[i]; e < g->inst_elem[i + 1]; ++e) inst_of[e] = i;
    // with shadow lanes any two consecutive elements of a segment and
    // instance meet in one wavefront; without (the Prony laws) only those
    // after a wavefront's first lane
    for (const int4 &t : g->h_etiles)
      for (int i = t.x; i < t.y; ++i) {
        const bool meet = (t.w & kTileShadows) ? (i > t.x || (t.w & kTileShadow)) : (i - t.x) % 64 != 0;
        if (!meet) continue;
        const int *a = conn + 8 * (size_t)g->perm[i - 1], *b = conn + 8 * (size_t)g->perm[i];
        if (inst_of[g->perm[i - 1]] == inst_of[g->perm[i]] && b[0] == a[1] && b[3] == a[2] && b[4] == a[5] &&
            b[7] == a[6]) {
          bits[i] |= 1;
          bits[i - 1] |= 2;
          ++pairs;
        }
      }
    // slots some element owns (it has the node at that local corner)
    std::vector<unsigned char> owned(nn, 0);
    for (long i = 0; i < 8L * ne; ++i) owned[conn[i]] |= (unsigned char)(1u << (i % 8));
    // An element without a right partner writes each xi+ corner to the
    // matching xi- slot of its node when no element owns that slot, else to
    // an extras entry added to it after the element pass (the same choices
    // in a batch as alone, so the instances' sums stay bitwise the same).
    // Every corner force then reaches a xi- slot.  Slots to zero before a
    // pass outside the loop: the free slots the loop writes; after it: the
    // xi+ slots the loop does not write.
    complete = pairs > 0;
    int n_extra = 0;
    for (int i = 0; i < ne && complete; ++i) {
      if (bits[i] & 2) continue;
      const int *c = conn + 8 * (size_t)g->perm[i];
      int free_bits = 0;
      for (int j = 0; j < 4; ++j) {
        const int k = (0x6521 >> (4 * j)) & 15;  // 1, 2, 5, 6
        if (!((owned[c[k]] >> (k ^ 1)) & 1)) {
          free_bits |= 1 << j;
          zero[c[k]] |= (unsigned char)(1u << (k ^ 1));
        } else {
          xadd.push_back(make_int2(3 * ((k ^ 1) * nn + c[k]), 4 * n_extra + j));
        }
      }
      if (free_bits == 15) {
        bits[i] |= 4;
      } else {
        bits[i] |= 8 | (free_bits << 4);
        xidx[i] = 4 * n_extra++;
      }
    }
    for (int i = 0; i < ne && complete; ++i) {
      const int *c = conn + 8 * (size_t)g->perm[i];
      for (int k : {1, 2, 5, 6}) zero[nn + (size_t)c[k]] |= (unsigned char)(1u << k);
    }
    n_extras = n_extra;
  }
  unsigned char *dbits = nullptr;
  int rc = dalloc(&dbits, (size_t)ne);
  if (rc) return rc;
  if (hipMemcpy(dbits, bits.data(), ne, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(dbits);
    return fail(FT_ERR_HIP, "pair bit upload failed");
  }
  hipLaunchKernelGGL(k_pair_bits, dim3((ne + kBlock - 1) / kBlock), dim3(kBlock), 0, g->stream, ne, dbits, g->einfo);
  const hipError_t e1 = hipGetLastError(), e2 = hipStreamSynchronize(g->stream);
  (void)hipFree(dbits);
  if (e1 != hipSuccess || e2 != hipSuccess) return fail(FT_ERR_HIP, "pair bits: %s", hipGetErrorString(e1 ? e1 : e2));
  if (pairs > 0 && !g->pzero && (rc = dalloc(&g->pzero, 2 * (size_t)nn))) return rc;
  if (pairs > 0) HIPCHK(hipMemcpy(g->pzero, zero.data(), 2 * (size_t)nn, hipMemcpyHostToDevice));
  for (void *p : {g->xf, (void *)g->xidx, (void *)g->xadd})
    if (p) (void)hipFree(p);
  g->xf = nullptr;
  g->xidx = nullptr;
  g->xadd = nullptr;
  g->n_xadd = 0;
  if (n_extras > 0) {
    const size_t esz = g->precision == FT_FP32 ? 4 : 8;
    if (hipMalloc(&g->xf, 12 * (size_t)n_extras * esz) != hipSuccess || hipMemset(g->xf, 0, 12 * (size_t)n_extras * esz) != hipSuccess ||
        (rc = dalloc(&g->xidx, (size_t)ne)) || (rc = dalloc(&g->xadd, xadd.size())))
      return rc ? rc : fail(FT_ERR_ALLOC, "corner-pair extras allocation failed");
    HIPCHK(hipMemcpy(g->xidx, xidx.data(), ne * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g->xadd, xadd.data(), xadd.size() * sizeof(int2), hipMemcpyHostToDevice));
    g->n_xadd = (int)xadd.size();
  }
  g->paired = pairs > 0;
  g->pair_complete = complete;
  g->n_pairs = pairs;
  // slots the new pairing leaves unwritten may hold earlier forces
  if ((rc = zero_slots(g, 0)) || (rc = zero_slots(g, 1))) return rc;
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

// Element tiles: at most sg.tile consecutive elements of one material
// segment (internal order) and one instance.
int build_element_tiles(ft_engine *g) {
  std::vector<int> inst_of_elem(g->ne);
  for (int i = 0; i < g->n_inst; ++i)
    for (int e = g->inst_elem[i]; e < g->inst_elem[i + 1]; ++e) inst_of_elem[e] = i;
  // corner pairs: 63 elements per wavefront (lane 0 recomputes the element
  // before them)
  bool shadows = false;
  if (g->layout == FT_LAYOUT_CORNER_PAIRS && g->corner && !g->eslot) {
    const int *conn = g->h_conn.data();
    for (const auto &sg : g->segs)
      for (int i = sg.begin + 1; i < sg.end && !shadows; ++i) {
        const int *a = conn + 8 * (size_t)g->perm[i - 1], *b = conn + 8 * (size_t)g->perm[i];
        shadows = b[0] == a[1] && b[3] == a[2] && b[4] == a[5] && b[7] == a[6];
      }
  }
  std::vector<int4> et;
  for (auto &sg : g->segs) {
    sg.tile_begin = (int)et.size();
    const bool seg_shadows = shadows && sg.mat != 5 && sg.mat != 7;
    const int cap = seg_shadows ? sg.tile / 64 * 63 : sg.tile;
    for (int i = sg.begin; i < sg.end;) {
      const int inst = inst_of_elem[g->perm[i]];
      int j = i + 1;
      while (j < sg.end && j - i < cap && inst_of_elem[g->perm[j]] == inst) ++j;
      const bool shadow_ok = seg_shadows && i > sg.begin && inst_of_elem[g->perm[i - 1]] == inst;
      et.push_back(make_int4(i, j, inst, (seg_shadows ? kTileShadows : 0) | (shadow_ok ? kTileShadow : 0)));
      i = j;
    }
    sg.tile_end = (int)et.size();
  }
  for (void *p : {(void *)g->etiles, (void *)g->cand_geom})
    if (p) (void)hipFree(p);
  g->etiles = nullptr;
  g->cand_geom = nullptr;
  g->n_etiles = (int)et.size();
  g->h_etiles = et;
  int rc;
  if ((rc = dalloc(&g->etiles, et.size())) || (rc = dalloc(&g->cand_geom, kCandStride * et.size())))
    return rc;
  HIPCHK(hipMemcpy(g->etiles, et.data(), et.size() * sizeof(int4), hipMemcpyHostToDevice));
  return build_pairs(g);  // wavefront boundaries moved
}

// kTileMotion on the node tiles that hold a constrained node (every tile
// while no boundary is known), uploaded with the tiles.
int mark_motion_tiles(ft_engine *g) {
  const bool known = g->h_code.size() == (size_t)g->nn;
  for (auto &t : g->h_ntiles) {
    bool bc = !known;
    for (int n = t.x; n < t.y && !bc; ++n) bc = g->h_code[n] != 0;
    t.w = (t.w & ~kTileMotion) | (bc ? kTileMotion : 0);
  }
  if (!g->ntiles || g->h_ntiles.empty()) return 0;
  HIPCHK(hipMemcpy(g->ntiles, g->h_ntiles.data(), g->h_ntiles.size() * sizeof(int4), hipMemcpyHostToDevice));
  return 0;
}

// Tiles of consecutive entries that never span two instances: elements per
// material segment, nodes (kBlock per tile) in order.
int build_tiles(ft_engine *g) {
  int rc;
  if ((rc = build_element_tiles(g))) return rc;
  std::vector<int4> nt;
  g->ntile_of_inst.assign(g->n_inst + 1, 0);
  for (int i = 0; i < g->n_inst; ++i) {
    g->ntile_of_inst[i] = (int)nt.size();
    for (int n = g->inst_node[i]; n < g->inst_node[i + 1]; n += kBlock)
      nt.push_back(make_int4(n, std::min(n + kBlock, g->inst_node[i + 1]), i,
                             n == g->inst_node[i] ? kTileFirst : 0));
  }
  g->ntile_of_inst[g->n_inst] = (int)nt.size();
  g->h_ntiles = nt;
  for (void *p : {(void *)g->ntiles, (void *)g->eacc})
    if (p) (void)hipFree(p);
  g->ntiles = nullptr;
  g->eacc = nullptr;
  g->n_ntiles = (int)nt.size();
  if ((rc = dalloc(&g->ntiles, nt.size())) ||
      (rc = dalloc(&g->eacc, kEaccStride * (size_t)std::max(1, g->n_ntiles))))
    return rc;
  HIPCHK(hipMemset(g->eacc, 0, kEaccStride * (size_t)std::max(1, g->n_ntiles) * sizeof(double)));
  return mark_motion_tiles(g);
}

// Per-instance arrays of the control state (re)allocated for g->n_inst.
int alloc_instances(ft_engine *g) {
  for (void *p : {(void *)g->ctl, (void *)g->dtbits, (void *)g->mrange, (void *)g->inst_cap, (void *)g->inst_tmax})
    if (p) (void)hipFree(p);
  g->ctl = nullptr;
  g->dtbits = nullptr;
  g->mrange = nullptr;
  g->inst_cap = nullptr;
  g->inst_tmax = nullptr;
  const size_t n = g->n_inst;
  int rc;
  if ((rc = dalloc(&g->ctl, n)) || (rc = dalloc(&g->dtbits, n)) || (rc = dalloc(&g->mrange, 2 * n)) ||
      (rc = dalloc(&g->inst_cap, n)) || (rc = dalloc(&g->inst_tmax, n)))
    return rc;
  HIPCHK(hipMemset(g->ctl, 0, n * sizeof(Ctl)));
  HIPCHK(hipMemset(g->mrange, 0, 2 * n * sizeof(int)));
  HIPCHK(hipMemset(g->inst_cap, 0, n * sizeof(int)));
  g->h_tmax.assign(n, 1e300);
  return 0;
}

// Boundary codes from the caller's table: each instance gets its own copies
// of the motions its DOFs use (the motion states depend on the instance's
// time), at most kMaxMotions of them, addressed instance-locally by the node
// codes (byte d = local index + 1).
int rebuild_boundary(ft_engine *g) {
  const int nn = g->nn;
  std::vector<int> code(nn, 0);
  std::vector<unsigned char> mask(nn, 0);
  std::vector<double> mexp;
  std::vector<int> minst, mr(2 * g->n_inst, 0);
  const int nm = (int)(g->h_motions.size() / FT_MOTION_STRIDE);
  std::vector<int> local(nm, -1);
  for (int i = 0; i < g->n_inst; ++i) {
    std::fill(local.begin(), local.end(), -1);
    mr[2 * i] = (int)minst.size();
    for (int n = g->inst_node[i]; n < g->inst_node[i + 1]; ++n)
      for (int d = 0; d < 3; ++d) {
        const int k = g->h_dof_motion.empty() ? -1 : g->h_dof_motion[3 * (size_t)n + d];
        if (k < 0) continue;
        if (local[k] < 0) {
          if (mr[2 * i + 1] == kMaxMotions)
            return fail(FT_ERR_ARG, "instance %d uses more than %d distinct motions", i, kMaxMotions);
          local[k] = mr[2 * i + 1]++;
          mexp.insert(mexp.end(), g->h_motions.begin() + FT_MOTION_STRIDE * (size_t)k,
                      g->h_motions.begin() + FT_MOTION_STRIDE * (size_t)(k + 1));
          minst.push_back(i);
        }
        mask[n] |= (unsigned char)(1 << d);
        code[n] |= (local[k] + 1) << (8 * d);
      }
  }
  g->n_motions = (int)minst.size();
  g->mcap = 1;
  for (int i = 0; i < g->n_inst; ++i) g->mcap = std::max(g->mcap, mr[2 * i + 1]);
  g->timed_motions = false;
  g->rot_motions = false;
  for (int k = 0; k < g->n_motions; ++k) {
    const int type = (int)mexp[FT_MOTION_STRIDE * (size_t)k];
    if (type != FT_MOTION_FIXED) g->timed_motions = true;
    if (type == FT_MOTION_ROT_HALFSINE) g->rot_motions = true;
  }
  for (void *p : {(void *)g->motions, (void *)g->motion_inst, (void *)g->mstate})
    if (p) (void)hipFree(p);
  g->motions = nullptr;
  g->motion_inst = nullptr;
  g->mstate = nullptr;
  int rc;
  if ((rc = dalloc(&g->motions, std::max<size_t>(1, mexp.size()))) ||
      (rc = dalloc(&g->motion_inst, std::max<size_t>(1, minst.size()))) ||
      (rc = dalloc(&g->mstate, 3 * std::max<size_t>(1, minst.size()))))
    return rc;
  if (!mexp.empty()) {
    HIPCHK(hipMemcpy(g->motions, mexp.data(), mexp.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g->motion_inst, minst.data(), minst.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemset(g->mstate, 0, 3 * std::max<size_t>(1, minst.size()) * sizeof(MotionT)));
  HIPCHK(hipMemcpy(g->mrange, mr.data(), mr.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(g->node_code, code.data(), nn * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(g->node_mask, mask.data(), nn, hipMemcpyHostToDevice));
  g->h_code = code;
  if ((rc = mark_motion_tiles(g))) return rc;
  hipLaunchKernelGGL(k_elem_flags, dim3((g->ne + 255) / 256), dim3(256), 0, g->stream, g->ne, g->conn,
                     g->node_mask, g->einfo);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

// Step caps of the next launches: cap[i] (host) for every instance.
int upload_caps(ft_engine *g, const std::vector<int> &cap) {
  HIPCHK(hipMemcpyAsync(g->inst_cap, cap.data(), cap.size() * sizeof(int), hipMemcpyHostToDevice, g->stream));
  return 0;
}

int read_slots(ft_engine *g, int parity, std::vector<StepSlot> &s) {
  std::vector<Ctl> c(g->n_inst);
  HIPCHK(hipMemcpyAsync(c.data(), g->ctl, c.size() * sizeof(Ctl), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  s.resize(g->n_inst);
  for (int i = 0; i < g->n_inst; ++i) s[i] = c[i].slot[parity];
  return 0;
}

int upload_u(ft_engine *g, const double *u) {
  if (u) return h2d_nodal(g, g->u, u);
  return 0;
}

// Device staging of `len` doubles for host-array entry points (kept between
// calls; grows on demand).
int aux_reserve(ft_engine *g, size_t len) {
  if (g->aux_len >= len) return 0;
  if (g->aux) (void)hipFree(g->aux);
  g->aux = nullptr;
  g->aux_len = 0;
  int rc = dalloc(&g->aux, len);
  if (rc) return rc;
  g->aux_len = len;
  return 0;
}

int check_status(ft_engine *g) {
  int st = 0;
  HIPCHK(hipMemcpyAsync(&st, g->status, sizeof(int), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  if (st & 8) return fail(FT_ERR_MATERIAL, "unknown material type");
  return 0;
}

int read_slot(ft_engine *g, int parity, StepSlot *s) {
  HIPCHK(hipMemcpyAsync(s, &g->ctl->slot[parity], sizeof(StepSlot), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

}  // namespace

extern "C" {

const char *ft_last_error(void) { return g_err.c_str(); }
int ft_version(void) { return 1; }

int ft_device_count(int *count) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  if (count) *count = c;
  return 0;
}

// Element-force layout (FT_LAYOUT_*): corner slots when every node is local
// corner k of at most one element for each k; node slots when every node
// has at most 8 elements (slot = rank of the element among the node's
// elements in the caller's element order, so the node sum adds them in the
// reference's scatter-add order); otherwise element-major slots read through
// the ELL.  AUTO takes the first that applies.  The force buffer is
// (re)allocated and zeroed: unwritten node-major slots read as zero for the
// whole run.
static int set_layout(ft_engine *g, int mode) {
  const int nn = g->nn, ne = g->ne;
  const int *conn = g->h_conn.data();
  bool corner_ok = true;
  {
    std::vector<char> seen(8 * (size_t)nn, 0);
    for (long i = 0; i < 8L * ne && corner_ok; ++i) {
      char &c = seen[(size_t)(i % 8) * nn + conn[i]];
      if (c) corner_ok = false;
      c = 1;
    }
  }
  const bool slots_ok = g->width <= 16;  // ranks 8-15 in overflow slots
  if (mode == FT_LAYOUT_AUTO)
    mode = corner_ok ? FT_LAYOUT_CORNER_PAIRS : slots_ok ? FT_LAYOUT_NODE_SLOTS : FT_LAYOUT_ELL;
  if ((mode == FT_LAYOUT_CORNER || mode == FT_LAYOUT_CORNER_PAIRS) && !corner_ok)
    return fail(FT_ERR_ARG, "corner-slot layout: a node is the same local corner of two elements");
  if (mode == FT_LAYOUT_NODE_SLOTS && !slots_ok)
    return fail(FT_ERR_ARG, "node-slot layout: a node has %d elements (at most 16)", g->width);
  if (mode != FT_LAYOUT_CORNER && mode != FT_LAYOUT_CORNER_PAIRS && mode != FT_LAYOUT_NODE_SLOTS &&
      mode != FT_LAYOUT_ELL)
    return fail(FT_ERR_ARG, "unknown layout %d", mode);
  // The new buffers are built first; the engine keeps its old layout if
  // any step fails.
  const bool corner = mode != FT_LAYOUT_ELL;
  const size_t esz = g->precision == FT_FP32 ? 4 : 8;
  const size_t fe_len = 24 * (size_t)(corner ? nn : ne);
  void *fe = nullptr, *fe_ovf = nullptr;
  unsigned *eslot = nullptr;
  int *ovf = nullptr, n_ovf = 0;
  std::vector<unsigned> h_eslot;
  std::vector<int> h_ovf;
  auto undo = [&](int rc) {
    for (void *p : {fe, (void *)eslot, (void *)ovf, fe_ovf})
      if (p) (void)hipFree(p);
    return rc;
  };
  if (hipMalloc(&fe, fe_len * esz) != hipSuccess || hipMemset(fe, 0, fe_len * esz) != hipSuccess)
    return undo(fail(FT_ERR_ALLOC, "element force buffer allocation failed"));
  if (mode == FT_LAYOUT_NODE_SLOTS) {
    h_eslot.assign(ne, 0u);
    std::vector<int> rank(nn, 0);
    for (int e = 0; e < ne; ++e)
      for (int k = 0; k < 8; ++k) h_eslot[e] |= (unsigned)rank[conn[8 * (size_t)e + k]]++ << (4 * k);
    std::vector<unsigned> dev(ne);
    for (int i = 0; i < ne; ++i) dev[i] = h_eslot[g->perm[i]];
    int rc = dalloc(&eslot, (size_t)ne);
    if (rc) return undo(rc);
    if (hipMemcpy(eslot, dev.data(), ne * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess)
      return undo(fail(FT_ERR_HIP, "node-slot upload failed"));
    if (g->width > 8) {
      // nodes with more than 8 elements: ranks 8.. in [rank - 8][overflow node][xyz]
      h_ovf.assign(nn, 0);
      for (int n = 0; n < nn; ++n)
        if (rank[n] > 8) h_ovf[n] = (n_ovf++ << 4) | (rank[n] - 8);
      if ((rc = dalloc(&ovf, (size_t)nn))) return undo(rc);
      const size_t len = 3 * (size_t)(g->width - 8) * n_ovf * esz;
      if (hipMemcpy(ovf, h_ovf.data(), nn * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
          hipMalloc(&fe_ovf, len) != hipSuccess || hipMemset(fe_ovf, 0, len) != hipSuccess)
        return undo(fail(FT_ERR_ALLOC, "overflow slot set-up failed"));
    }
  }
  for (void *p : {g->fe, (void *)g->eslot, (void *)g->ovf, g->fe_ovf})
    if (p) (void)hipFree(p);
  g->fe = fe;
  g->eslot = eslot;
  g->ovf = ovf;
  g->fe_ovf = fe_ovf;
  g->n_ovf = n_ovf;
  g->h_eslot.swap(h_eslot);
  g->h_ovf.swap(h_ovf);
  g->corner = corner;
  g->layout = mode;
  // corner pairs change the elements per tile
  return g->h_etiles.empty() ? build_pairs(g) : build_element_tiles(g);
}

int ft_engine_set_layout(ft_engine *g, int mode) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (g->run_active) return fail(FT_ERR_STATE, "the layout is chosen before ft_run_init");
  HIPCHK(hipSetDevice(g->device));
  HIPCHK(hipStreamSynchronize(g->stream));
  drop_graphs(g);  // captured launches hold the old force buffer
  return set_layout(g, mode);
}

int ft_engine_create(int n_nodes, int n_elems, const double *coords, const int *conn,
                     const int *pid, int n_parts, const int *material_id, const double *props,
                     int precision, ft_engine **out) {
  if (!out) return fail(FT_ERR_ARG, "out is null");
  *out = nullptr;
  if (n_nodes <= 0 || n_elems <= 0 || n_parts <= 0 || !coords || !conn || !pid || !material_id ||
      !props)
    return fail(FT_ERR_ARG, "invalid mesh arguments");
  if (precision != FT_FP64 && precision != FT_FP32) return fail(FT_ERR_ARG, "bad precision");
  // 24 n stays below 2^31: element and node arrays are addressed with 32-bit
  // byte offsets from scalar bases in the kernels.
  if ((long)n_elems * 24 > (long)INT32_MAX || (long)n_nodes * 24 > (long)INT32_MAX || n_parts > kPidMask)
    return fail(FT_ERR_ARG, "mesh too large");
  for (long i = 0; i < 8L * n_elems; ++i)
    if (conn[i] < 0 || conn[i] >= n_nodes) return fail(FT_ERR_INPUT, "connectivity index out of range at %ld", i);
  for (int e = 0; e < n_elems; ++e)
    if (pid[e] < 0 || pid[e] >= n_parts) return fail(FT_ERR_INPUT, "part id out of range at element %d", e);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(FT_ERR_HIP, "no HIP device available");
  ft_engine *g = new ft_engine;
  g->nn = n_nodes;
  g->ne = n_elems;
  g->nparts = n_parts;
  g->precision = precision;
  auto bail = [&](int code) {
    ft_engine_destroy(g);
    return code;
  };
  // a HIP failure past this point destroys the partly built engine
#define HIPCHK_BAIL(call)                                                                   \
  do {                                                                                      \
    const hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                                   \
      return bail(fail(FT_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)));         \
  } while (0)
  HIPCHK_BAIL(hipGetDevice(&g->device));
  int rc = 0;
  if (hipStreamCreate(&g->stream) != hipSuccess) return bail(fail(FT_ERR_HIP, "stream creation failed"));
  if ((rc = scratch_address(g->device, &g->scratch))) return bail(rc);
  g->lstream = g->stream;
  // parts (CalculateTimeStep.cpp:13-19 wave speed kept in FP64)
  g->h_parts.resize(n_parts);
  for (int p = 0; p < n_parts; ++p) {
    PartDev &q = g->h_parts[p];
    memset(&q, 0, sizeof(q));
    const double *pr = props + 9 * (size_t)p;
    q.mat = material_id[p];
    if (q.mat < 0 || q.mat > 8) return bail(fail(FT_ERR_MATERIAL, "Unknown material type %d for part %d", q.mat, p));
    q.visco = q.mat == 5 || q.mat == 7;
    q.rho = pr[0];
    q.mu = pr[1];
    q.lambda = pr[2];
    q.k1 = pr[3];
    q.k2 = pr[4];
    q.g1 = pr[5];
    q.t1 = pr[6];
    q.g2 = pr[7];
    q.t2 = pr[8];
    q.alpha = pr[3];  // Ogden (6, 7)
    q.kappa = pr[5];  // anisotropic HGO (8)
    if ((q.mat == 6 || q.mat == 7) && q.alpha == 0.0)
      return bail(fail(FT_ERR_INPUT, "Ogden exponent of part %d is zero", p));
    double nu = 0.5 * q.lambda / (q.lambda + q.mu);
    q.ce = sqrt(q.lambda * (1.0 / nu - 1.0) / q.rho);
    q.inv_ce = 1.0 / q.ce;
    if (q.visco) g->has_visco = true;
  }
  const size_t nd = 3 * (size_t)n_nodes;
  if ((rc = dalloc(&g->X, nd)) || (rc = dalloc(&g->conn, 8 * (size_t)n_elems)) ||
      (rc = dalloc(&g->einfo, (size_t)n_elems)) || (rc = dalloc(&g->parts, (size_t)n_parts)) ||
      (rc = dalloc(&g->u, nd)) || (rc = dalloc(&g->vs, nd)) || (rc = dalloc(&g->fi, nd)) ||
      (rc = dalloc(&g->tmp3, nd)) || (rc = dalloc(&g->mass, (size_t)n_nodes)) ||
      (rc = dalloc(&g->node_code, (size_t)n_nodes)) || (rc = dalloc(&g->node_mask, (size_t)n_nodes)) ||
      (rc = dalloc(&g->status, 1)))
    return bail(rc);
  const size_t esz = precision == FT_FP32 ? 4 : 8;
  g->h_conn.assign(conn, conn + 8 * (size_t)n_elems);
  if (g->has_visco && hipMalloc(&g->hist, hist_len(n_elems) * esz) != hipSuccess)
    return bail(fail(FT_ERR_ALLOC, "history allocation failed"));
  g->nodal_grid = grid_for(n_nodes);
  // Group elements by material (stable counting sort on the material id).
  g->perm.resize(n_elems);
  g->inv.resize(n_elems);
  {
    constexpr int kMats = 9;
    int cnt[kMats + 1] = {0};
    for (int e = 0; e < n_elems; ++e) cnt[g->h_parts[pid[e]].mat + 1]++;
    for (int m = 0; m < kMats; ++m) cnt[m + 1] += cnt[m];
    for (int m = 0; m < kMats; ++m)
      if (cnt[m + 1] > cnt[m]) g->segs.push_back({m, cnt[m], cnt[m + 1]});
    std::vector<int> fill(cnt, cnt + kMats);
    for (int e = 0; e < n_elems; ++e) {
      int i = fill[g->h_parts[pid[e]].mat]++;
      g->perm[i] = e;
      g->inv[e] = i;
    }
  }
  // node -> (slot, element) CSR with entries slot * ne + element (the index of
  // the element-force triple in the slot-major buffer); each node's entries
  // follow the original element order, i.e. the reference's scatter-add
  // order (counting sort: O(N)).
  std::vector<int> val(n_nodes, 0);
  for (long i = 0; i < 8L * n_elems; ++i) val[conn[i]]++;
  g->width = std::max(8, *std::max_element(val.begin(), val.end()));
  std::vector<int> hell((size_t)g->width * n_nodes, -1);
  {
    std::fill(val.begin(), val.end(), 0);
    for (int e = 0; e < n_elems; ++e)
      for (int k = 0; k < 8; ++k) {
        const int n = conn[8 * e + k];
        hell[(size_t)val[n]++ * n_nodes + n] = k * n_elems + g->inv[e];
      }
  }
  if ((rc = dalloc(&g->ell, hell.size()))) return bail(rc);
  if ((rc = set_layout(g, FT_LAYOUT_AUTO))) return bail(rc);
  std::vector<int> hinfo(n_elems), hconn(8 * (size_t)n_elems);
  for (int i = 0; i < n_elems; ++i) {
    const int e = g->perm[i];
    hinfo[i] = pid[e];
    for (int k = 0; k < 8; ++k) hconn[8 * (size_t)i + k] = conn[8 * (size_t)e + k];
  }
  {
    std::vector<double> soa = to_soa(coords, n_nodes);
    HIPCHK_BAIL(hipMemcpy(g->X, soa.data(), nd * sizeof(double), hipMemcpyHostToDevice));
  }
  HIPCHK_BAIL(hipMemcpy(g->conn, hconn.data(), 8 * (size_t)n_elems * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK_BAIL(hipMemcpy(g->einfo, hinfo.data(), (size_t)n_elems * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK_BAIL(hipMemcpy(g->parts, g->h_parts.data(), n_parts * sizeof(PartDev), hipMemcpyHostToDevice));
  HIPCHK_BAIL(hipMemcpy(g->ell, hell.data(), hell.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK_BAIL(hipMemset(g->u, 0, nd * sizeof(double)));
  HIPCHK_BAIL(hipMemset(g->vs, 0, nd * sizeof(double)));
  HIPCHK_BAIL(hipMemset(g->fi, 0, nd * sizeof(double)));
  HIPCHK_BAIL(hipMemset(g->node_code, 0, (size_t)n_nodes * sizeof(int)));
  HIPCHK_BAIL(hipMemset(g->node_mask, 0, (size_t)n_nodes));
  HIPCHK_BAIL(hipMemset(g->status, 0, sizeof(int)));
  if (g->hist) HIPCHK_BAIL(hipMemset(g->hist, 0, hist_len(n_elems) * esz));
  g->h_dof_motion.assign(nd, -1);
  // one instance covering the whole mesh
  g->n_inst = 1;
  g->inst_node = {0, n_nodes};
  g->inst_elem = {0, n_elems};
  if ((rc = alloc_instances(g)) || (rc = build_tiles(g)) || (rc = rebuild_boundary(g))) return bail(rc);
  HIPCHK_BAIL(hipDeviceSynchronize());
  *out = g;
  return 0;
#undef HIPCHK_BAIL
}

int ft_engine_destroy(ft_engine *g) {
  if (!g) return 0;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  drop_graphs(g);
  if (g->mlp) ft::mlp_destroy(g->mlp);
  for (void *p : {(void *)g->surr_norm, (void *)g->surr_feat, (void *)g->surr_y, (void *)g->surr_tally,
                  (void *)g->ds_x, (void *)g->ds_t, (void *)g->p_arrive, (void *)g->p_running})
    if (p) (void)hipFree(p);
  void *ptrs[] = {g->X, g->conn, g->einfo, g->parts, g->u, g->vs, g->fi, g->tmp3, g->mass,
                  g->node_code, g->node_mask, g->motions, g->mstate, g->ctl, g->dtbits, g->status,
                  g->ell, g->fe, g->eslot, g->ovf, g->fe_ovf, g->hist, g->fib, g->eacc, g->hist_rows, g->aux, g->inj_flags, g->inj_perm,
                  g->etiles, g->cand_geom, g->pzero, g->xf, g->xidx, g->xadd, g->gp_fields, g->perm_dev, g->ntiles, g->mrange, g->motion_inst, g->inst_cap, g->inst_tmax,
                  g->inj_ps_old, g->inj_psxsr, g->inj_vol, g->inj_part, g->inj_scal, g->inj_vpart,
                  g->inj_vsum, g->inj_sel, g->inj_hist};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (int c = 0; c < kTimingClasses; ++c)
    for (auto &pr : g->ev[c]) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  for (auto e : g->ev_pool) (void)hipEventDestroy(e);
  if (g->done_ev) (void)hipEventDestroy(g->done_ev);
  if (g->copy_stream) (void)hipStreamSynchronize(g->copy_stream);
  for (hipEvent_t e : {g->snap_taken, g->snap_done})
    if (e) (void)hipEventDestroy(e);
  if (g->copy_stream) (void)hipStreamDestroy(g->copy_stream);
  for (void *p : {(void *)g->snap_host, (void *)g->snap_ctl_host})
    if (p) (void)hipHostFree(p);
  for (void *p : {(void *)g->snap_dev, (void *)g->snap_ctl_dev})
    if (p) (void)hipFree(p);
  if (g->h_ctl) (void)hipHostFree(g->h_ctl);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return 0;
}

int ft_engine_info(ft_engine *g, int *n_nodes, int *n_elems, int *precision) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (n_nodes) *n_nodes = g->nn;
  if (n_elems) *n_elems = g->ne;
  if (precision) *precision = g->precision;
  return 0;
}

int ft_engine_layout(ft_engine *g, int *layout, int *ell_width) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (layout) *layout = g->layout == FT_LAYOUT_CORNER_PAIRS && !g->paired ? FT_LAYOUT_CORNER : g->layout;
  if (ell_width) *ell_width = g->width;
  return 0;
}

int ft_set_boundary(ft_engine *g, const int *dof_motion, int n_motions, const double *motions) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  drop_graphs(g);
  if (n_motions < 0) return fail(FT_ERR_ARG, "negative motion count");
  if (n_motions > 0 && !motions) return fail(FT_ERR_ARG, "motions is null");
  const size_t nd = 3 * (size_t)g->nn;
  std::vector<int> dm(nd, -1);
  for (size_t i = 0; i < nd; ++i) {
    int k = dof_motion ? dof_motion[i] : -1;
    if (k >= n_motions || k < -1) return fail(FT_ERR_ARG, "dof %zu: motion index %d out of range", i, k);
    dm[i] = k;
  }
  for (int k = 0; k < n_motions; ++k) {
    int type = (int)motions[FT_MOTION_STRIDE * k];
    if (type < 0 || type > 3) return fail(FT_ERR_ARG, "motion %d: unknown type %d", k, type);
  }
  g->h_dof_motion.assign(dm.begin(), dm.end());
  g->h_motions.assign(motions, motions + FT_MOTION_STRIDE * (size_t)n_motions);
  return rebuild_boundary(g);
}

int ft_set_instances(ft_engine *g, int n_inst, const int *node_offsets, const int *elem_offsets) {
  if (!g || !node_offsets || !elem_offsets) return fail(FT_ERR_ARG, "null argument");
  drop_graphs(g);
  if (n_inst < 1) return fail(FT_ERR_ARG, "at least one instance");
  if (node_offsets[0] != 0 || node_offsets[n_inst] != g->nn || elem_offsets[0] != 0 ||
      elem_offsets[n_inst] != g->ne)
    return fail(FT_ERR_ARG, "offsets must span all nodes and elements");
  for (int i = 0; i < n_inst; ++i)
    if (node_offsets[i + 1] <= node_offsets[i] || elem_offsets[i + 1] <= elem_offsets[i])
      return fail(FT_ERR_ARG, "instance %d is empty or the offsets decrease", i);
  for (int i = 0; i < n_inst; ++i)
    for (long e = elem_offsets[i]; e < elem_offsets[i + 1]; ++e)
      for (int k = 0; k < 8; ++k) {
        const int n = g->h_conn[8 * e + k];
        if (n < node_offsets[i] || n >= node_offsets[i + 1])
          return fail(FT_ERR_ARG, "element %ld of instance %d uses node %d of another instance", e, i, n);
      }
  // the step history and the injury criteria follow one instance's step
  if (n_inst > 1 && g->hist_cap > 0)
    return fail(FT_ERR_STATE, "the step history is a single-instance feature (ft_set_history 0 first)");
  if (n_inst > 1 && g->inj_on)
    return fail(FT_ERR_STATE, "the injury criteria are a single-instance feature (ft_injury_disable first)");
  if (n_inst > 1 && g->surr_on)
    return fail(FT_ERR_STATE, "the surrogate is a single-instance mode (ft_surrogate_enable 0 first)");
  HIPCHK(hipStreamSynchronize(g->stream));
  g->n_inst = n_inst;
  g->tmax_per_instance = false;  // end times belong to the previous instances
  g->inst_node.assign(node_offsets, node_offsets + n_inst + 1);
  g->inst_elem.assign(elem_offsets, elem_offsets + n_inst + 1);
  g->run_active = false;
  int rc;
  if ((rc = alloc_instances(g)) || (rc = build_tiles(g)) || (rc = rebuild_boundary(g))) return rc;
  return 0;
}

int ft_set_instance_end_times(ft_engine *g, const double *tmax) {
  if (!g || !tmax) return fail(FT_ERR_ARG, "null argument");
  g->h_tmax.assign(tmax, tmax + g->n_inst);
  g->tmax_per_instance = true;
  return 0;
}

int ft_reset_history(ft_engine *g) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (g->hist) {
    const size_t esz = g->precision == FT_FP32 ? 4 : 8;
    HIPCHK(hipMemsetAsync(g->hist, 0, hist_len(g->ne) * esz, g->stream));
    HIPCHK(hipStreamSynchronize(g->stream));
  }
  return 0;
}

int ft_set_state(ft_engine *g, const double *u, const double *v) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  const size_t nb = 3 * (size_t)g->nn * sizeof(double);
  int rc;
  if (u) {
    if ((rc = h2d_nodal(g, g->u, u))) return rc;
  } else {
    HIPCHK(hipMemsetAsync(g->u, 0, nb, g->stream));
  }
  if (v) {
    if ((rc = h2d_nodal(g, g->vs, v))) return rc;
  } else {
    HIPCHK(hipMemsetAsync(g->vs, 0, nb, g->stream));
  }
  HIPCHK(hipStreamSynchronize(g->stream));
  g->run_active = false;
  return 0;
}

int ft_assemble_lumped_mass(ft_engine *g, double *mass_out) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  double *me = nullptr;
  int rc = dalloc(&me, 8 * (size_t)g->ne);
  if (rc) return rc;
  hipLaunchKernelGGL(k_mass_elem, dim3(grid_for(g->ne)), dim3(kBlock), 0, g->stream, g->nn, g->ne, g->conn,
                     g->einfo, g->X, g->parts, me);
  hipLaunchKernelGGL(k_mass_gather, dim3(grid_for(g->nn)), dim3(kBlock), 0, g->stream, g->nn, g->ell,
                     g->width, me, g->mass);
  hipError_t e = hipGetLastError();
  (void)hipStreamSynchronize(g->stream);
  (void)hipFree(me);
  if (e != hipSuccess) return fail(FT_ERR_HIP, "mass kernels: %s", hipGetErrorString(e));
  g->mass_ready = true;
  if (mass_out) {
    std::vector<double> m(g->nn);
    HIPCHK(hipMemcpy(m.data(), g->mass, g->nn * sizeof(double), hipMemcpyDeviceToHost));
    for (int n = 0; n < g->nn; ++n) mass_out[3 * n] = mass_out[3 * n + 1] = mass_out[3 * n + 2] = m[n];
  }
  return 0;
}

int ft_stable_time_step(ft_engine *g, const double *u, const int *boundary, double failure_dt,
                        double *dt_out) {
  if (!g || !dt_out) return fail(FT_ERR_ARG, "null argument");
  int rc = upload_u(g, u);
  if (rc) return rc;
  if (boundary) {
    std::vector<unsigned char> mask(g->nn, 0);
    for (int n = 0; n < g->nn; ++n)
      for (int d = 0; d < 3; ++d)
        if (boundary[3 * n + d]) mask[n] |= (unsigned char)(1 << d);
    HIPCHK(hipMemcpyAsync(g->node_mask, mask.data(), g->nn, hipMemcpyHostToDevice, g->stream));
    hipLaunchKernelGGL(k_elem_flags, dim3((g->ne + 255) / 256), dim3(256), 0, g->stream, g->ne, g->conn,
                       g->node_mask, g->einfo);
    HIPCHK(hipGetLastError());
  }
  std::vector<unsigned long long> bits(g->n_inst, kInfBits);
  HIPCHK(hipMemcpyAsync(g->dtbits, bits.data(), bits.size() * sizeof(bits[0]), hipMemcpyHostToDevice, g->stream));
  if ((rc = element_pass(g, false, true, false, 0.0, g->dtbits, 0))) return rc;
  HIPCHK(hipMemcpyAsync(bits.data(), g->dtbits, bits.size() * sizeof(bits[0]), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  const unsigned long long bmin = *std::min_element(bits.begin(), bits.end());
  double dt;
  memcpy(&dt, &bmin, sizeof(dt));
  if (bmin == kInfBits) dt = 1e20;  // every element skipped: GlobalVariables.h huge
  *dt_out = dt;
  if (dt < failure_dt) return fail(FT_ERR_DT, "Timestep too small, dt = %15.9e", dt);
  return 0;
}

// fi = node sums of the element forces (GetForce_3D.cpp:39-44), into g->fi.
static int gather_forces(ft_engine *g) {
  {
    Timed t(g, 1);
    const bool w8 = g->width == 8;
    if (g->corner) {
      if (g->precision == FT_FP32)
        hipLaunchKernelGGL((k_gather<float, -1>), dim3(g->nodal_grid), dim3(kBlock), 0, g->stream, g->nn,
                           g->ell, g->width, (const float *)g->fe, g->fi,
                           Overflow<float>{(const float *)g->fe_ovf, g->ovf, g->n_ovf});
      else
        hipLaunchKernelGGL((k_gather<double, -1>), dim3(g->nodal_grid), dim3(kBlock), 0, g->stream, g->nn,
                           g->ell, g->width, (const double *)g->fe, g->fi,
                           Overflow<double>{(const double *)g->fe_ovf, g->ovf, g->n_ovf});
    } else if (g->precision == FT_FP32) {
      if (w8)
        hipLaunchKernelGGL((k_gather<float, 8>), dim3(g->nodal_grid), dim3(kBlock), 0, g->stream, g->nn,
                           g->ell, g->width, (const float *)g->fe, g->fi,
                           Overflow<float>{(const float *)g->fe_ovf, g->ovf, g->n_ovf});
      else
        hipLaunchKernelGGL((k_gather<float, 0>), dim3(g->nodal_grid), dim3(kBlock), 0, g->stream, g->nn,
                           g->ell, g->width, (const float *)g->fe, g->fi,
                           Overflow<float>{(const float *)g->fe_ovf, g->ovf, g->n_ovf});
    } else {
      if (w8)
        hipLaunchKernelGGL((k_gather<double, 8>), dim3(g->nodal_grid), dim3(kBlock), 0, g->stream, g->nn,
                           g->ell, g->width, (const double *)g->fe, g->fi,
                           Overflow<double>{(const double *)g->fe_ovf, g->ovf, g->n_ovf});
      else
        hipLaunchKernelGGL((k_gather<double, 0>), dim3(g->nodal_grid), dim3(kBlock), 0, g->stream, g->nn,
                           g->ell, g->width, (const double *)g->fe, g->fi,
                           Overflow<double>{(const double *)g->fe_ovf, g->ovf, g->n_ovf});
    }
    HIPCHK(hipGetLastError());
  }
  return 0;
}
int ft_element_forces(ft_engine *g, const double *u, double dt, int update_history, double *fe_out) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  int rc = upload_u(g, u);
  if (rc) return rc;
  Unpaired unpaired_pass(g);
  if ((rc = element_pass(g, true, false, false, dt, nullptr, update_history))) return rc;
  if ((rc = check_status(g))) return rc;
  if (fe_out) {
    const size_t n = 24 * (size_t)(g->corner ? g->nn : g->ne);
    std::vector<double> dev(n);
    if (g->precision == FT_FP32) {
      std::vector<float> tmp(n);
      HIPCHK(hipMemcpy(tmp.data(), g->fe, n * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) dev[i] = tmp[i];
    } else {
      HIPCHK(hipMemcpy(dev.data(), g->fe, n * sizeof(double), hipMemcpyDeviceToHost));
    }
    std::vector<double> ovf(3 * (size_t)std::max(g->width - 8, 0) * g->n_ovf);
    if (!ovf.empty()) {
      if (g->precision == FT_FP32) {
        std::vector<float> tmp(ovf.size());
        HIPCHK(hipMemcpy(tmp.data(), g->fe_ovf, tmp.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < ovf.size(); ++i) ovf[i] = tmp[i];
      } else {
        HIPCHK(hipMemcpy(ovf.data(), g->fe_ovf, ovf.size() * sizeof(double), hipMemcpyDeviceToHost));
      }
    }
    for (int e = 0; e < g->ne; ++e)
      for (int k = 0; k < 8; ++k) {
        const int node = g->h_conn[8 * (size_t)e + k];
        const size_t sk = g->h_eslot.empty() ? (size_t)k : (g->h_eslot[e] >> (4 * k)) & 15u;
        const double *src;
        if (sk >= 8)  // overflow slot of a node with more than 8 elements
          src = &ovf[3 * ((sk - 8) * g->n_ovf + (size_t)(g->h_ovf[node] >> 4))];
        else
          src = &dev[3 * (g->corner ? sk * g->nn + node : (size_t)k * g->ne + g->inv[e])];
        for (int d = 0; d < 3; ++d) fe_out[24 * (size_t)e + 3 * k + d] = src[d];
      }
  }
  return 0;
}

int ft_get_force(ft_engine *g, const double *u, double dt, double *fi_out) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  int rc = upload_u(g, u);
  if (rc) return rc;
  Unpaired unpaired_pass(g);
  if ((rc = element_pass(g, true, false, false, dt, nullptr, 1))) return rc;
  if ((rc = gather_forces(g))) return rc;
  if ((rc = check_status(g))) return rc;
  if (fi_out) return d2h_nodal(g, fi_out, g->fi);
  return 0;
}

int ft_get_force_fields(ft_engine *g, const double *u, double dt, double *fi_out, double *F_out,
                        double *detF_out, double *pk2_out) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  Unpaired unpaired_pass(g);
  const size_t ne = (size_t)g->ne;
  int rc;
  if (!g->gp_fields) {
    if ((rc = dalloc(&g->gp_fields, 128 * ne)) || (rc = dalloc(&g->perm_dev, ne))) return rc;
    HIPCHK(hipMemcpy(g->perm_dev, g->perm.data(), ne * sizeof(int), hipMemcpyHostToDevice));
  }
  if ((rc = upload_u(g, u))) return rc;
  {
    Timed t(g, 0);
    auto launch = [&](auto zero) -> int {
      using R = decltype(zero);
      ElemArgs<R> a = elem_args<R>(g, 0, g->n_etiles, false, dt, nullptr, 1);
      a.gp_F = g->gp_fields;
      a.gp_detF = g->gp_fields + 72 * ne;
      a.gp_pk2 = g->gp_fields + 80 * ne;
      a.perm = g->perm_dev;
      hipLaunchKernelGGL((k_element_gp<R>), dim3(grid_for_tiles(g->n_etiles)), dim3(kBlock), 0, g->stream, a);
      HIPCHK(hipGetLastError());
      return 0;
    };
    if ((rc = g->precision == FT_FP32 ? launch(0.0f) : launch(0.0))) return rc;
  }
  if ((rc = gather_forces(g))) return rc;
  if ((rc = check_status(g))) return rc;
  if (F_out) HIPCHK(hipMemcpy(F_out, g->gp_fields, 72 * ne * sizeof(double), hipMemcpyDeviceToHost));
  if (detF_out) HIPCHK(hipMemcpy(detF_out, g->gp_fields + 72 * ne, 8 * ne * sizeof(double), hipMemcpyDeviceToHost));
  if (pk2_out) HIPCHK(hipMemcpy(pk2_out, g->gp_fields + 80 * ne, 48 * ne * sizeof(double), hipMemcpyDeviceToHost));
  if (fi_out) return d2h_nodal(g, fi_out, g->fi);
  return 0;
}

int ft_run_init(ft_engine *g, double tmax, double dt_factor, double failure_dt) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  drop_graphs(g);
  if (!g->mass_ready) return fail(FT_ERR_STATE, "ft_assemble_lumped_mass must run before ft_run_init");
  if (!(dt_factor > 0)) return fail(FT_ERR_ARG, "dt_factor must be positive");
  g->tmax = tmax;
  g->factor = dt_factor;
  g->failure_dt = failure_dt;
  g->parity = 0;
  g->run_active = false;
  // Boundary motions at t = 0 (ex9.cpp:84) on the constrained DOFs.
  const size_t nd = 3 * (size_t)g->nn;
  if (g->n_motions > 0) {
    std::vector<double> X(nd), u(nd), v(nd);
    int rc;
    if ((rc = d2h_nodal(g, X.data(), g->X)) || (rc = d2h_nodal(g, u.data(), g->u)) ||
        (rc = d2h_nodal(g, v.data(), g->vs)))
      return rc;
    for (int n = 0; n < g->nn; ++n)
      for (int d = 0; d < 3; ++d) {
        int mi = g->h_dof_motion[3 * n + d];
        if (mi < 0) continue;
        double ub[3], vb[3], ab[3];
        motion_eval(&g->h_motions[FT_MOTION_STRIDE * mi], &X[3 * n], 0.0, ub, vb, ab);
        u[3 * n + d] = ub[d];
        v[3 * n + d] = vb[d];
      }
    if ((rc = h2d_nodal(g, g->u, u.data())) || (rc = h2d_nodal(g, g->vs, v.data()))) return rc;
  }
  HIPCHK(hipMemset(g->eacc, 0, kEaccStride * (size_t)std::max(1, g->n_ntiles) * sizeof(double)));
  HIPCHK(hipMemset(g->status, 0, sizeof(int)));
  if (!g->tmax_per_instance) g->h_tmax.assign(g->n_inst, tmax);
  HIPCHK(hipMemcpy(g->inst_tmax, g->h_tmax.data(), g->n_inst * sizeof(double), hipMemcpyHostToDevice));
  // dt_1 = factor * StableTimeStep() at u_0 (ex9.cpp:87), per instance
  int rc;
  std::vector<unsigned long long> bits(g->n_inst, kInfBits);
  HIPCHK(hipMemcpy(g->dtbits, bits.data(), bits.size() * sizeof(bits[0]), hipMemcpyHostToDevice));
  if ((rc = element_pass(g, false, true, false, 0.0, g->dtbits, 0))) return rc;
  HIPCHK(hipMemcpy(bits.data(), g->dtbits, bits.size() * sizeof(bits[0]), hipMemcpyDeviceToHost));
  std::vector<Ctl> c(g->n_inst);
  for (int i = 0; i < g->n_inst; ++i) {
    memset(&c[i], 0, sizeof(Ctl));
    double dtmin = 1e20;
    if (bits[i] != kInfBits) memcpy(&dtmin, &bits[i], sizeof(dtmin));
    if (dtmin < failure_dt)
      return fail(FT_ERR_DT, "Timestep too small, dt = %15.9e (instance %d)", dtmin, i);
    StepSlot &s0 = c[i].slot[0];
    s0.t = 0.0;
    s0.t_prev = 0.0;
    s0.dt = dt_factor * dtmin;
    memcpy(&s0.dtmin_bits, &dtmin, sizeof(dtmin));
    s0.dt_guide = dtmin;
    c[i].slot[1].dtmin_bits = kInfBits;
    for (StepSlot &sl : c[i].slot) {
      sl.cand_bits = kNoCandidate;
      sl.settled = 0;
      sl.far_bits = ~0u;
    }
  }
  HIPCHK(hipMemcpy(g->ctl, c.data(), c.size() * sizeof(Ctl), hipMemcpyHostToDevice));
  // Initial GetForce + CalculateAccelerations (ex9.cpp:89-92); the Prony
  // update of this pass uses dt_1 like the reference's global dt.
  if ((rc = element_pass(g, true, false, true, 0.0, nullptr, 1))) return rc;
  if ((rc = launch_pair_extras(g))) return rc;
  if ((rc = upload_caps(g, std::vector<int>(g->n_inst, 0)))) return rc;
  if ((rc = launch_nodal(g, kInitial))) return rc;  // stops: state of step 0
  if ((rc = check_status(g))) return rc;
  g->run_active = true;
  return 0;
}

// One step of every running instance: element pass (forces at u_k, the next
// stable step), optional injury update, nodal update.
int loop_body(ft_engine *g) {
  g->sample = (g->loop_iter++ % g->timing_every) == 0;
  int rc;
  if ((rc = element_pass(g, true, true, true, 0.0, nullptr, 1))) return rc;
  if ((rc = launch_pair_extras(g))) return rc;
  // u still holds this step's displacement here (the nodal kernel below
  // advances it): the injury update of ex5.cpp:240 for t_k
  if (g->inj_on && (rc = injury_launch(g, &g->ctl->slot[g->parity], 0.0, 0.0))) return rc;
  if ((rc = launch_nodal(g, kStep))) return rc;
  if (g->hist_cap > 0) {
    Timed t(g, 2);
    hipLaunchKernelGGL(k_history, dim3(1), dim3(kBlock), 0, g->stream, g->eacc, g->ntile_of_inst[1], g->ctl,
                       g->parity, g->hist_rows, g->hist_cap);
    HIPCHK(hipGetLastError());
  }
  g->sample = true;
  return 0;
}

// graph_steps loop bodies from the current parity, captured and
// instantiated once (the loop issues no memset and no host synchronisation).
int capture_chunk(ft_engine *g) {
  const int p0 = g->parity;
  HIPCHK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int k = 0; k < g->graph_steps && !rc; ++k) rc = loop_body(g);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(g->stream, &graph);
  g->parity = p0;
  if (rc || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc ? rc : fail(FT_ERR_HIP, "stream capture failed: %s", hipGetErrorString(e));
  }
  const hipError_t ei = hipGraphInstantiate(&g->gexec[p0], graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    g->gexec[p0] = nullptr;
    return fail(FT_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ei));
  }
  return 0;
}

// Early stop of the launch loop: every kDoneCheck steps the control slots of
// the next step are copied to pinned host memory behind the launches; a
// later check that finds the copy complete and every instance done (end
// time, step cap or no stable step) ends the loop, so a call asked for more
// steps than the run has left does not keep launching kernels that exit at
// once.  At most two check intervals of such launches are made.
constexpr int kDoneCheck = 128;

struct DoneWatch {
  ft_engine *g;
  bool pending = false;
  int parity = 0;
  int init() {
    if (!g->done_ev && hipEventCreateWithFlags(&g->done_ev, hipEventDisableTiming) != hipSuccess)
      return fail(FT_ERR_HIP, "event creation failed");
    if (g->h_ctl_n < g->n_inst) {
      if (g->h_ctl) (void)hipHostFree(g->h_ctl);
      g->h_ctl = nullptr;
      g->h_ctl_n = 0;
      if (hipHostMalloc((void **)&g->h_ctl, g->n_inst * sizeof(Ctl)) != hipSuccess)
        return fail(FT_ERR_ALLOC, "pinned control copy allocation failed");
      g->h_ctl_n = g->n_inst;
    }
    return 0;
  }
  // true when an earlier copy shows every instance done; then enqueues the
  // next copy
  bool all_done() {
    if (pending && hipEventQuery(g->done_ev) == hipSuccess) {
      pending = false;
      bool all = true;
      for (int i = 0; i < g->n_inst && all; ++i) all = g->h_ctl[i].slot[parity].done != 0;
      if (all) return true;
    }
    if (!pending &&
        hipMemcpyAsync(g->h_ctl, g->ctl, g->n_inst * sizeof(Ctl), hipMemcpyDeviceToHost, g->stream) == hipSuccess &&
        hipEventRecord(g->done_ev, g->stream) == hipSuccess) {
      pending = true;
      parity = g->parity;
    }
    return false;
  }
};

int ft_run_steps(ft_engine *g, int max_steps, int *steps_done) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (steps_done) *steps_done = 0;
  if (!g->run_active) return fail(FT_ERR_STATE, "ft_run_init must run first");
  if (max_steps < 0) return fail(FT_ERR_ARG, "max_steps < 0");
  std::vector<StepSlot> s;
  int rc;
  if ((rc = read_slots(g, g->parity, s))) return rc;  // all stopped between calls
  std::vector<int> before(g->n_inst), cap(g->n_inst);
  bool any = false;
  for (int i = 0; i < g->n_inst; ++i) {
    before[i] = s[i].step;
    const bool runs = max_steps > 0 && s[i].t < g->h_tmax[i] && !s[i].failed;
    cap[i] = runs ? (int)std::min<long>((long)before[i] + max_steps, INT32_MAX - 1) : before[i];
    any = any || runs;
  }
  if (!any) return 0;
  if ((rc = upload_caps(g, cap))) return rc;
  if ((rc = launch_nodal(g, kResume))) return rc;  // start step before + 1
  int k = 0;
  if (g->persistent && !g->timing && !g->surr_on && !g->inj_on) {
    if ((rc = launch_persistent(g, max_steps))) return rc;
    g->parity ^= max_steps & 1;
    k = max_steps;
  }
  DoneWatch watch{g};
  if (k < max_steps && (rc = watch.init())) return rc;
  int since = 0;
  bool stop = false;
  if (g->graph_steps > 0 && !g->timing && !g->surr_on) {
    for (; k + g->graph_steps <= max_steps; k += g->graph_steps) {
      if (!g->gexec[g->parity] && (rc = capture_chunk(g))) return rc;
      HIPCHK(hipGraphLaunch(g->gexec[g->parity], g->stream));
      g->parity ^= g->graph_steps & 1;
      if ((since += g->graph_steps) >= kDoneCheck) {
        since = 0;
        if ((stop = watch.all_done())) break;
      }
    }
  }
  for (; k < max_steps && !stop; ++k) {
    if ((rc = loop_body(g))) return rc;
    if (++since >= kDoneCheck) {
      since = 0;
      stop = watch.all_done();
    }
  }
  if ((rc = read_slots(g, g->parity, s))) return rc;
  int done = 0;
  for (int i = 0; i < g->n_inst; ++i) done = std::max(done, s[i].step - before[i]);
  if (steps_done) *steps_done = done;
  int st = 0;
  HIPCHK(hipMemcpy(&st, g->status, sizeof(int), hipMemcpyDeviceToHost));
  if (st & kStatusBarrier) return fail(FT_ERR_HIP, "persistent launch: a grid barrier wait ran out; the run is void");
  if (st & 8) return fail(FT_ERR_MATERIAL, "Unknown material type");
  // A batch keeps going when one of its instances stops without a stable
  // step (ft_get_instance reports it); a single run ends as the reference's
  // StableTimeStep does (code 19), or with 53 for non-finite forces.
  if (g->n_inst > 1) return 0;
  if (st & FT_STATUS_DT_COLLAPSE) {
    const double d = s[0].dt_done / g->factor;
    return fail(FT_ERR_DT, "Timestep too small, dt = %15.9e", d < 1e308 ? d : 0.0);
  }
  if (st & FT_STATUS_NONFINITE)
    return fail(FT_ERR_NUMERIC, "non-finite element force (an element inverted or a law left its domain)");
  return 0;
}

int ft_get_time(ft_engine *g, double *time, double *dt, int *step) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  StepSlot s;
  int rc = read_slot(g, g->parity, &s);
  if (rc) return rc;
  if (time) *time = s.t;
  if (dt) *dt = g->run_active ? s.dt_done : 0.0;
  if (step) *step = s.step;
  return 0;
}

// Component-major [3][nn] to node-major [nn][3].
__global__ void k_to_node_major(int nn, const double *soa, double *aos) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  aos[3 * (size_t)n] = soa[n];
  aos[3 * (size_t)n + 1] = soa[(size_t)nn + n];
  aos[3 * (size_t)n + 2] = soa[2 * (size_t)nn + n];
}

// A node-major copy of the current displacement (and the control slots) is
// made on the engine's stream behind the queued work, then moved to pinned
// host memory on a second stream while later calls compute.
int ft_snapshot_begin(ft_engine *g) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (g->snap_pending) return fail(FT_ERR_STATE, "a snapshot is pending: ft_snapshot_end first");
  HIPCHK(hipSetDevice(g->device));
  const size_t nb = 3 * (size_t)g->nn * sizeof(double);
  if (!g->snap_dev) {
    int rc = dalloc(&g->snap_dev, 3 * (size_t)g->nn);
    if (rc) return rc;
    if (hipHostMalloc((void **)&g->snap_host, nb) != hipSuccess) return fail(FT_ERR_ALLOC, "pinned snapshot allocation failed");
    HIPCHK(hipStreamCreate(&g->copy_stream));
    HIPCHK(hipEventCreateWithFlags(&g->snap_taken, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&g->snap_done, hipEventDisableTiming));
  }
  if (g->snap_ctl_n < g->n_inst) {
    if (g->snap_ctl_dev) (void)hipFree(g->snap_ctl_dev);
    if (g->snap_ctl_host) (void)hipHostFree(g->snap_ctl_host);
    g->snap_ctl_dev = nullptr;
    g->snap_ctl_host = nullptr;
    g->snap_ctl_n = 0;
    int rc = dalloc(&g->snap_ctl_dev, (size_t)g->n_inst);
    if (rc) return rc;
    if (hipHostMalloc((void **)&g->snap_ctl_host, g->n_inst * sizeof(Ctl)) != hipSuccess)
      return fail(FT_ERR_ALLOC, "pinned snapshot allocation failed");
    g->snap_ctl_n = g->n_inst;
  }
  const size_t cb = g->n_inst * sizeof(Ctl);
  hipLaunchKernelGGL(k_to_node_major, dim3((g->nn + 255) / 256), dim3(256), 0, g->stream, g->nn, g->u, g->snap_dev);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(g->snap_ctl_dev, g->ctl, cb, hipMemcpyDeviceToDevice, g->stream));
  HIPCHK(hipEventRecord(g->snap_taken, g->stream));
  HIPCHK(hipStreamWaitEvent(g->copy_stream, g->snap_taken, 0));
  HIPCHK(hipMemcpyAsync(g->snap_host, g->snap_dev, nb, hipMemcpyDeviceToHost, g->copy_stream));
  HIPCHK(hipMemcpyAsync(g->snap_ctl_host, g->snap_ctl_dev, cb, hipMemcpyDeviceToHost, g->copy_stream));
  HIPCHK(hipEventRecord(g->snap_done, g->copy_stream));
  g->snap_parity = g->parity;
  g->snap_pending = true;
  return 0;
}

const double *ft_snapshot_view(ft_engine *g) {
  if (!g || !g->snap_host || g->snap_pending) return nullptr;
  return g->snap_host;
}

int ft_snapshot_end(ft_engine *g, double *u, double *time, int *step) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (!g->snap_pending) return fail(FT_ERR_STATE, "no snapshot pending: ft_snapshot_begin first");
  HIPCHK(hipEventSynchronize(g->snap_done));
  g->snap_pending = false;
  if (u) memcpy(u, g->snap_host, 3 * (size_t)g->nn * sizeof(double));
  const StepSlot &s = g->snap_ctl_host[0].slot[g->snap_parity];
  if (time) *time = s.t;
  if (step) *step = s.step;
  return 0;
}

int ft_get_state(ft_engine *g, double *u, double *v, double *a, double *fi) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  const size_t nb = 3 * (size_t)g->nn * sizeof(double);
  HIPCHK(hipStreamSynchronize(g->stream));
  (void)nb;
  int rc0;
  if (u && (rc0 = d2h_nodal(g, u, g->u))) return rc0;
  if (v && (rc0 = d2h_nodal(g, v, g->vs))) return rc0;
  if (fi && (rc0 = d2h_nodal(g, fi, g->fi))) return rc0;
  if (a) {
    int rc;
    hipLaunchKernelGGL(k_state_out, dim3(std::max(1, g->n_ntiles)), dim3(kBlock), 0, g->stream, g->nn, g->ntiles,
                       g->n_ntiles, g->ctl, g->parity, g->fi, g->mass, g->X, g->node_code, g->motions, g->mrange,
                       g->tmp3);
    HIPCHK(hipGetLastError());
    if ((rc = d2h_nodal(g, a, g->tmp3))) return rc;
  }
  return 0;
}

namespace {
// Energy sums of instance i: its node tiles' accumulators in tile order.
int instance_energy(ft_engine *g, int i, double out4[4]) {
  const int t0 = g->ntile_of_inst[i], t1 = g->ntile_of_inst[i + 1];
  std::vector<double> e(kEaccStride * (size_t)(t1 - t0));
  HIPCHK(hipStreamSynchronize(g->stream));
  HIPCHK(hipMemcpy(e.data(), g->eacc + kEaccStride * (size_t)t0, e.size() * sizeof(double), hipMemcpyDeviceToHost));
  double w = 0, x = 0, k = 0;
  for (int b = 0; b < t1 - t0; ++b) {
    w += e[kEaccStride * b];
    x += e[kEaccStride * b + 1];
    k += e[kEaccStride * b + 2];
  }
  out4[0] = w;
  out4[1] = x;
  out4[2] = k;
  out4[3] = fabs(k + w - x);
  return 0;
}
}  // namespace

int ft_get_energy(ft_engine *g, double out4[4]) {
  if (!g || !out4) return fail(FT_ERR_ARG, "null argument");
  return instance_energy(g, 0, out4);
}

int ft_get_instance(ft_engine *g, int i, double out8[8]) {
  if (!g || !out8) return fail(FT_ERR_ARG, "null argument");
  if (i < 0 || i >= g->n_inst) return fail(FT_ERR_ARG, "instance %d out of range", i);
  std::vector<Ctl> c(1);
  HIPCHK(hipStreamSynchronize(g->stream));
  HIPCHK(hipMemcpy(c.data(), g->ctl + i, sizeof(Ctl), hipMemcpyDeviceToHost));
  const StepSlot &s = c[0].slot[g->parity];
  out8[0] = s.t;
  out8[1] = g->run_active ? s.dt_done : 0.0;
  out8[2] = s.step;
  out8[3] = s.failed ? 2.0 : ((s.t >= g->h_tmax[i]) ? 1.0 : 0.0);
  return instance_energy(g, i, out8 + 4);
}

int ft_set_history(ft_engine *g, int capacity) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  drop_graphs(g);
  if (capacity < 0) return fail(FT_ERR_ARG, "capacity < 0");
  if (capacity > 0 && g->n_inst != 1) return fail(FT_ERR_STATE, "the step history is a single-instance feature");
  HIPCHK(hipStreamSynchronize(g->stream));
  if (g->hist_rows) HIPCHK(hipFree(g->hist_rows));
  g->hist_rows = nullptr;
  g->hist_cap = 0;
  if (capacity > 0) {
    int rc = dalloc(&g->hist_rows, 6 * (size_t)capacity);
    if (rc) return rc;
    g->hist_cap = capacity;
  }
  int zero = 0;
  HIPCHK(hipMemcpy(&g->ctl->hist_count, &zero, sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

int ft_get_history(ft_engine *g, double *time, double *dt, double *energy, int *n) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  HIPCHK(hipStreamSynchronize(g->stream));
  int cnt = 0;
  HIPCHK(hipMemcpy(&cnt, &g->ctl->hist_count, sizeof(int), hipMemcpyDeviceToHost));
  cnt = std::min(cnt, g->hist_cap);
  if (n) *n = cnt;
  if (cnt == 0) return 0;
  std::vector<double> rows(6 * (size_t)cnt);
  HIPCHK(hipMemcpy(rows.data(), g->hist_rows, rows.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int i = 0; i < cnt; ++i) {
    if (time) time[i] = rows[6 * i];
    if (dt) dt[i] = rows[6 * i + 1];
    if (energy)
      for (int k = 0; k < 4; ++k) energy[4 * i + k] = rows[6 * i + 2 + k];
  }
  return 0;
}

int ft_get_status(ft_engine *g, int *status) {
  if (!g || !status) return fail(FT_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(g->stream));
  HIPCHK(hipMemcpy(status, g->status, sizeof(int), hipMemcpyDeviceToHost));
  return 0;
}

int ft_get_visco_history(ft_engine *g, double *hist) {
  if (!g || !hist) return fail(FT_ERR_ARG, "null argument");
  const size_t n = 144 * (size_t)g->ne, len = hist_len(g->ne);
  if (!g->hist) {
    memset(hist, 0, n * sizeof(double));
    return 0;
  }
  HIPCHK(hipStreamSynchronize(g->stream));
  std::vector<double> cm(len);
  if (g->precision == FT_FP32) {
    std::vector<float> t(len);
    HIPCHK(hipMemcpy(t.data(), g->hist, len * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < len; ++i) cm[i] = t[i];
  } else {
    HIPCHK(hipMemcpy(cm.data(), g->hist, len * sizeof(double), hipMemcpyDeviceToHost));
  }
  // 64-element blocks, internal order -> element-major original order
  for (size_t e = 0; e < (size_t)g->ne; ++e)
    for (int c = 0; c < 144; ++c) hist[e * 144 + c] = cm[hist_index(c, g->inv[e])];
  return 0;
}

int ft_set_visco_history(ft_engine *g, const double *hist) {
  if (!g || !hist) return fail(FT_ERR_ARG, "null argument");
  if (!g->hist) return 0;  // no Prony law in the mesh
  drop_graphs(g);
  const size_t len = hist_len(g->ne);
  std::vector<double> cm(len, 0.0);
  for (size_t e = 0; e < (size_t)g->ne; ++e)
    for (int c = 0; c < 144; ++c) cm[hist_index(c, g->inv[e])] = hist[e * 144 + c];
  HIPCHK(hipStreamSynchronize(g->stream));
  if (g->precision == FT_FP32) {
    std::vector<float> t(cm.begin(), cm.end());
    HIPCHK(hipMemcpy(g->hist, t.data(), len * sizeof(float), hipMemcpyHostToDevice));
  } else {
    HIPCHK(hipMemcpy(g->hist, cm.data(), len * sizeof(double), hipMemcpyHostToDevice));
  }
  return 0;
}

// ---- constitutive surrogate ----------------------------------------------
int ft_surrogate_reserve(ft_engine *g, long capacity) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (capacity < 0) return fail(FT_ERR_ARG, "capacity < 0");
  for (void *p : {(void *)g->ds_x, (void *)g->ds_t})
    if (p) (void)hipFree(p);
  g->ds_x = g->ds_t = nullptr;
  g->ds_n = g->ds_cap = 0;
  if (capacity == 0) return 0;
  int rc;
  if ((rc = dalloc(&g->ds_x, (size_t)capacity * kSurrFeat)) || (rc = dalloc(&g->ds_t, (size_t)capacity * kSurrOut)))
    return rc;
  g->ds_cap = capacity;
  return 0;
}

int ft_surrogate_collect(ft_engine *g, long *total) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (!g->ds_cap) return fail(FT_ERR_STATE, "ft_surrogate_reserve must run first");
  const float *fibf = g->precision == FT_FP32 ? reinterpret_cast<const float *>(g->fib) : nullptr;
  const double *fibd = g->precision == FT_FP64 ? reinterpret_cast<const double *>(g->fib) : nullptr;
  for (const auto &sg : g->segs) {
    if (!is_surr_mat(sg.mat)) continue;
    const int n = sg.end - sg.begin;
    if (g->ds_n + 8L * n > g->ds_cap) break;
    const dim3 grid((n + kBlock - 1) / kBlock);
    if (sg.mat == 7)
      hipLaunchKernelGGL(k_surr_sample<7>, grid, dim3(kBlock), 0, g->stream, g->nn, g->ne, g->conn, g->X, g->u, fibf,
                         fibd, g->parts, g->einfo, sg.begin, sg.end, g->ds_n, g->ds_x, g->ds_t);
    else
      hipLaunchKernelGGL(k_surr_sample<8>, grid, dim3(kBlock), 0, g->stream, g->nn, g->ne, g->conn, g->X, g->u, fibf,
                         fibd, g->parts, g->einfo, sg.begin, sg.end, g->ds_n, g->ds_x, g->ds_t);
    HIPCHK(hipGetLastError());
    g->ds_n += 8L * n;
  }
  HIPCHK(hipStreamSynchronize(g->stream));
  if (total) *total = g->ds_n;
  return 0;
}

namespace {

// Normalisation block from the collected samples: feature and output mean
// and spread, and the training box (per-feature range of the raw features).
int surr_statistics(ft_engine *g, std::vector<float> &norm) {
  const long n = g->ds_n;
  const long len = 1L << 16;
  const int nch = (int)((n + len - 1) / len);
  double *part = nullptr;
  int rc;
  if ((rc = dalloc(&part, (size_t)nch * (kSurrFeat + kSurrOut) * 4))) return rc;
  hipLaunchKernelGGL(k_col_stats, dim3(kSurrFeat, nch), dim3(kBlock), 0, g->stream, g->ds_x, n, kSurrFeat, len, part);
  hipLaunchKernelGGL(k_col_stats, dim3(kSurrOut, nch), dim3(kBlock), 0, g->stream, g->ds_t, n, kSurrOut, len,
                     part + (size_t)nch * kSurrFeat * 4);
  std::vector<double> h((size_t)nch * (kSurrFeat + kSurrOut) * 4);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(part);
  if (e != hipSuccess) return fail(FT_ERR_HIP, "statistics failed: %s", hipGetErrorString(e));
  norm.assign(kSurrNorm, 0.0f);
  auto column = [&](const double *base, int w, int col, double &mean, double &sd, double &lo, double &hi) {
    double s = 0, q = 0;
    lo = 1e300;
    hi = -1e300;
    for (int c = 0; c < nch; ++c) {
      const double *o = base + ((size_t)c * w + col) * 4;
      s += o[0];
      q += o[1];
      lo = std::min(lo, o[2]);
      hi = std::max(hi, o[3]);
    }
    mean = s / n;
    sd = sqrt(std::max(q / n - mean * mean, 0.0));
  };
  for (int k = 0; k < kSurrFeat; ++k) {
    double mean, sd, lo, hi;
    column(h.data(), kSurrFeat, k, mean, sd, lo, hi);
    const bool used = k < 12 && sd > 0;
    norm[k] = used ? (float)mean : 0.0f;
    norm[16 + k] = used ? (float)sd : 1.0f;
    norm[32 + k] = (float)lo;
    norm[48 + k] = (float)hi;
  }
  for (int k = 0; k < kSurrOut; ++k) {
    double mean, sd, lo, hi;
    column(h.data() + (size_t)nch * kSurrFeat * 4, kSurrOut, k, mean, sd, lo, hi);
    norm[64 + k] = k < 6 ? (float)mean : 0.0f;
    norm[72 + k] = (k < 6 && sd > 0) ? (float)sd : 1.0f;
  }
  return 0;
}

int surr_upload_norm(ft_engine *g) {
  if (!g->surr_norm) {
    int rc = dalloc(&g->surr_norm, kSurrNorm);
    if (rc) return rc;
  }
  HIPCHK(hipMemcpy(g->surr_norm, g->h_surr_norm.data(), kSurrNorm * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

int ft_surrogate_train(ft_engine *g, int steps, int batch, double lr, unsigned long long seed, double out[4]) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (g->ds_n < 1024) return fail(FT_ERR_STATE, "too few collected samples (ft_surrogate_collect)");
  if (steps < 0 || batch < 1) return fail(FT_ERR_ARG, "steps < 0 or batch < 1");
  drop_graphs(g);
  int rc;
  // targets per unit strain magnitude (the samples keep S_iso; a copy)
  float *ts_raw = g->ds_t;
  float *ts_per = nullptr;
  if ((rc = dalloc(&ts_per, (size_t)g->ds_n * kSurrOut))) return rc;
  HIPCHK(hipMemcpy(ts_per, ts_raw, (size_t)g->ds_n * kSurrOut * sizeof(float), hipMemcpyDeviceToDevice));
  hipLaunchKernelGGL(k_surr_per_strain, dim3((unsigned)((g->ds_n + kBlock - 1) / kBlock)), dim3(kBlock), 0, g->stream,
                     g->ds_x, ts_per, g->ds_n);
  g->ds_t = ts_per;
  std::vector<float> norm;
  rc = surr_statistics(g, norm);
  if (rc) {
    g->ds_t = ts_raw;
    (void)hipFree(ts_per);
    return rc;
  }
  g->h_surr_norm = norm;
  if ((rc = surr_upload_norm(g))) return rc;
  const long n = g->ds_n;
  __bf16 *xb = nullptr;
  float *tb = nullptr;
  if ((rc = dalloc(&xb, (size_t)n * kSurrFeat)) || (rc = dalloc(&tb, (size_t)n * kSurrOut))) {
    if (xb) (void)hipFree(xb);
    return rc;
  }
  hipLaunchKernelGGL(k_surr_normalise, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, g->stream,
                     g->ds_x, g->ds_t, n, g->surr_norm, xb, tb);
  (void)hipStreamSynchronize(g->stream);
  g->ds_t = ts_raw;
  (void)hipFree(ts_per);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) rc = fail(FT_ERR_HIP, "normalisation failed: %s", hipGetErrorString(e));
  if (!rc && !g->mlp) rc = ft::mlp_create(seed, &g->mlp);
  double loss = 0.0;
  if (!rc) rc = ft::mlp_train(g->mlp, xb, tb, n, steps, batch, (float)lr, seed, g->stream, &loss);
  // error of the trained network on (up to) 2^16 samples spread over the set
  double e2 = 0, r2 = 0;
  const long m = std::min(n, 65536L), stride = n / m;
  __bf16 *vx = nullptr;
  float *vt = nullptr, *vy = nullptr, *vs = nullptr;
  if (!rc && !(rc = dalloc(&vx, (size_t)m * kSurrFeat)) && !(rc = dalloc(&vt, (size_t)m * kSurrOut)) &&
      !(rc = dalloc(&vy, (size_t)m * kSurrOut)) && !(rc = dalloc(&vs, (size_t)m))) {
    hipLaunchKernelGGL(k_gather_stride, dim3((unsigned)((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, g->stream, xb,
                       tb, g->ds_x, stride, m, vx, vt, vs);
    rc = ft::mlp_infer(g->mlp, vx, m, vy, g->stream);
    std::vector<float> ht((size_t)m * kSurrOut), hy((size_t)m * kSurrOut), hs((size_t)m);
    if (!rc && (hipMemcpyAsync(ht.data(), vt, ht.size() * sizeof(float), hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
                hipMemcpyAsync(hy.data(), vy, hy.size() * sizeof(float), hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
                hipMemcpyAsync(hs.data(), vs, hs.size() * sizeof(float), hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
                hipStreamSynchronize(g->stream) != hipSuccess))
      rc = fail(FT_ERR_HIP, "validation failed");
    for (long i = 0; i < m && !rc; ++i)
      for (int k = 0; k < 6; ++k) {
        const double w = k < 3 ? 1.0 : 2.0, sc = norm[72 + k], mu = norm[64 + k];
        const double yv = (hy[i * kSurrOut + k] * sc + mu) * hs[i], tv = (ht[i * kSurrOut + k] * sc + mu) * hs[i];
        e2 += w * (yv - tv) * (yv - tv);
        r2 += w * tv * tv;
      }
  }
  for (void *p : {(void *)vx, (void *)vt, (void *)vy, (void *)vs})
    if (p) (void)hipFree(p);
  (void)hipStreamSynchronize(g->stream);
  (void)hipFree(xb);
  (void)hipFree(tb);
  if (rc) return rc;
  g->surr_ready = true;
  if (out) {
    out[0] = loss;
    out[1] = (double)n;
    out[2] = r2 > 0 ? sqrt(e2 / r2) : 0.0;
    out[3] = steps;
  }
  return 0;
}

int ft_surrogate_samples(ft_engine *g, long first, long count, float *x, float *t) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (first < 0 || count < 0 || first + count > g->ds_n) return fail(FT_ERR_ARG, "sample range outside the set");
  if (count == 0) return 0;
  HIPCHK(hipStreamSynchronize(g->stream));
  if (x) HIPCHK(hipMemcpy(x, g->ds_x + first * kSurrFeat, count * kSurrFeat * sizeof(float), hipMemcpyDeviceToHost));
  if (t) HIPCHK(hipMemcpy(t, g->ds_t + first * kSurrOut, count * kSurrOut * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int ft_surrogate_get(ft_engine *g, float *params, float *norm) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (!g->surr_ready) return fail(FT_ERR_STATE, "no trained surrogate");
  if (norm) memcpy(norm, g->h_surr_norm.data(), kSurrNorm * sizeof(float));
  return params ? ft::mlp_get_params(g->mlp, params) : 0;
}

int ft_surrogate_set(ft_engine *g, const float *params, const float *norm) {
  if (!g || !params || !norm) return fail(FT_ERR_ARG, "null argument");
  drop_graphs(g);
  int rc;
  if (!g->mlp && (rc = ft::mlp_create(1, &g->mlp))) return rc;
  if ((rc = ft::mlp_set_params(g->mlp, params))) return rc;
  g->h_surr_norm.assign(norm, norm + kSurrNorm);
  if ((rc = surr_upload_norm(g))) return rc;
  g->surr_ready = true;
  return 0;
}

int ft_surrogate_enable(ft_engine *g, int enable, int audit_every, int audit_permille) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  drop_graphs(g);
  if (!enable) {
    g->surr_on = false;
    return 0;
  }
  if (!g->surr_ready) return fail(FT_ERR_STATE, "no trained surrogate (ft_surrogate_train or ft_surrogate_set)");
  if (g->precision != FT_FP32) return fail(FT_ERR_STATE, "the surrogate runs in the FP32 engine");
  if (g->n_inst != 1) return fail(FT_ERR_STATE, "the surrogate is a single-instance mode");
  if (audit_every < 1 || audit_permille < 0 || audit_permille > 1000) return fail(FT_ERR_ARG, "bad audit schedule");
  long gps = 0;
  for (const auto &sg : g->segs)
    if (is_surr_mat(sg.mat)) gps += 8L * (sg.end - sg.begin);
  if (gps == 0) return fail(FT_ERR_STATE, "no element of material 7 or 8");
  if (gps != g->surr_gps) {
    for (void *p : {(void *)g->surr_feat, (void *)g->surr_y})
      if (p) (void)hipFree(p);
    g->surr_feat = nullptr;
    g->surr_y = nullptr;
    int rc;
    if ((rc = dalloc(&g->surr_feat, (size_t)gps * kSurrFeat)) || (rc = dalloc(&g->surr_y, (size_t)gps * kSurrOut)))
      return rc;
    g->surr_gps = gps;
  }
  if (g->surr_tally) (void)hipFree(g->surr_tally);
  g->surr_tally = nullptr;
  int rc;
  if ((rc = dalloc(&g->surr_tally, (size_t)std::max(1, g->n_etiles) * 6))) return rc;
  HIPCHK(hipMemset(g->surr_tally, 0, (size_t)std::max(1, g->n_etiles) * 6 * sizeof(double)));
  g->surr_audit_every = audit_every;
  g->surr_permille = audit_permille;
  g->surr_iter = 0;
  g->surr_on = true;
  return 0;
}

int ft_surrogate_stats(ft_engine *g, double out[6]) {
  if (!g || !out) return fail(FT_ERR_ARG, "null argument");
  for (int k = 0; k < 6; ++k) out[k] = 0.0;
  if (!g->surr_tally) return 0;
  std::vector<double> t((size_t)std::max(1, g->n_etiles) * 6);
  HIPCHK(hipStreamSynchronize(g->stream));
  HIPCHK(hipMemcpy(t.data(), g->surr_tally, t.size() * sizeof(double), hipMemcpyDeviceToHost));
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (size_t i = 0; i < t.size() / 6; ++i)
    for (int k = 0; k < 6; ++k) v[k] += t[i * 6 + k];
  out[0] = v[1] > 0 ? sqrt(v[0] / v[1]) : 0.0;  // audited relative error of S_iso
  out[1] = v[2] > 0 ? sqrt(v[0] / v[2]) : 0.0;  // the same relative to the full PK2 stress
  out[2] = v[4] > 0 ? v[3] / v[4] : 0.0;        // fallback share of the Gauss-point evaluations
  out[3] = v[5];                                // audited Gauss-point evaluations
  out[4] = v[4];                                // surrogate-segment Gauss-point evaluations
  out[5] = v[3];                                // evaluations that fell back to the analytic law
  return 0;
}

int ft_set_persistent(ft_engine *g, int enable) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  drop_graphs(g);
  g->persistent = enable != 0;
  return 0;
}

int ft_set_graph(ft_engine *g, int steps) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (steps < 0) return fail(FT_ERR_ARG, "steps < 0");
  drop_graphs(g);
  g->graph_steps = steps;
  return 0;
}

// Element-kernel launch shapes by measurement on this mesh: per material
// segment, every tile size (64, 128, 256 elements) and, for the FP32 kernel,
// every occupancy variant (1..5 waves per SIMD) runs `repeats` timed
// launches of the run loop's force + dt kernel; the fastest is kept.  The
// launches write only the element forces (rewritten by every step), a
// scratch minimum and a scratch copy of the Prony history (updated as in the
// loop, so its traffic is part of the timing); the status word is restored,
// so tuning between run calls leaves the run unchanged.
int ft_autotune(ft_engine *g, int repeats, double *report, int *n_segments) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (repeats < 1) return fail(FT_ERR_ARG, "repeats < 1");
  drop_graphs(g);
  int rc;
  if ((rc = collect_timing(g))) return rc;
  const bool timing = g->timing;
  g->timing = false;
  int status0 = 0;
  HIPCHK(hipMemcpy(&status0, g->status, sizeof(int), hipMemcpyDeviceToHost));
  const bool fp32 = g->precision == FT_FP32;
  // the history the timed launches update: a copy of the engine's
  void *hist_run = g->hist, *hist_copy = nullptr;
  if (hist_run) {
    const size_t bytes = hist_len(g->ne) * (fp32 ? 4 : 8);
    if ((rc = dalloc((char **)&hist_copy, bytes))) {
      g->timing = timing;
      return rc;
    }
    if (hipMemcpy(hist_copy, hist_run, bytes, hipMemcpyDeviceToDevice) != hipSuccess) {
      (void)hipFree(hist_copy);
      g->timing = timing;
      return fail(FT_ERR_HIP, "history copy failed");
    }
    g->hist = hist_copy;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(FT_ERR_HIP, "event creation failed");
  auto time_one = [&](size_t si, int waves, double &us) -> int {
    const ft_engine::Seg sg = g->segs[si];
    auto launch = [&]() {
      return fp32 ? force_segment<float, true>(g, sg, waves, false, 1e-7, g->dtbits, 1)
                  : force_segment<double, true>(g, sg, waves, false, 1e-7, g->dtbits, 1);
    };
    int r;
    if ((r = launch())) return r;  // warm-up
    HIPCHK(hipEventRecord(e0, g->stream));
    for (int k = 0; k < repeats; ++k)
      if ((r = launch())) return r;
    HIPCHK(hipEventRecord(e1, g->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    us = 1e3 * ms / repeats;
    return 0;
  };
  const int tiles[3] = {64, 128, 256};
  const size_t ns = g->segs.size();
  std::vector<double> rows(6 * ns, 0.0);
  for (size_t si = 0; si < ns && !rc; ++si) {
    auto &sg = g->segs[si];
    double best = 1e300, t_def = 0;
    int bw = 0, bt = kBlock;
    for (int tile : tiles) {
      sg.tile = tile;
      if ((rc = build_element_tiles(g))) break;
      if (tile == kBlock) {
        // The default shape (launch bounds of the compiler's choosing) is a
        // candidate like the others.
        if ((rc = time_one(si, 0, t_def))) break;
        if (t_def < best) {
          best = t_def;
          bw = 0;
          bt = tile;
        }
      }
      for (int w = fp32 ? 1 : 0; w <= (fp32 ? kMaxTuneWaves : 0) && !rc; ++w) {
        // a variant that spills more than a few registers to scratch memory
        // is not a candidate (kilobyte spills ran 20x slower; a few dozen
        // bytes can still win through the higher occupancy)
        if (fp32 && elem_scratch(sg.mat, w) > kTuneSpillMax) continue;
        double us = 0;
        if ((rc = time_one(si, w, us))) break;
        if (us < best) {
          best = us;
          bw = w;
          bt = tile;
        }
      }
      if (rc) break;
    }
    sg.tile = bt;
    sg.waves = bw;
    double *r = &rows[6 * si];
    r[0] = sg.mat;
    r[1] = sg.end - sg.begin;
    r[2] = bw;
    r[3] = bt;
    r[4] = best;
    r[5] = t_def;
  }
  if (!rc) rc = build_element_tiles(g);
  if (hist_copy) {
    (void)hipDeviceSynchronize();
    (void)hipFree(hist_copy);
    g->hist = hist_run;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  g->timing = timing;
  if (rc) return rc;
  HIPCHK(hipMemcpy(g->status, &status0, sizeof(int), hipMemcpyHostToDevice));
  if (report) memcpy(report, rows.data(), rows.size() * sizeof(double));
  if (n_segments) *n_segments = (int)ns;
  return 0;
}

int ft_get_segments(ft_engine *g, int *n_segments, int *desc) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (n_segments) *n_segments = (int)g->segs.size();
  if (desc)
    for (size_t i = 0; i < g->segs.size(); ++i) {
      const auto &sg = g->segs[i];
      int *d = desc + 5 * i;
      d[0] = sg.mat;
      d[1] = sg.begin;
      d[2] = sg.end;
      d[3] = sg.waves;
      d[4] = sg.tile;
    }
  return 0;
}

int ft_set_segment_shape(ft_engine *g, int segment, int waves, int tile) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (segment < 0 || segment >= (int)g->segs.size()) return fail(FT_ERR_ARG, "segment %d out of range", segment);
  if (waves < 0 || waves > kMaxTuneWaves || (waves > 0 && g->precision != FT_FP32))
    return fail(FT_ERR_ARG, "occupancy variant %d not available", waves);
  if (tile != 64 && tile != 128 && tile != 256) return fail(FT_ERR_ARG, "tile size must be 64, 128 or 256");
  drop_graphs(g);
  g->segs[segment].waves = waves;
  g->segs[segment].tile = tile;
  return build_element_tiles(g);
}

int ft_set_timing(ft_engine *g, int enable) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  int rc = collect_timing(g);
  if (rc) return rc;
  g->timing = enable != 0;
  g->timing_every = std::max(1, enable);
  g->loop_iter = 0;
  for (int c = 0; c < kTimingClasses; ++c) {
    g->ms[c] = 0;
    g->launches[c] = 0;
  }
  return 0;
}

int ft_get_timing(ft_engine *g, double ms[3], int launches[3]) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  int rc = collect_timing(g);
  if (rc) return rc;
  for (int c = 0; c < kTimingClasses; ++c) {
    if (ms) ms[c] = g->ms[c];
    if (launches) launches[c] = g->launches[c];
  }
  return 0;
}

int ft_element_strain(ft_engine *g, const double *u, double *eavg, double *principal) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  int rc = upload_u(g, u);
  if (rc) return rc;
  double *de = nullptr, *dp = nullptr;
  if ((rc = dalloc(&de, 9 * (size_t)g->ne)) || (rc = dalloc(&dp, 3 * (size_t)g->ne))) {
    if (de) (void)hipFree(de);
    return rc;
  }
  hipLaunchKernelGGL(k_strain, dim3((g->ne + kBlock - 1) / kBlock), dim3(kBlock), 0, g->stream, g->nn, g->ne,
                     g->conn, g->X, g->u, de, dp);
  hipError_t err = hipGetLastError();
  std::vector<double> he(9 * (size_t)g->ne), hp(3 * (size_t)g->ne);
  if (err == hipSuccess) err = hipStreamSynchronize(g->stream);
  if (err == hipSuccess) err = hipMemcpy(he.data(), de, he.size() * sizeof(double), hipMemcpyDeviceToHost);
  if (err == hipSuccess) err = hipMemcpy(hp.data(), dp, hp.size() * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(de);
  (void)hipFree(dp);
  if (err != hipSuccess) return fail(FT_ERR_HIP, "strain kernel: %s", hipGetErrorString(err));
  for (int e = 0; e < g->ne; ++e) {
    const size_t i = g->inv[e];
    if (eavg) memcpy(eavg + 9 * (size_t)e, &he[9 * i], 9 * sizeof(double));
    if (principal) memcpy(principal + 3 * (size_t)e, &hp[3 * i], 3 * sizeof(double));
  }
  return 0;
}

int ft_accelerations(ft_engine *g, const double *f_net, const int *boundary, double *a) {
  if (!g || !f_net || !boundary || !a) return fail(FT_ERR_ARG, "null argument");
  if (!g->mass_ready) return fail(FT_ERR_STATE, "lumped mass not assembled");
  const size_t nd = 3 * (size_t)g->nn;
  int rc = aux_reserve(g, 3 * nd);
  if (rc) return rc;
  double *df = g->aux, *da = g->aux + nd;
  int *db = (int *)(g->aux + 2 * nd);
  HIPCHK(hipMemcpyAsync(df, f_net, nd * sizeof(double), hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipMemcpyAsync(da, a, nd * sizeof(double), hipMemcpyHostToDevice, g->stream));
  HIPCHK(hipMemcpyAsync(db, boundary, nd * sizeof(int), hipMemcpyHostToDevice, g->stream));
  hipLaunchKernelGGL(k_accel, dim3(grid_for((long)nd)), dim3(kBlock), 0, g->stream, (int)nd, df, db, g->mass,
                     da);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(a, da, nd * sizeof(double), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

int ft_energy_terms(ft_engine *g, const double *u, const double *u_prev, const double *v,
                    const double *a, const double *a_prev, const double *fi, const double *fi_prev,
                    const double *fe, const double *fe_prev, const int *boundary, double out3[3]) {
  const double *in[] = {u, u_prev, v, a, a_prev, fi, fi_prev, fe, fe_prev};
  if (!g || !boundary || !out3) return fail(FT_ERR_ARG, "null argument");
  for (const double *p : in)
    if (!p) return fail(FT_ERR_ARG, "null argument");
  if (!g->mass_ready) return fail(FT_ERR_STATE, "lumped mass not assembled");
  const size_t nd = 3 * (size_t)g->nn;
  int rc = aux_reserve(g, 10 * nd + 3 * kEnergyBlocks + 3);
  if (rc) return rc;
  double *d[9];
  for (int k = 0; k < 9; ++k) {
    d[k] = g->aux + k * nd;
    HIPCHK(hipMemcpyAsync(d[k], in[k], nd * sizeof(double), hipMemcpyHostToDevice, g->stream));
  }
  int *db = (int *)(g->aux + 9 * nd);
  double *part = g->aux + 10 * nd, *res = part + 3 * kEnergyBlocks;
  HIPCHK(hipMemcpyAsync(db, boundary, nd * sizeof(int), hipMemcpyHostToDevice, g->stream));
  EnergyIn ei{d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], g->mass, db};
  hipLaunchKernelGGL(k_energy_terms, dim3(kEnergyBlocks), dim3(kBlock), 0, g->stream, (int)nd, ei, part);
  hipLaunchKernelGGL(k_sum3, dim3(1), dim3(64), 0, g->stream, kEnergyBlocks, part, res);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out3, res, 3 * sizeof(double), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

// ft_energy_terms on device arrays, with no allocation: the partial sums go
// to this GPU's copy of the static `scratch` (every call overwrites all of
// it, so nothing needs clearing); calls that overlap on one GPU share it and
// race.  ndof must be 3 * nodes of the engine.
int ft_energy_terms_noalloc(ft_engine *g, long ndof, const double *u, const double *u_prev,
                            const double *v, const double *a, const double *a_prev, const double *fi,
                            const double *fi_prev, const double *fe, const double *fe_prev,
                            const int *boundary, double out3[3]) {
  const double *in[] = {u, u_prev, v, a, a_prev, fi, fi_prev, fe, fe_prev};
  if (!g || !boundary || !out3) return fail(FT_ERR_ARG, "null argument");
  for (const double *p : in)
    if (!p) return fail(FT_ERR_ARG, "null argument");
  if (ndof != 3L * g->nn) return fail(FT_ERR_ARG, "ndof %ld is not 3 x %d nodes", ndof, g->nn);
  if (!g->mass_ready) return fail(FT_ERR_STATE, "lumped mass not assembled");
  if (!g->scratch) return fail(FT_ERR_STATE, "scratch address not set");
  // The partials of the fixed kEnergyBlocks-block grid and the totals fill
  // the scratch exactly, whatever the mesh size (the kernel strides).
  static_assert(3 * (size_t)kEnergyBlocks + 3 <= kScratchDoubles, "scratch too small");
  double *part = g->scratch, *res = g->scratch + 3 * kEnergyBlocks;
  EnergyIn ei{u, u_prev, v, a, a_prev, fi, fi_prev, fe, fe_prev, g->mass, boundary};
  hipLaunchKernelGGL(k_energy_terms, dim3(kEnergyBlocks), dim3(kBlock), 0, g->stream, (int)ndof, ei, part);
  hipLaunchKernelGGL(k_sum3, dim3(1), dim3(64), 0, g->stream, kEnergyBlocks, part, res);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out3, res, 3 * sizeof(double), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

int ft_set_fibres(ft_engine *g, const double *dir) {
  if (!g || !dir) return fail(FT_ERR_ARG, "null argument");
  drop_graphs(g);
  const size_t ne = g->ne;
  const size_t esz = g->precision == FT_FP32 ? 4 : 8;
  if (!g->fib && hipMalloc(&g->fib, 3 * ne * esz) != hipSuccess)
    return fail(FT_ERR_ALLOC, "fibre allocation failed");
  std::vector<double> soa(3 * ne);
  for (size_t e = 0; e < ne; ++e) {
    const double *a = dir + 3 * e;
    const double n = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    if (!(n > 0)) return fail(FT_ERR_ARG, "fibre direction of element %zu is zero", e);
    for (int d = 0; d < 3; ++d) soa[d * ne + g->inv[e]] = a[d] / n;
  }
  if (g->precision == FT_FP32) {
    std::vector<float> f(soa.begin(), soa.end());
    HIPCHK(hipMemcpy(g->fib, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
  } else {
    HIPCHK(hipMemcpy(g->fib, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return 0;
}

int ft_injury_enable(ft_engine *g, const int *include_part) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  drop_graphs(g);
  if (g->n_inst != 1) return fail(FT_ERR_STATE, "the injury criteria are a single-instance feature");
  const size_t ne = g->ne;
  const int nblk = (g->ne + kBlock - 1) / kBlock;
  int rc;
  if (!g->inj_flags) {
    if ((rc = dalloc(&g->inj_flags, ne)) || (rc = dalloc(&g->inj_perm, ne)) || (rc = dalloc(&g->inj_ps_old, ne)) ||
        (rc = dalloc(&g->inj_psxsr, ne)) || (rc = dalloc(&g->inj_vol, ne)) ||
        (rc = dalloc(&g->inj_part, 8 * (size_t)nblk)) || (rc = dalloc(&g->inj_scal, kInjScal)) ||
        (rc = dalloc(&g->inj_vpart, 5 * (size_t)nblk)) || (rc = dalloc(&g->inj_vsum, 5)) ||
        (rc = dalloc(&g->inj_sel, 4)) || (rc = dalloc(&g->inj_hist, 512)))
      return rc;
  }
  // included elements: parts with include_part[p] != 0 (ex5.cpp:1262-1281)
  std::vector<int> fl(ne);
  long n = 0;
  {
    std::vector<int> hinfo(ne);
    HIPCHK(hipMemcpy(hinfo.data(), g->einfo, ne * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < ne; ++i) {
      const int p = hinfo[i] & kPidMask;
      fl[i] = (!include_part || include_part[p]) ? kInjIncluded : 0;
      n += fl[i] ? 1 : 0;
    }
  }
  g->inj_n = n;
  HIPCHK(hipMemcpy(g->inj_flags, fl.data(), ne * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(g->inj_perm, g->perm.data(), ne * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(g->inj_ps_old, 0, ne * sizeof(double)));  // ex5 leaves PS_Old unset; 0 here
  HIPCHK(hipMemset(g->inj_psxsr, 0, ne * sizeof(double)));
  double sc[kInjScal];
  for (int k = 0; k < kInjScal; ++k) sc[k] = 0.0;
  sc[1] = sc[4] = sc[7] = sc[10] = -1.0;  // no element yet
  HIPCHK(hipMemcpy(g->inj_scal, sc, sizeof(sc), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_element_volume, dim3(nblk), dim3(kBlock), 0, g->stream, g->nn, g->ne, g->conn,
                     g->X, g->inj_vol);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(g->stream));
  g->inj_on = true;
  return 0;
}

int ft_injury_disable(ft_engine *g) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  drop_graphs(g);
  g->inj_on = false;
  return 0;
}

int ft_injury_update(ft_engine *g, const double *u, double time, double dt) {
  if (!g) return fail(FT_ERR_ARG, "null engine");
  if (!g->inj_flags) return fail(FT_ERR_STATE, "ft_injury_enable must run first");
  if (!(dt > 0)) return fail(FT_ERR_ARG, "dt must be positive");
  int rc = upload_u(g, u);
  if (rc) return rc;
  if ((rc = injury_launch(g, nullptr, time, dt))) return rc;
  HIPCHK(hipStreamSynchronize(g->stream));
  return 0;
}

int ft_injury_results(ft_engine *g, double out[21], int *flags) {
  if (!g || !out) return fail(FT_ERR_ARG, "null argument");
  if (!g->inj_flags) return fail(FT_ERR_STATE, "ft_injury_enable must run first");
  const int nblk = (g->ne + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_injury_volumes, dim3(nblk), dim3(kBlock), 0, g->stream, g->ne, g->inj_flags, g->inj_vol,
                     g->inj_vpart);
  hipLaunchKernelGGL(k_sum_tuples, dim3(1), dim3(64), 0, g->stream, nblk, 5, g->inj_vpart, g->inj_vsum);
  HIPCHK(hipGetLastError());
  double sc[kInjScal], vs[5];
  HIPCHK(hipMemcpyAsync(sc, g->inj_scal, sizeof(sc), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipMemcpyAsync(vs, g->inj_vsum, sizeof(vs), hipMemcpyDeviceToHost, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  for (int k = 0; k < 16; ++k) out[k] = sc[k];
  for (int k = 0; k < 4; ++k) out[16 + k] = vs[1 + k] / vs[0];
  out[20] = vs[0];
  if (flags) {
    std::vector<int> fl(g->ne);
    HIPCHK(hipMemcpy(fl.data(), g->inj_flags, fl.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < g->ne; ++i) flags[g->perm[i]] = fl[i];
  }
  return 0;
}

int ft_copy_probe(double megabytes, int repeats, double *gbs) {
  if (!(megabytes > 0) || repeats < 1) return fail(FT_ERR_ARG, "bad copy probe size");
  const size_t n = (size_t)(megabytes * 1e6 / 16.0);
  float4 *a = nullptr, *b = nullptr;
  int rc;
  if ((rc = dalloc(&a, n)) || (rc = dalloc(&b, n))) {
    if (a) (void)hipFree(a);
    return rc;
  }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipMemset(a, 0, n * sizeof(float4)));
  hipLaunchKernelGGL(k_copy_probe, dim3(2048), dim3(kBlock), 0, 0, a, b, n);  // warm
  HIPCHK(hipEventRecord(e0, 0));
  for (int r = 0; r < repeats; ++r) hipLaunchKernelGGL(k_copy_probe, dim3(2048), dim3(kBlock), 0, 0, a, b, n);
  HIPCHK(hipEventRecord(e1, 0));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  if (gbs) *gbs = 2.0 * n * sizeof(float4) * repeats / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

int ft_traffic_model(ft_engine *g, double bytes[2]) {
  if (!g || !bytes) return fail(FT_ERR_ARG, "null argument");
  const double es = g->precision == FT_FP32 ? 4.0 : 8.0;
  const double ne = g->ne, nn = g->nn;
  // element kernel: connectivity 32 + einfo 4 + force write 24 es per element;
  // X (3 doubles) + u (3 doubles) per node; Prony history read + write and
  // fibre directions for the elements that carry them.
  // (corner pairs: the right partner's four shared corners are not written)
  double elem = ne * (32.0 + 4.0 + 24.0 * es) + nn * (24.0 + 24.0) - (g->paired ? g->n_pairs * 12.0 * es : 0.0);
  if (g->eslot) elem += ne * 4.0;  // node-slot numbers
  for (const auto &sg : g->segs) {
    const double n = sg.end - sg.begin;
    if (sg.mat == 5 || sg.mat == 7) elem += n * 2.0 * 144.0 * es;  // Prony history read + write
    if (sg.mat == 8) elem += n * 3.0 * es;                         // fibre direction
  }
  // nodal kernel: element forces, 8 slots of 3 per node (corner and node
  // slots; with overflow slots their index per node and the extra ranks of
  // the nodes with more than 8 elements) or 24 es per element + the ELL incidences 4 width per node; per
  // node mass 8, motion code 4, u and v_half read + write 2 x 48 (X only at
  // constrained nodes and fi only when a run stops: not counted).
  // (complete corner pairs: the four xi- slots)
  double nodal = (g->corner ? nn * ((g->pair_complete ? 12.0 : 24.0) * es) : ne * (24.0 * es) + nn * 4.0 * g->width) +
                 (g->ovf ? nn * 4.0 + g->n_ovf * 3.0 * es * (g->width - 8) : 0.0) +
                 nn * (8.0 + 4.0 + 2.0 * 48.0);
  bytes[0] = elem;
  bytes[1] = nodal;
  return 0;
}

}  // extern "C"
