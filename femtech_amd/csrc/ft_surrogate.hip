// ft_surrogate.hip -- the constitutive surrogate network on the matrix cores
// (no counterpart in the reference).
//
//   k_mlp_infer  all five layers of a 64-sample slab per wavefront in one
//                launch: BF16 operands on v_mfma_f32_32x32x16_bf16, FP32
//                accumulation, bias and SiLU on the accumulators, and the
//                layer-to-layer hand-over kept in registers (the weights of
//                layers 1-4 are stored with their input columns permuted to
//                the accumulator layout, so one layer's output tile is the
//                next layer's B operand as it stands).
//   k_gemm       general BF16 MFMA product C = A B with strided operands and
//                split-K partials (training: forward, input and weight
//                gradients).
//   elementwise  bias + SiLU, loss, SiLU backward, column sums, Adam.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/femtech_amd.h"
#include "ft_mlp.h"

extern "C" void ft_internal_set_error(const char *msg);

namespace ft {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

int mfail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int mfail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  ft_internal_set_error(buf);
  return code;
}

#define MCHK(call)                                                                      \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return mfail(FT_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                   __FILE__, __LINE__);                                                 \
  } while (0)

// Flat FP32 parameter vector: per layer W_l [out][in] row-major, then b_l.
__host__ __device__ constexpr int layer_in(int l) { return l == 0 ? kMlpIn : kMlpHid; }
__host__ __device__ constexpr int layer_out(int l) { return l == kMlpLayers - 1 ? kMlpOut : kMlpHid; }
__host__ __device__ constexpr int w_off(int l) {
  return l == 0 ? 0 : w_off(l - 1) + layer_out(l - 1) * (layer_in(l - 1) + 1);
}
__host__ __device__ constexpr int b_off(int l) { return w_off(l) + layer_out(l) * layer_in(l); }
constexpr int kParams = w_off(kMlpLayers);
static_assert(kParams == 52744, "parameter layout");

// Inference copy (BF16): layer 0 [128][16] as is; layers 1-3 [128][128] and
// layer 4 [32][128] (rows past 8 zero) with permuted input columns.
__host__ __device__ constexpr int wi_off(int l) { return l == 0 ? 0 : kMlpHid * kMlpIn + (l - 1) * kMlpHid * kMlpHid; }
constexpr int kWi = wi_off(kMlpLayers - 1) + 32 * kMlpHid;

// Accumulator layout of the 32x32 product (lane l, h = l / 32): element
// 4 j + r of the 16 holds row 8 j + 4 h + r, column l % 32.  The B operand of
// a 16-deep step takes 8 values per lane at depth 8 h + i.  Handing the 16
// accumulators of output tile t over as two B operands (values 0-7 and
// 8-15) feeds depth position p of step s = 2 t + q with the feature below;
// the next layer's weights are stored with their columns in this order.
__host__ __device__ inline int chain_feature(int kpos) {
  const int s = kpos >> 4, p = kpos & 15, h = p >> 3, pp = p & 7;
  return 32 * (s >> 1) + 16 * (s & 1) + 8 * (pp >> 2) + 4 * h + (pp & 3);
}

__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ inline float silu_grad(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

__global__ void k_refresh(const float *P, __bf16 *Pb, __bf16 *Wi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kParams) Pb[i] = (__bf16)P[i];
  if (i < kWi) {
    int l, row, col;
    if (i < wi_off(1)) {
      l = 0;
      row = i / kMlpIn;
      col = i % kMlpIn;
    } else {
      l = 1 + (i - wi_off(1)) / (kMlpHid * kMlpHid);
      const int r = i - wi_off(l);
      row = r / kMlpHid;
      col = chain_feature(r % kMlpHid);
    }
    float v = 0.0f;
    if (row < layer_out(l)) v = P[w_off(l) + row * layer_in(l) + col];
    Wi[i] = (__bf16)v;
  }
}

// SiLU for inference: hardware exponential and reciprocal (1 ulp), the
// BF16 rounding of the result dominates the error.
__device__ inline float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

// Accumulators of the four output tiles of a layer started at the bias (the
// lane's rows 8 j + 4 h + r hold 4 consecutive biases: one 16-byte read).
__device__ inline void bias_init(f32x16 (&acc)[4], const float *b, int h) {
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float4 v = *reinterpret_cast<const float4 *>(b + 32 * t + 8 * j + 4 * h);
      acc[t][4 * j] = v.x;
      acc[t][4 * j + 1] = v.y;
      acc[t][4 * j + 2] = v.z;
      acc[t][4 * j + 3] = v.w;
    }
}

// SiLU on the four output tiles of one 32-sample column tile (bias already
// in the accumulators), then the hand-over to the next layer's eight B
// operands.
__device__ inline void act_chain(const f32x16 (&acc)[4], bf16x8 (&B)[8]) {
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) B[2 * t + (j >> 1)][(j & 1) * 4 + r] = (__bf16)silu_fast(acc[t][4 * j + r]);
}

// y[n][8] = network(x[n][16]).  Persistent blocks of 4 wavefronts: the
// block stages all weights (BF16, 108 KB) and biases in LDS once, then each
// wavefront walks 64-sample slabs (two column tiles), the A operands coming
// from LDS.
// NT column tiles of 32 samples per wavefront, THREADS per block.  Two
// tiles at one wavefront per SIMD measured fastest: one tile at two or four
// wavefronts per SIMD reads every weight fragment from LDS once per tile and
// ran 25 % slower.
constexpr int kInferTiles = 2;
constexpr int kInferThreads = 256;

template <int NT, int THREADS>
__global__ __launch_bounds__(THREADS) void k_mlp_infer(const __bf16 *x, long n, const __bf16 *wi, const float *P,
                                                       float *y) {
  __shared__ __attribute__((aligned(16))) __bf16 sw[kWi];
  __shared__ float sb[kMlpLayers * kMlpHid];
  for (int i = threadIdx.x; i < kWi / 8; i += blockDim.x)
    reinterpret_cast<bf16x8 *>(sw)[i] = reinterpret_cast<const bf16x8 *>(wi)[i];
  for (int i = threadIdx.x; i < kMlpLayers * kMlpHid; i += blockDim.x) {
    const int l = i / kMlpHid, r = i % kMlpHid;
    sb[i] = r < layer_out(l) ? P[b_off(l) + r] : 0.0f;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, h = lane >> 5, c32 = lane & 31;
  constexpr int waves = THREADS / 64, slab_n = 32 * NT;
  const long slabs = (n + slab_n - 1) / slab_n;
  const bf16x8 zero8 = {};
  const f32x16 zero16 = {};
  for (long slab = (long)blockIdx.x * waves + (threadIdx.x >> 6); slab < slabs; slab += (long)gridDim.x * waves) {
    const long base = slab * slab_n;
    bf16x8 B[NT][8];
#pragma unroll
    for (int c = 0; c < NT; ++c) {
      const long s = base + 32 * c + c32;
      B[c][0] = s < n ? *reinterpret_cast<const bf16x8 *>(x + s * kMlpIn + 8 * h) : zero8;
    }
    f32x16 acc[NT][4];
    bias_init(acc[0], sb, h);
#pragma unroll
    for (int c = 1; c < NT; ++c)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[c][t] = acc[0][t];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const bf16x8 a = *reinterpret_cast<const bf16x8 *>(sw + (32 * t + c32) * kMlpIn + 8 * h);
#pragma unroll
      for (int c = 0; c < NT; ++c) acc[c][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, B[c][0], acc[c][t], 0, 0, 0);
    }
#pragma unroll
    for (int c = 0; c < NT; ++c) act_chain(acc[c], B[c]);
#pragma unroll 1
    for (int l = 1; l < kMlpLayers - 1; ++l) {
      const __bf16 *w = sw + wi_off(l);
      bias_init(acc[0], sb + l * kMlpHid, h);
#pragma unroll
      for (int c = 1; c < NT; ++c)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[c][t] = acc[0][t];
      // depth-step outer, tiles inner: independent accumulation chains in
      // flight behind each other on the matrix core
#pragma unroll
      for (int s = 0; s < 8; ++s)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const bf16x8 a = *reinterpret_cast<const bf16x8 *>(w + (32 * t + c32) * kMlpHid + 16 * s + 8 * h);
#pragma unroll
          for (int c = 0; c < NT; ++c)
            acc[c][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, B[c][s], acc[c][t], 0, 0, 0);
        }
#pragma unroll
      for (int c = 0; c < NT; ++c) act_chain(acc[c], B[c]);
    }
    const __bf16 *w = sw + wi_off(kMlpLayers - 1);
    f32x16 o[NT];
#pragma unroll
    for (int c = 0; c < NT; ++c) o[c] = zero16;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const bf16x8 a = *reinterpret_cast<const bf16x8 *>(w + c32 * kMlpHid + 16 * s + 8 * h);
#pragma unroll
      for (int c = 0; c < NT; ++c) o[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, B[c][s], o[c], 0, 0, 0);
    }
    const float *b4 = sb + (kMlpLayers - 1) * kMlpHid;
#pragma unroll
    for (int c = 0; c < NT; ++c) {
      const long s = base + 32 * c + c32;
      if (s < n) {  // rows 4 h + r of the first row group are outputs 4 h + r
        float4 v;
        v.x = o[c][0] + b4[4 * h];
        v.y = o[c][1] + b4[4 * h + 1];
        v.z = o[c][2] + b4[4 * h + 2];
        v.w = o[c][3] + b4[4 * h + 3];
        *reinterpret_cast<float4 *>(y + s * kMlpOut + 4 * h) = v;
      }
    }
  }
}

// C (+ blockIdx.y partial) [M][ldc] = A[M][K] B[K][N] over depths
// [blockIdx.y kchunk, +kchunk); A(m, k) = A[m sam + k sak], B(k, n) =
// B[k sbk + n sbn].  One 32x32 tile per wavefront.
struct GemmArgs {
  int M, N, K;
  const __bf16 *A;
  long sam, sak;
  const __bf16 *B;
  long sbk, sbn;
  float *C;
  int ldc;
  int kchunk;
};

__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  const int lane = threadIdx.x & 63, h = lane >> 5, c32 = lane & 31;
  const int tilesN = (g.N + 31) / 32, tilesM = (g.M + 31) / 32;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= tilesM * tilesN) return;  // uniform over the wavefront
  const int m0 = (tile / tilesN) * 32, n0 = (tile % tilesN) * 32;
  const int kb = blockIdx.y * g.kchunk, ke = min(g.K, kb + g.kchunk);
  const int m = m0 + c32, n = n0 + c32;
  f32x16 acc = {};
  for (int k0 = kb; k0 < ke; k0 += 16) {
    bf16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = k0 + 8 * h + i;
      a[i] = (m < g.M && k < ke) ? g.A[m * g.sam + k * g.sak] : (__bf16)0.0f;
      b[i] = (n < g.N && k < ke) ? g.B[k * g.sbk + n * g.sbn] : (__bf16)0.0f;
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  float *C = g.C + (size_t)blockIdx.y * g.M * g.ldc;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = m0 + 8 * j + 4 * h + r;
      if (row < g.M && n < g.N) C[(size_t)row * g.ldc + n] = acc[4 * j + r];
    }
}

// out[i] = sum_z parts[z count + i], z in order (deterministic split-K).
__global__ void k_sum_parts(const float *parts, int nz, long count, float *out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.0f;
  for (int z = 0; z < nz; ++z) s += parts[(size_t)z * count + i];
  out[i] = s;
}

// Z[n][w] += b; A = BF16(SiLU(Z)) when act, else A = nothing.
__global__ void k_bias_act(float *Z, const float *b, long n, int w, int act, __bf16 *A) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * w) return;
  const float z = Z[i] + b[i % w];
  Z[i] = z;
  if (act) A[i] = (__bf16)silu(z);
}

// Mean squared error over the 6 used outputs: per-block partial sums of
// (y - t)^2, dY = 2 (y - t) / (6 n) (FP32 and BF16).
__global__ void k_loss(const float *Y, const float *T, long n, float *dY, __bf16 *dYb, double *part) {
  __shared__ double sh[256];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double e2 = 0.0;
  if (i < n * kMlpOut) {
    const int k = (int)(i % kMlpOut);
    float d = 0.0f;
    if (k < kMlpUsedOut) {
      const float r = Y[i] - T[i];
      e2 = (double)r * r;
      d = 2.0f * r / (float)(kMlpUsedOut * n);
    }
    dY[i] = d;
    dYb[i] = (__bf16)d;
  }
  sh[threadIdx.x] = e2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// dZ = dA * SiLU'(Z) (FP32 and BF16).
__global__ void k_act_bwd(const float *dA, const float *Z, long count, float *dZ, __bf16 *dZb) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float d = dA[i] * silu_grad(Z[i]);
  dZ[i] = d;
  dZb[i] = (__bf16)d;
}

// db[c] = sum_s dZ[s][c], one block per column, fixed order.
__global__ void k_colsum(const float *dZ, long n, int w, float *db) {
  __shared__ float sh[256];
  const int c = blockIdx.x;
  float s = 0.0f;
  for (long i = threadIdx.x; i < n; i += blockDim.x) s += dZ[i * w + c];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] = sh[0];
}

__global__ void k_adam(float *P, const float *G, float *m, float *v, float lr, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  const float g = G[i];
  const float mi = 0.9f * m[i] + 0.1f * g;
  const float vi = 0.999f * v[i] + 0.001f * g * g;
  m[i] = mi;
  v[i] = vi;
  P[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
}

__device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Batch of `batch` samples drawn uniformly (with replacement) from the data.
__global__ void k_gather_batch(const __bf16 *x, const float *t, long n_data, int batch, uint64_t key, __bf16 *xb,
                               float *tb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch) return;
  const long s = (long)(mix64(key ^ (uint64_t)i * 0xD1B54A32D192ED03ull) % (uint64_t)n_data);
  *reinterpret_cast<bf16x8 *>(xb + (size_t)i * kMlpIn) = *reinterpret_cast<const bf16x8 *>(x + s * kMlpIn);
  *reinterpret_cast<bf16x8 *>(xb + (size_t)i * kMlpIn + 8) = *reinterpret_cast<const bf16x8 *>(x + s * kMlpIn + 8);
  *reinterpret_cast<float4 *>(tb + (size_t)i * kMlpOut) = *reinterpret_cast<const float4 *>(t + s * kMlpOut);
  *reinterpret_cast<float4 *>(tb + (size_t)i * kMlpOut + 4) = *reinterpret_cast<const float4 *>(t + s * kMlpOut + 4);
}

int blocks(long n, int b = 256) { return (int)((n + b - 1) / b); }

}  // namespace

struct MlpNet {
  float *P = nullptr, *G = nullptr, *m = nullptr, *v = nullptr;
  __bf16 *Pb = nullptr, *Wi = nullptr;
  int adam_t = 0;
  int n_cu = 0;  // compute units (persistent inference grid)
  // training workspace for batches up to `cap` samples
  int cap = 0;
  float *Z[kMlpLayers] = {};
  __bf16 *A[kMlpLayers - 1] = {};
  float *dZ = nullptr, *dA = nullptr, *parts = nullptr;
  __bf16 *dZb = nullptr, *xb = nullptr;
  float *tb = nullptr, *dY = nullptr;
  __bf16 *dYb = nullptr;
  double *lpart = nullptr;
  size_t parts_len = 0;
};

namespace {

void free_ws(MlpNet *n) {
  for (auto &p : n->Z) (void)hipFree(p), p = nullptr;
  for (auto &p : n->A) (void)hipFree(p), p = nullptr;
  for (void *p : {(void *)n->dZ, (void *)n->dA, (void *)n->parts, (void *)n->dZb, (void *)n->xb, (void *)n->tb,
                  (void *)n->dY, (void *)n->dYb, (void *)n->lpart})
    if (p) (void)hipFree(p);
  n->dZ = n->dA = n->parts = n->tb = n->dY = nullptr;
  n->dZb = n->xb = n->dYb = nullptr;
  n->lpart = nullptr;
  n->cap = 0;
}

int ensure_ws(MlpNet *n, int batch) {
  if (batch <= n->cap) return 0;
  free_ws(n);
  const size_t b = (size_t)batch;
  for (int l = 0; l < kMlpLayers; ++l) MCHK(hipMalloc(&n->Z[l], b * layer_out(l) * sizeof(float)));
  for (int l = 0; l < kMlpLayers - 1; ++l) MCHK(hipMalloc(&n->A[l], b * layer_out(l) * sizeof(__bf16)));
  MCHK(hipMalloc(&n->dZ, b * kMlpHid * sizeof(float)));
  MCHK(hipMalloc(&n->dA, b * kMlpHid * sizeof(float)));
  MCHK(hipMalloc(&n->dZb, b * kMlpHid * sizeof(__bf16)));
  MCHK(hipMalloc(&n->xb, b * kMlpIn * sizeof(__bf16)));
  MCHK(hipMalloc(&n->tb, b * kMlpOut * sizeof(float)));
  MCHK(hipMalloc(&n->dY, b * kMlpOut * sizeof(float)));
  MCHK(hipMalloc(&n->dYb, b * kMlpOut * sizeof(__bf16)));
  MCHK(hipMalloc(&n->lpart, (size_t)blocks(b * kMlpOut) * sizeof(double)));
  n->parts_len = (size_t)((batch + 255) / 256) * kMlpHid * kMlpHid;
  MCHK(hipMalloc(&n->parts, n->parts_len * sizeof(float)));
  n->cap = batch;
  return 0;
}

// C[M][ldc] = A B, split over the depth in chunks of `kchunk` when K is long.
int gemm(MlpNet *net, GemmArgs g, hipStream_t st) {
  const int tiles = ((g.M + 31) / 32) * ((g.N + 31) / 32);
  const int nz = (g.K + g.kchunk - 1) / g.kchunk;
  float *out = g.C;
  const int ldc = g.ldc;
  if (nz > 1) {
    if ((size_t)nz * g.M * g.N > net->parts_len) return mfail(FT_ERR_ARG, "split-K workspace too small");
    g.C = net->parts;
    g.ldc = g.N;
  }
  hipLaunchKernelGGL(k_gemm, dim3((tiles + 3) / 4, nz), dim3(256), 0, st, g);
  MCHK(hipGetLastError());
  if (nz > 1) {
    if (ldc != g.N) return mfail(FT_ERR_ARG, "split-K output must be dense");
    const long count = (long)g.M * g.N;
    hipLaunchKernelGGL(k_sum_parts, dim3(blocks(count)), dim3(256), 0, st, net->parts, nz, count, out);
    MCHK(hipGetLastError());
  }
  return 0;
}

int refresh(MlpNet *net, hipStream_t st) {
  hipLaunchKernelGGL(k_refresh, dim3(blocks(std::max(kParams, kWi))), dim3(256), 0, st, net->P, net->Pb, net->Wi);
  MCHK(hipGetLastError());
  return 0;
}

}  // namespace

int mlp_param_count() { return kParams; }

int mlp_create(unsigned long long seed, MlpNet **out) {
  MlpNet *n = new MlpNet();
  *out = nullptr;
  auto bail = [&](int rc) {
    mlp_destroy(n);
    return rc;
  };
  if (hipMalloc(&n->P, kParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&n->G, kParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&n->m, kParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&n->v, kParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&n->Pb, kParams * sizeof(__bf16)) != hipSuccess ||
      hipMalloc(&n->Wi, kWi * sizeof(__bf16)) != hipSuccess)
    return bail(mfail(FT_ERR_ALLOC, "network allocation failed"));
  // Glorot-uniform weights from a 64-bit mix of (seed, index), zero biases
  std::vector<float> h(kParams, 0.0f);
  for (int l = 0; l < kMlpLayers; ++l) {
    const int fi = l == 0 ? kMlpUsedIn : layer_in(l), fo = l == kMlpLayers - 1 ? kMlpUsedOut : layer_out(l);
    const double a = sqrt(6.0 / (fi + fo));
    for (int o = 0; o < fo; ++o)
      for (int i = 0; i < fi; ++i) {
        uint64_t z = seed * 0x9E3779B97F4A7C15ull + (uint64_t)(w_off(l) + o * layer_in(l) + i) + 1;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
        h[w_off(l) + o * layer_in(l) + i] = (float)((2.0 * u - 1.0) * a);
      }
  }
  int rc = mlp_set_params(n, h.data());
  if (rc) return bail(rc);
  *out = n;
  return 0;
}

void mlp_destroy(MlpNet *n) {
  if (!n) return;
  (void)hipDeviceSynchronize();
  free_ws(n);
  for (void *p : {(void *)n->P, (void *)n->G, (void *)n->m, (void *)n->v, (void *)n->Pb, (void *)n->Wi})
    if (p) (void)hipFree(p);
  delete n;
}

int mlp_get_params(MlpNet *net, float *host) {
  MCHK(hipMemcpy(host, net->P, kParams * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int mlp_set_params(MlpNet *net, const float *host) {
  MCHK(hipMemcpy(net->P, host, kParams * sizeof(float), hipMemcpyHostToDevice));
  MCHK(hipMemset(net->m, 0, kParams * sizeof(float)));
  MCHK(hipMemset(net->v, 0, kParams * sizeof(float)));
  MCHK(hipMemset(net->G, 0, kParams * sizeof(float)));
  net->adam_t = 0;
  int rc = refresh(net, nullptr);
  if (rc) return rc;
  MCHK(hipDeviceSynchronize());
  return 0;
}

int mlp_copy_gradients(MlpNet *net, float *host) {
  MCHK(hipMemcpy(host, net->G, kParams * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int mlp_infer(MlpNet *net, const __bf16 *x, long n, float *y, hipStream_t st) {
  if (n <= 0) return 0;
  if (!net->n_cu) {
    int dev = 0;
    MCHK(hipGetDevice(&dev));
    MCHK(hipDeviceGetAttribute(&net->n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  }
  const long slabs = (n + 32 * kInferTiles - 1) / (32 * kInferTiles), per_block = kInferThreads / 64;
  const long blocks_needed = (slabs + per_block - 1) / per_block;
  const unsigned grid = (unsigned)std::min<long>(blocks_needed, (long)net->n_cu);
  hipLaunchKernelGGL((k_mlp_infer<kInferTiles, kInferThreads>), dim3(grid), dim3(kInferThreads), 0, st, x, n,
                     net->Wi, net->P, y);
  MCHK(hipGetLastError());
  return 0;
}

int mlp_gradients(MlpNet *net, const __bf16 *x, const float *t, int n, hipStream_t st, double *loss) {
  if (n <= 0) return mfail(FT_ERR_ARG, "empty batch");
  int rc;
  if ((rc = ensure_ws(net, n))) return rc;
  // forward: Z_l = A_{l-1} W_l^T + b_l, A_l = SiLU(Z_l)
  const __bf16 *Aprev = x;
  for (int l = 0; l < kMlpLayers; ++l) {
    const int in = layer_in(l), out = layer_out(l);
    GemmArgs g{n, out, in, Aprev, in, 1, net->Pb + w_off(l), 1, in, net->Z[l], out, in};
    if ((rc = gemm(net, g, st))) return rc;
    const bool act = l < kMlpLayers - 1;
    hipLaunchKernelGGL(k_bias_act, dim3(blocks((long)n * out)), dim3(256), 0, st, net->Z[l], net->P + b_off(l),
                       (long)n, out, act ? 1 : 0, act ? net->A[l] : nullptr);
    MCHK(hipGetLastError());
    if (act) Aprev = net->A[l];
  }
  const int nb = blocks((long)n * kMlpOut);
  hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, st, net->Z[kMlpLayers - 1], t, (long)n, net->dY, net->dYb,
                     net->lpart);
  MCHK(hipGetLastError());
  // backward
  const float *dZ = net->dY;
  const __bf16 *dZb = net->dYb;
  for (int l = kMlpLayers - 1; l >= 0; --l) {
    const int in = layer_in(l), out = layer_out(l);
    const __bf16 *Ain = l == 0 ? x : net->A[l - 1];
    // dW_l[o][i] = sum_s dZ[s][o] A_{l-1}[s][i]
    GemmArgs gw{out, in, n, dZb, 1, out, Ain, in, 1, net->G + w_off(l), in, 256};
    if ((rc = gemm(net, gw, st))) return rc;
    hipLaunchKernelGGL(k_colsum, dim3(out), dim3(256), 0, st, dZ, (long)n, out, net->G + b_off(l));
    MCHK(hipGetLastError());
    if (l == 0) break;
    // dA_{l-1}[s][i] = sum_o dZ[s][o] W_l[o][i]; dZ_{l-1} = dA SiLU'(Z_{l-1})
    GemmArgs gd{n, in, out, dZb, out, 1, net->Pb + w_off(l), in, 1, net->dA, in, out};
    if ((rc = gemm(net, gd, st))) return rc;
    hipLaunchKernelGGL(k_act_bwd, dim3(blocks((long)n * in)), dim3(256), 0, st, net->dA, net->Z[l - 1],
                       (long)n * in, net->dZ, net->dZb);
    MCHK(hipGetLastError());
    dZ = net->dZ;
    dZb = net->dZb;
  }
  if (loss) {
    std::vector<double> part(nb);
    MCHK(hipMemcpyAsync(part.data(), net->lpart, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    double s = 0.0;
    for (double p : part) s += p;
    *loss = s / ((double)n * kMlpUsedOut);
  }
  return 0;
}

int mlp_adam(MlpNet *net, float lr, hipStream_t st) {
  net->adam_t += 1;
  const float c1 = (float)(1.0 / (1.0 - pow(0.9, net->adam_t)));
  const float c2 = (float)(1.0 / (1.0 - pow(0.999, net->adam_t)));
  hipLaunchKernelGGL(k_adam, dim3(blocks(kParams)), dim3(256), 0, st, net->P, net->G, net->m, net->v, lr, c1, c2);
  MCHK(hipGetLastError());
  return refresh(net, st);
}

int mlp_train(MlpNet *net, const __bf16 *x, const float *t, long n_data, int steps, int batch, float lr,
              unsigned long long seed, hipStream_t st, double *final_loss) {
  if (n_data <= 0 || batch <= 0 || steps < 0) return mfail(FT_ERR_ARG, "empty training set or batch");
  int rc;
  if ((rc = ensure_ws(net, batch))) return rc;
  double acc = 0.0;
  int nacc = 0;
  for (int k = 0; k < steps; ++k) {
    const uint64_t key = seed * 0x9E3779B97F4A7C15ull ^ ((uint64_t)k << 20);
    hipLaunchKernelGGL(k_gather_batch, dim3(blocks(batch)), dim3(256), 0, st, x, t, n_data, batch, key, net->xb,
                       net->tb);
    MCHK(hipGetLastError());
    // the loss is read back (a synchronisation) only over the last 1 % of steps
    const bool read = final_loss && k >= steps - std::max(1, steps / 100);
    double l = 0.0;
    if ((rc = mlp_gradients(net, net->xb, net->tb, batch, st, read ? &l : nullptr))) return rc;
    if (read) {
      acc += l;
      ++nacc;
    }
    const float lr_k = lr * (0.02f + 0.98f * 0.5f * (1.0f + cosf(3.14159265f * (float)k / (float)steps)));
    if ((rc = mlp_adam(net, lr_k, st))) return rc;
  }
  MCHK(hipStreamSynchronize(st));
  if (final_loss) *final_loss = nacc ? acc / nacc : 0.0;
  return 0;
}

}  // namespace ft

// ---------------------------------------------------------------------------
// C ABI of the stand-alone network (host arrays; tests and tools)
// ---------------------------------------------------------------------------
struct ft_mlp {
  ft::MlpNet *net = nullptr;
};

namespace {

int to_device_bf16(const float *host, size_t count, __bf16 **dev) {
  std::vector<__bf16> h(count);
  for (size_t i = 0; i < count; ++i) h[i] = (__bf16)host[i];
  if (hipMalloc(dev, count * sizeof(__bf16)) != hipSuccess) return ft::mfail(FT_ERR_ALLOC, "allocation failed");
  if (hipMemcpy(*dev, h.data(), count * sizeof(__bf16), hipMemcpyHostToDevice) != hipSuccess)
    return ft::mfail(FT_ERR_HIP, "copy failed");
  return 0;
}

int to_device_f32(const float *host, size_t count, float **dev) {
  if (hipMalloc(dev, count * sizeof(float)) != hipSuccess) return ft::mfail(FT_ERR_ALLOC, "allocation failed");
  if (hipMemcpy(*dev, host, count * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return ft::mfail(FT_ERR_HIP, "copy failed");
  return 0;
}

int device_ok() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return ft::mfail(FT_ERR_HIP, "no HIP device available");
  return 0;
}

}  // namespace

extern "C" {

int ft_mlp_param_count(void) { return ft::kParams; }

int ft_mlp_create(unsigned long long seed, ft_mlp **out) {
  if (!out) return ft::mfail(FT_ERR_ARG, "null argument");
  *out = nullptr;
  int rc;
  if ((rc = device_ok())) return rc;
  ft_mlp *m = new ft_mlp();
  if ((rc = ft::mlp_create(seed, &m->net))) {
    delete m;
    return rc;
  }
  *out = m;
  return 0;
}

int ft_mlp_destroy(ft_mlp *m) {
  if (!m) return 0;
  ft::mlp_destroy(m->net);
  delete m;
  return 0;
}

int ft_mlp_get_params(ft_mlp *m, float *params) {
  if (!m || !params) return ft::mfail(FT_ERR_ARG, "null argument");
  return ft::mlp_get_params(m->net, params);
}

int ft_mlp_set_params(ft_mlp *m, const float *params) {
  if (!m || !params) return ft::mfail(FT_ERR_ARG, "null argument");
  return ft::mlp_set_params(m->net, params);
}

int ft_mlp_forward(ft_mlp *m, const float *x, int n, float *y) {
  if (!m || !x || !y || n < 0) return ft::mfail(FT_ERR_ARG, "bad argument");
  if (n == 0) return 0;
  __bf16 *dx = nullptr;
  float *dy = nullptr;
  int rc = to_device_bf16(x, (size_t)n * ft::kMlpIn, &dx);
  if (!rc && hipMalloc(&dy, (size_t)n * ft::kMlpOut * sizeof(float)) != hipSuccess)
    rc = ft::mfail(FT_ERR_ALLOC, "allocation failed");
  if (!rc) rc = ft::mlp_infer(m->net, dx, n, dy, nullptr);
  if (!rc && hipMemcpy(y, dy, (size_t)n * ft::kMlpOut * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    rc = ft::mfail(FT_ERR_HIP, "copy failed");
  if (dx) (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  return rc;
}

int ft_mlp_gradients(ft_mlp *m, const float *x, const float *t, int n, double *loss, float *grad) {
  if (!m || !x || !t || n <= 0) return ft::mfail(FT_ERR_ARG, "bad argument");
  __bf16 *dx = nullptr;
  float *dt = nullptr;
  int rc = to_device_bf16(x, (size_t)n * ft::kMlpIn, &dx);
  if (!rc) rc = to_device_f32(t, (size_t)n * ft::kMlpOut, &dt);
  if (!rc) rc = ft::mlp_gradients(m->net, dx, dt, n, nullptr, loss);
  if (!rc && grad) rc = ft::mlp_copy_gradients(m->net, grad);
  if (dx) (void)hipFree(dx);
  if (dt) (void)hipFree(dt);
  return rc;
}

int ft_mlp_train(ft_mlp *m, const float *x, const float *t, int n_data, int steps, int batch, double lr,
                 unsigned long long seed, double *final_loss) {
  if (!m || !x || !t || n_data <= 0) return ft::mfail(FT_ERR_ARG, "bad argument");
  __bf16 *dx = nullptr;
  float *dt = nullptr;
  int rc = to_device_bf16(x, (size_t)n_data * ft::kMlpIn, &dx);
  if (!rc) rc = to_device_f32(t, (size_t)n_data * ft::kMlpOut, &dt);
  if (!rc) rc = ft::mlp_train(m->net, dx, dt, n_data, steps, batch, (float)lr, seed, nullptr, final_loss);
  if (dx) (void)hipFree(dx);
  if (dt) (void)hipFree(dt);
  return rc;
}

}  // extern "C"
