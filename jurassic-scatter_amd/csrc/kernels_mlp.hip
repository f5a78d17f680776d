// kernels_mlp.hip -- matrix-core kernels of the scattering-source surrogate
// (a 24 -> 4 x 256 SiLU -> 1 perceptron): an LDS-tiled BF16 GEMM on
// v_mfma_f32_32x32x16_bf16 with split-K and a fused bias + SiLU epilogue,
// the element-wise and reduction pieces of the backward pass, Adam, and a
// fused inference kernel that keeps a 64-sample block's activations in LDS
// through all five layers.
//
// Operand layout of v_mfma_f32_32x32x16_bf16 (one wave, C[32][32] +=
// A[32][16] B[16][32]): lane l supplies A[l % 32][8 (l / 32) + 0..7] and
// B[8 (l / 32) + 0..7][l % 32]; accumulator r of lane l holds
// C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
#include "jur_mlp.h"

namespace {

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ inline uint16_t f2bf(float f) { return mlp_f2bf(f); }
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// SiLU and its derivative with the hardware exponential and reciprocal
// (v_exp_f32, v_rcp_f32: about 1 ulp); the results are rounded to BF16
// (8-bit mantissa) or feed BF16 products, far coarser than these errors.
__device__ inline float sigmoid_fast(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigmoid_fast(z); }
__device__ inline float dsilu(float z) {
  const float s = sigmoid_fast(z);
  return s * (1.f + z * (1.f - s));
}

constexpr int TS = 64;        // block tile (M and N)
constexpr int TK = 32;        // K per LDS stage
constexpr int LDK = TK + 8;   // padded LDS row, BF16 elements (80 B)
constexpr int LDH = MLP_H + 8;

__device__ inline f32x16 zero16() {
  f32x16 z;
#pragma unroll
  for (int i = 0; i < 16; i++) z[i] = 0.f;
  return z;
}

}  // namespace

__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  __shared__ __align__(16) uint16_t As[TS][LDK];
  __shared__ __align__(16) uint16_t Bs[TS][LDK];  // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * TS, n0 = blockIdx.x * TS;
  const int kper = g.K / g.ksplit;
  const int kbeg = blockIdx.z * kper, kend = kbeg + kper;
  const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
  const bool active = (m0 + wm < g.M) && (n0 + wn < g.N);
  f32x16 acc = zero16();
  for (int k0 = kbeg; k0 < kend; k0 += TK) {
    // stage A[m0..+64][k0..+32] and B[k0..+32][n0..+64], walking the
    // contiguous dimension with consecutive threads
    if (g.sak == 1) {
      for (int i = tid; i < TS * TK; i += 256) {
        const int r = i >> 5, c = i & 31, m = m0 + r;
        As[r][c] = m < g.M ? g.A[(long)m * g.sam + (k0 + c)] : (uint16_t)0;
      }
    } else {
      for (int i = tid; i < TS * TK; i += 256) {
        const int r = i & 63, c = i >> 6, m = m0 + r;
        As[r][c] = m < g.M ? g.A[(long)m * g.sam + (long)(k0 + c) * g.sak] : (uint16_t)0;
      }
    }
    if (g.sbn == 1) {
      for (int i = tid; i < TS * TK; i += 256) {
        const int c = i & 63, r = i >> 6, n = n0 + c;
        Bs[c][r] = n < g.N ? g.B[(long)(k0 + r) * g.sbk + n] : (uint16_t)0;
      }
    } else {
      for (int i = tid; i < TS * TK; i += 256) {
        const int r = i & 31, c = i >> 5, n = n0 + c;
        Bs[c][r] = n < g.N ? g.B[(long)(k0 + r) * g.sbk + (long)n * g.sbn] : (uint16_t)0;
      }
    }
    __syncthreads();
    if (active) {
      const int h = lane >> 5, r = lane & 31;
#pragma unroll
      for (int kk = 0; kk < TK; kk += 16) {
        const bf16x8 a = *(const bf16x8*)&As[wm + r][kk + 8 * h];
        const bf16x8 b = *(const bf16x8*)&Bs[wn + r][kk + 8 * h];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
      }
    }
    __syncthreads();
  }
  if (!active) return;
  const int j = n0 + wn + (lane & 31);
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int i = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (g.epi == 0) {
      g.C[(long)blockIdx.z * g.cslab + (long)i * g.ldc + j] = acc[r];
    } else {
      const float z = acc[r] + g.bias[j];
      g.C[(long)i * g.ldc + j] = z;
      g.act[(long)i * g.ldact + j] = f2bf(silu(z));
    }
  }
}

// The same GEMM with 16-byte global loads and a double-buffered 64-deep K
// stage (the next stage's loads in flight during the current stage's matrix
// operations).  Operands whose contiguous dimension is K are staged as read;
// those contiguous in M or N are transposed on the way into LDS (a thread's
// 8 elements to 8 rows; consecutive threads take consecutive k, so the 2-byte
// LDS writes of a wave fall in distinct banks).  The K chunks of 16 reach
// the matrix core in the same order as in k_gemm: identical results.
// Requires M and N multiples of 8, K per split a multiple of TK2 and
// 16-byte aligned rows (checked by the launcher).
constexpr int TK2 = 64;
constexpr int LDK2 = TK2 + 8;  // 144-byte LDS rows
typedef unsigned int u32x4v __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_gemm_v(GemmArgs g) {
  __shared__ __align__(16) uint16_t As[2][TS][LDK2];
  __shared__ __align__(16) uint16_t Bs[2][TS][LDK2];  // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * TS, n0 = blockIdx.x * TS;
  const int kper = g.K / g.ksplit;
  const int kbeg = blockIdx.z * kper, kend = kbeg + kper;
  const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
  const bool active = (m0 + wm < g.M) && (n0 + wn < g.N);
  const bool adirect = g.sak == 1, bdirect = g.sbk == 1;
  u32x4v ra[2], rb[2];
  // direct: vector v = (row, 8 k); transposed: v = (8 rows, k) with k fastest
  auto gload = [&](int k0) {
#pragma unroll
    for (int v = 0; v < 2; v++) {
      const int idx = tid + v * 256;
      const u32x4v z = {0u, 0u, 0u, 0u};
      if (adirect) {
        const int r = idx >> 3, c = (idx & 7) * 8, m = m0 + r;
        ra[v] = m < g.M ? *(const u32x4v*)&g.A[(long)m * g.sam + k0 + c] : z;
      } else {
        const int k = idx & 63, mc = (idx >> 6) * 8;
        ra[v] = m0 + mc < g.M ? *(const u32x4v*)&g.A[(long)(m0 + mc) + (long)(k0 + k) * g.sak] : z;
      }
      if (bdirect) {
        const int r = idx >> 3, c = (idx & 7) * 8, n = n0 + r;
        rb[v] = n < g.N ? *(const u32x4v*)&g.B[(long)n * g.sbn + k0 + c] : z;
      } else {
        const int k = idx & 63, nc = (idx >> 6) * 8;
        rb[v] = n0 + nc < g.N ? *(const u32x4v*)&g.B[(long)(k0 + k) * g.sbk + n0 + nc] : z;
      }
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int v = 0; v < 2; v++) {
      const int idx = tid + v * 256;
      if (adirect) {
        *(u32x4v*)&As[buf][idx >> 3][(idx & 7) * 8] = ra[v];
      } else {
        const int k = idx & 63, mc = (idx >> 6) * 8;
#pragma unroll
        for (int j = 0; j < 8; j++) As[buf][mc + j][k] = (uint16_t)(ra[v][j >> 1] >> (16 * (j & 1)));
      }
      if (bdirect) {
        *(u32x4v*)&Bs[buf][idx >> 3][(idx & 7) * 8] = rb[v];
      } else {
        const int k = idx & 63, nc = (idx >> 6) * 8;
#pragma unroll
        for (int j = 0; j < 8; j++) Bs[buf][nc + j][k] = (uint16_t)(rb[v][j >> 1] >> (16 * (j & 1)));
      }
    }
  };
  f32x16 acc = zero16();
  gload(kbeg);
  lstore(0);
  __syncthreads();
  int cur = 0;
  for (int k0 = kbeg; k0 < kend; k0 += TK2) {
    const bool more = k0 + TK2 < kend;
    if (more) gload(k0 + TK2);
    if (active) {
      const int h = lane >> 5, r = lane & 31;
#pragma unroll
      for (int kk = 0; kk < TK2; kk += 16) {
        const bf16x8 a = *(const bf16x8*)&As[cur][wm + r][kk + 8 * h];
        const bf16x8 b = *(const bf16x8*)&Bs[cur][wn + r][kk + 8 * h];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
      }
    }
    if (more) lstore(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
  if (!active) return;
  const int j = n0 + wn + (lane & 31);
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int i = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (g.epi == 0) {
      g.C[(long)blockIdx.z * g.cslab + (long)i * g.ldc + j] = acc[r];
    } else {
      const float z = acc[r] + g.bias[j];
      g.C[(long)i * g.ldc + j] = z;
      g.act[(long)i * g.ldact + j] = f2bf(silu(z));
    }
  }
}

__global__ void k_splitk_sum(const float* part, int nslab, long slab, long n, float* out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // slabs summed in order; unrolled so that their loads are in flight together
  float s = 0.f;
#pragma unroll 16
  for (int z = 0; z < nslab; z++) s += part[(long)z * slab + i];
  out[i] = s;
}

// Batch b of training step t takes row (t B + b) mod ntot (the rows are
// shuffled once on upload).
__global__ void k_gather_batch(const uint16_t* X, const float* Y, long ntot, int B, const int* step,
                               uint16_t* xb, float* yb) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const long row = ((long)(*step) * B + b) % ntot;
  const uint4* src = (const uint4*)(X + row * MLP_INP);
  uint4* dst = (uint4*)(xb + (long)b * MLP_INP);
#pragma unroll
  for (int i = 0; i < MLP_INP * 2 / 16; i++) dst[i] = src[i];
  yb[b] = Y[row];
}

// y[b] = act[b][:] . W5 + b5, one wave per sample.
__global__ void k_out_fwd(const uint16_t* act, const uint16_t* w5, const float* b5, int B, float* y) {
  const int b = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= B) return;
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < MLP_H / 64; q++) {
    const int j = lane + 64 * q;
    s += bf2f(act[(long)b * MLP_H + j]) * bf2f(w5[j]);
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) y[b] = s + b5[0];
}

// Mean squared error: dy = 2 (y - t) / B, loss_hist[step & mask] = mean (y - t)^2.
__global__ void __launch_bounds__(1024) k_loss_grad(const float* y, const float* t, int B, float* dy,
                                                    float* loss_hist, const int* step, int mask) {
  __shared__ float red[16];
  float s = 0.f;
  for (int b = threadIdx.x; b < B; b += blockDim.x) {
    const float d = y[b] - t[b];
    dy[b] = 2.f * d / (float)B;
    s += d * d;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float tot = 0.f;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) tot += red[i];
    if (loss_hist) loss_hist[(*step) & mask] = tot / (float)B;
  }
}

// dz4 = dy W5 silu'(z4) (BF16).
__global__ void k_out_dz(const float* z4, const float* dy, const uint16_t* w5, int B, uint16_t* dz4) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)B * MLP_H) return;
  const int b = (int)(i / MLP_H), j = (int)(i % MLP_H);
  dz4[i] = f2bf(dy[b] * bf2f(w5[j]) * dsilu(z4[i]));
}

// gW5[j] = sum_b act4[b][j] dy[b] (block j < 256), gb5 = sum_b dy[b] (block 256).
// Output-layer weight and bias gradients, part 1: block z sums rows
// 64 z .. 64 z + 63 of act4 x dy per column (column MLP_H: dy itself, the
// bias); the blocks' partial sums are added in order by k_splitk_sum.
__global__ void __launch_bounds__(256) k_out_wgrad_part(const uint16_t* act4, const float* dy, int B,
                                                        float* part) {
  const int j = threadIdx.x, r0 = blockIdx.x * 64;
  const int r1 = r0 + 64 < B ? r0 + 64 : B;
  float s = 0.f, sb = 0.f;
#pragma unroll 16
  for (int r = r0; r < r1; r++) {
    const float d = dy[r];
    s += bf2f(act4[(long)r * MLP_H + j]) * d;
    sb += d;
  }
  part[(long)blockIdx.x * (MLP_H + 1) + j] = s;
  if (j == 0) part[(long)blockIdx.x * (MLP_H + 1) + MLP_H] = sb;
}

__global__ void k_silu_back(const float* da, const float* z, long n, uint16_t* dz) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dz[i] = f2bf(da[i] * dsilu(z[i]));
}

// Column sums of dz [B][256] over 64-row chunks: part[chunk][j].
__global__ void __launch_bounds__(256) k_colsum_part(const uint16_t* dz, int B, float* part) {
  const int j = threadIdx.x, r0 = blockIdx.x * 64;
  float s = 0.f;
  const int r1 = r0 + 64 < B ? r0 + 64 : B;
#pragma unroll 16
  for (int r = r0; r < r1; r++) s += bf2f(dz[(long)r * MLP_H + j]);
  part[(long)blockIdx.x * MLP_H + j] = s;
}

__global__ void k_adam(float* w, const float* g, float* m, float* v, uint16_t* wbf, long n, const int* step,
                       float lr, float b1, float b2, float eps) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float t = (float)(*step + 1);
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - powf(b1, t));
  const float vh = vi / (1.f - powf(b2, t));
  const float wi = w[i] - lr * mh / (sqrtf(vh) + eps);
  w[i] = wi;
  wbf[i] = f2bf(wi);
}

__global__ void k_step_inc(int* step) { step[0] = step[0] + 1; }

__global__ void k_f2bf(const float* a, uint16_t* b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

__global__ void k_transpose_bf(const uint16_t* a, int rows, int cols, uint16_t* at) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)rows * cols) return;
  const int r = (int)(i / cols), c = (int)(i % cols);
  at[(long)c * rows + r] = a[i];
}

__global__ void k_fill_f32(float* p, long n, float v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Fused inference: 128 samples per block, 4 waves.  Default (variant 1,
// hidden_layer_w below): wave w computes all 128 samples (four 32-sample
// tiles) for neurons 64 w .. +64 (two 32-neuron tiles) of every hidden
// layer, so each 16-byte weight fragment is read from L2 once per block and
// feeds four matrix-core operations.  Variant 0 (hidden_layer): wave w
// computes samples 64 (w & 1) .. +64 and neurons 128 (w >> 1) .. +128, each
// fragment read by two waves and feeding two operations.  The layers are
// inlined into the kernel (as out-of-line calls they cost a scratch stack
// and 40 % of the time).  The weights are
// the A operand and the activations the B operand, so a lane's accumulator
// holds four consecutive neurons of one sample: the epilogue converts them
// with the packed BF16 instruction and stores 8 bytes at a time.  The
// activations stay in one LDS buffer (BF16, sample-major): a layer's
// outputs wait in the accumulators until every wave has read its inputs,
// then overwrite them.  (Staging the weight slices through LDS instead, one
// block per CU, measured 1.7x slower: the per-slice barriers and the lost
// second block per CU cost more than the halved L2 reads.)
// ---------------------------------------------------------------------------
namespace {
constexpr int INF_M = 128;  // samples per block
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Two FP32 values to packed BF16, round to nearest even (v_cvt_pk_bf16_f32).
__device__ inline uint32_t pack_bf16(float lo, float hi) {
  f32x2 v = {lo, hi};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

template <int K>
__device__ __forceinline__ void hidden_layer(uint16_t (*act)[LDH], const uint16_t* wt, const float* bias, int wm,
                                             int nb, int lane) {
  const int h = lane >> 5, r = lane & 31;
  f32x16 acc[2][4];  // [sample tile][neuron tile]
#pragma unroll
  for (int u = 0; u < 2; u++)
#pragma unroll
    for (int t = 0; t < 4; t++) acc[u][t] = zero16();
  // weight fragments of the next K step are loaded while the current step's
  // matrix operations run
  const uint16_t* wrow = wt + (long)(nb + r) * K + 8 * h;
  bf16x8 wn[4];
#pragma unroll
  for (int t = 0; t < 4; t++) wn[t] = *(const bf16x8*)&wrow[(long)32 * t * K];
  for (int kk = 0; kk < K; kk += 16) {
    bf16x8 wc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) wc[t] = wn[t];
    if (kk + 16 < K) {
#pragma unroll
      for (int t = 0; t < 4; t++) wn[t] = *(const bf16x8*)&wrow[(long)32 * t * K + kk + 16];
    }
    const bf16x8 x0 = *(const bf16x8*)&act[wm + r][kk + 8 * h];
    const bf16x8 x1 = *(const bf16x8*)&act[wm + 32 + r][kk + 8 * h];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wc[t], x0, acc[0][t], 0, 0, 0);
      acc[1][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wc[t], x1, acc[1][t], 0, 0, 0);
    }
  }
  __syncthreads();  // every wave has read this layer's inputs
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int g = 0; g < 4; g++) {
      // neurons n0 .. n0 + 3 of accumulator entries 4 g .. 4 g + 3
      const int n0 = nb + 32 * t + 8 * g + 4 * h;
      const float4 bj = *(const float4*)&bias[n0];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const f32x16& c = acc[u][t];
        uint2 v;
        v.x = pack_bf16(silu(c[4 * g] + bj.x), silu(c[4 * g + 1] + bj.y));
        v.y = pack_bf16(silu(c[4 * g + 2] + bj.z), silu(c[4 * g + 3] + bj.w));
        *(uint2*)&act[wm + 32 * u + r][n0] = v;
      }
    }
  __syncthreads();
}
// The same layer with wave w computing every sample of the block (four
// 32-sample tiles) for neurons 64 w .. +64 (two 32-neuron tiles): each
// weight fragment is read from L2 once per block (not once per wave pair)
// and feeds four matrix-core operations; the weight fragments are loaded
// two K steps ahead.
template <int K>
__device__ __forceinline__ void hidden_layer_w(uint16_t (*act)[LDH], const uint16_t* wt, const float* bias, int nb,
                                               int lane) {
  const int h = lane >> 5, r = lane & 31;
  f32x16 acc[4][2];  // [sample tile][neuron tile]
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int t = 0; t < 2; t++) acc[u][t] = zero16();
  const uint16_t* wrow = wt + (long)(nb + r) * K + 8 * h;
  bf16x8 w1[2], w2[2];
#pragma unroll
  for (int t = 0; t < 2; t++) {
    w1[t] = *(const bf16x8*)&wrow[(long)32 * t * K];
    if (16 < K) w2[t] = *(const bf16x8*)&wrow[(long)32 * t * K + 16];
  }
  for (int kk = 0; kk < K; kk += 16) {
    bf16x8 wc[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      wc[t] = w1[t];
      w1[t] = w2[t];
    }
    if (kk + 32 < K) {
#pragma unroll
      for (int t = 0; t < 2; t++) w2[t] = *(const bf16x8*)&wrow[(long)32 * t * K + kk + 32];
    }
    bf16x8 x[4];
#pragma unroll
    for (int u = 0; u < 4; u++) x[u] = *(const bf16x8*)&act[32 * u + r][kk + 8 * h];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int t = 0; t < 2; t++)
        acc[u][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wc[t], x[u], acc[u][t], 0, 0, 0);
  }
  __syncthreads();  // every wave has read this layer's inputs
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const int n0 = nb + 32 * t + 8 * g + 4 * h;
      const float4 bj = *(const float4*)&bias[n0];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const f32x16& c = acc[u][t];
        uint2 v;
        v.x = pack_bf16(silu(c[4 * g] + bj.x), silu(c[4 * g + 1] + bj.y));
        v.y = pack_bf16(silu(c[4 * g + 2] + bj.z), silu(c[4 * g + 3] + bj.w));
        *(uint2*)&act[32 * u + r][n0] = v;
      }
    }
  __syncthreads();
}
}  // namespace

// two waves per SIMD (at most 256 registers per lane): two blocks per CU
// hide each other's weight loads.  (Measured and dropped: 64-sample blocks
// at four per CU, 6.9 ms against 4.8 on the 8.3 M-sample call, the weight
// reads per sample doubled; a ping-pong schedule over two 64-sample halves
// with one half's SiLU epilogue between the other's matrix operations,
// 6.9 ms: nine barriers per block and no weight prefetch.)
template <bool W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) k_mlp_infer(InferArgs a) {
  __shared__ __align__(16) uint16_t A[INF_M][LDH];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long s0 = (long)blockIdx.x * INF_M;
  // x rows (32 BF16 each) into A[:, 0..31]
  for (int i = tid; i < INF_M * (MLP_INP / 8); i += 256) {
    const int r = i / (MLP_INP / 8), c = (i % (MLP_INP / 8)) * 8;
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; e++) v[e] = 0;
    if (s0 + r < a.n) v = *(const bf16x8*)&a.x[(s0 + r) * MLP_INP + c];
    *(bf16x8*)&A[r][c] = v;
  }
  __syncthreads();
  if (W) {
    const int nb = w * 64;
    hidden_layer_w<MLP_INP>(A, a.w1t, a.params + MLP_B1, nb, lane);
    hidden_layer_w<MLP_H>(A, a.w2t, a.params + MLP_B2, nb, lane);
    hidden_layer_w<MLP_H>(A, a.w3t, a.params + MLP_B3, nb, lane);
    hidden_layer_w<MLP_H>(A, a.w4t, a.params + MLP_B4, nb, lane);
  } else {
    const int wm = (w & 1) * 64, nb = (w >> 1) * 128;
    hidden_layer<MLP_INP>(A, a.w1t, a.params + MLP_B1, wm, nb, lane);
    hidden_layer<MLP_H>(A, a.w2t, a.params + MLP_B2, wm, nb, lane);
    hidden_layer<MLP_H>(A, a.w3t, a.params + MLP_B3, wm, nb, lane);
    hidden_layer<MLP_H>(A, a.w4t, a.params + MLP_B4, wm, nb, lane);
  }
  // output: 2 threads per sample, 128 columns each
  const int smp = tid >> 1, part = tid & 1;
  float s = 0.f;
  for (int j = part * 128; j < part * 128 + 128; j++) s += bf2f(A[smp][j]) * bf2f(a.wbf[MLP_W5 + j]);
  s += __shfl_xor(s, 1, 64);
  if (part == 0 && s0 + smp < a.n) a.y[s0 + smp] = s + a.params[MLP_B5];
}

// ---- launchers --------------------------------------------------------------
static inline unsigned nb256(long n) { return (unsigned)((n + 255) / 256); }

void launch_gemm(const GemmArgs& g, hipStream_t s) {
  dim3 grid((g.N + TS - 1) / TS, (g.M + TS - 1) / TS, g.ksplit);
  // the vectorised kernel where its layout conditions hold
  const bool al = ((uintptr_t)g.A % 16 == 0) && ((uintptr_t)g.B % 16 == 0);
  const bool la = g.sak == 1 ? g.sam % 8 == 0 : (g.sam == 1 && g.sak % 8 == 0);
  const bool lb = g.sbk == 1 ? g.sbn % 8 == 0 : (g.sbn == 1 && g.sbk % 8 == 0);
  if (al && la && lb && g.M % 8 == 0 && g.N % 8 == 0 && g.K % g.ksplit == 0 && (g.K / g.ksplit) % TK2 == 0)
    hipLaunchKernelGGL(k_gemm_v, grid, dim3(256), 0, s, g);
  else
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, g);
}
void launch_splitk_sum(const float* part, int nslab, long slab, long n, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_splitk_sum, dim3(nb256(n)), dim3(256), 0, s, part, nslab, slab, n, out);
}
void launch_gather_batch(const uint16_t* X, const float* Y, long ntot, int B, const int* step, unsigned seed,
                         uint16_t* xb, float* yb, hipStream_t s) {
  (void)seed;
  hipLaunchKernelGGL(k_gather_batch, dim3(nb256(B)), dim3(256), 0, s, X, Y, ntot, B, step, xb, yb);
}
void launch_out_fwd(const uint16_t* act, const uint16_t* w5, const float* b5, int B, float* y, hipStream_t s) {
  hipLaunchKernelGGL(k_out_fwd, dim3((B + 3) / 4), dim3(256), 0, s, act, w5, b5, B, y);
}
void launch_loss_grad(const float* y, const float* t, int B, float* dy, float* loss_hist, const int* step,
                      int mask, hipStream_t s) {
  hipLaunchKernelGGL(k_loss_grad, dim3(1), dim3(1024), 0, s, y, t, B, dy, loss_hist, step, mask);
}
void launch_out_bwd(const uint16_t* act4, const float* z4, const float* dy, const uint16_t* w5, int B,
                    float* gw5, float* gb5, uint16_t* dz4, float* part, hipStream_t s) {
  hipLaunchKernelGGL(k_out_dz, dim3(nb256((long)B * MLP_H)), dim3(256), 0, s, z4, dy, w5, B, dz4);
  // gw5 and gb5 are neighbours in the parameter layout (MLP_B5 = MLP_W5 + MLP_H)
  const int nblk = (B + 63) / 64;
  hipLaunchKernelGGL(k_out_wgrad_part, dim3(nblk), dim3(256), 0, s, act4, dy, B, part);
  hipLaunchKernelGGL(k_splitk_sum, dim3(nb256(MLP_H + 1)), dim3(256), 0, s, part, nblk, (long)(MLP_H + 1),
                     (long)(MLP_H + 1), gw5);
  (void)gb5;
}
void launch_silu_back(const float* da, const float* z, long n, uint16_t* dz, hipStream_t s) {
  hipLaunchKernelGGL(k_silu_back, dim3(nb256(n)), dim3(256), 0, s, da, z, n, dz);
}
void launch_colsum(const uint16_t* dz, int B, int N, float* part, hipStream_t s) {
  (void)N;
  hipLaunchKernelGGL(k_colsum_part, dim3((B + 63) / 64), dim3(MLP_H), 0, s, dz, B, part);
}
void launch_adam(float* w, const float* g, float* m, float* v, uint16_t* wbf, long n, const int* step, float lr,
                 float b1, float b2, float eps, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3(nb256(n)), dim3(256), 0, s, w, g, m, v, wbf, n, step, lr, b1, b2, eps);
}
void launch_step_inc(int* step, hipStream_t s) { hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(1), 0, s, step); }
void launch_f2bf(const float* a, uint16_t* b, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_f2bf, dim3(nb256(n)), dim3(256), 0, s, a, b, n);
}
void launch_transpose_bf(const uint16_t* a, int rows, int cols, uint16_t* at, hipStream_t s) {
  hipLaunchKernelGGL(k_transpose_bf, dim3(nb256((long)rows * cols)), dim3(256), 0, s, a, rows, cols, at);
}
void launch_fill_f32(float* p, long n, float v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_f32, dim3(nb256(n)), dim3(256), 0, s, p, n, v);
}
void launch_mlp_infer(const InferArgs& a, hipStream_t s) {
  if (a.n <= 0) return;
  if (a.variant == 1)
    hipLaunchKernelGGL(k_mlp_infer<true>, dim3((unsigned)((a.n + INF_M - 1) / INF_M)), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_mlp_infer<false>, dim3((unsigned)((a.n + INF_M - 1) / INF_M)), dim3(256), 0, s, a);
}
