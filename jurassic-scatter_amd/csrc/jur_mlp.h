// jur_mlp.h -- the scattering-source surrogate: a 24 -> 4 x 256 (SiLU) -> 1
// multilayer perceptron with BF16 weights and activations and FP32
// accumulation on the CDNA4 matrix cores (v_mfma_f32_32x32x16_bf16).
//
// Parameter layout (one FP32 master array, one BF16 copy of the same layout,
// gradients and Adam moments alike):
//   W1 [32][256] (rows 24..31 stay zero: the 24 inputs are padded to 32)
//   b1 [256], W2 [256][256], b2, W3, b3, W4, b4, W5 [256], b5 [1]
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// FP32 -> BF16, round to nearest even (NaN stays NaN), host and device.
__host__ __device__ inline uint16_t mlp_f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

#define MLP_IN 24
#define MLP_INP 32   // padded input width
#define MLP_H 256
#define MLP_NHID 4

enum {
  MLP_W1 = 0,
  MLP_B1 = MLP_W1 + MLP_INP * MLP_H,
  MLP_W2 = MLP_B1 + MLP_H,
  MLP_B2 = MLP_W2 + MLP_H * MLP_H,
  MLP_W3 = MLP_B2 + MLP_H,
  MLP_B3 = MLP_W3 + MLP_H * MLP_H,
  MLP_W4 = MLP_B3 + MLP_H,
  MLP_B4 = MLP_W4 + MLP_H * MLP_H,
  MLP_W5 = MLP_B4 + MLP_H,
  MLP_B5 = MLP_W5 + MLP_H,
  MLP_NPARAM = MLP_B5 + 1
};

// C[M][N] (+)= A[M][K] B[K][N]; A, B in BF16 with arbitrary strides (a
// transposed operand is a stride swap).  M, N, K multiples of 32.  Split-K:
// gridDim.z slabs each write their partial product to C + z * cslab, summed
// by launch_splitk_sum.  Epilogue 1 adds bias[n], stores z = acc + bias to C
// and silu(z) as BF16 to act.
struct GemmArgs {
  int M, N, K, ksplit;
  const uint16_t* A;
  long sam, sak;
  const uint16_t* B;
  long sbk, sbn;
  float* C;
  long ldc, cslab;
  int epi;
  const float* bias;
  uint16_t* act;
  long ldact;
};

void launch_gemm(const GemmArgs& g, hipStream_t s);
void launch_splitk_sum(const float* part, int nslab, long slab, long n, float* out, hipStream_t s);

// Training-step pieces (batch rows B, hidden H = 256).
void launch_gather_batch(const uint16_t* X, const float* Y, long ntot, int B, const int* step, unsigned seed,
                         uint16_t* xb, float* yb, hipStream_t s);
void launch_out_fwd(const uint16_t* act, const uint16_t* w5, const float* b5, int B, float* y, hipStream_t s);
void launch_loss_grad(const float* y, const float* t, int B, float* dy, float* loss_hist, const int* step,
                      int mask, hipStream_t s);
void launch_out_bwd(const uint16_t* act4, const float* z4, const float* dy, const uint16_t* w5, int B,
                    float* gw5, float* gb5, uint16_t* dz4, float* part, hipStream_t s);
void launch_silu_back(const float* da, const float* z, long n, uint16_t* dz, hipStream_t s);
void launch_colsum(const uint16_t* dz, int B, int N, float* out, hipStream_t s);
void launch_adam(float* w, const float* g, float* m, float* v, uint16_t* wbf, long n, const int* step, float lr,
                 float b1, float b2, float eps, hipStream_t s);
void launch_step_inc(int* step, hipStream_t s);
void launch_f2bf(const float* a, uint16_t* b, long n, hipStream_t s);
void launch_transpose_bf(const uint16_t* a, int rows, int cols, uint16_t* at, hipStream_t s);
void launch_fill_f32(float* p, long n, float v, hipStream_t s);

// Fused inference: y[n] for x [n][32] (BF16, normalised features), with
// transposed BF16 weights (W1t [256][32], W2t..W4t [256][256]), FP32 biases
// and the BF16 output weights: the same arithmetic as the training forward
// pass.
struct InferArgs {
  int n;
  const uint16_t* x;
  const uint16_t* w1t;
  const uint16_t* w2t;
  const uint16_t* w3t;
  const uint16_t* w4t;
  const float* params;   // FP32 master (biases, b5)
  const uint16_t* wbf;   // BF16 copy (W5)
  float* y;
  int variant;           // 1: a wave takes all 128 samples of its block x 64 neurons; 0: 64 samples x 128 neurons
};
void launch_mlp_infer(const InferArgs& a, hipStream_t s);
