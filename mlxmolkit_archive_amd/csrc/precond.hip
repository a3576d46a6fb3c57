// Learned per-atom initial inverse Hessian for the BFGS minimizer, gfx950.
//
// A small MLP maps per-atom features (MMFF type one-hot, degree, ring flags, summed bond
// and angle force constants; MLP_IN = 104) through two hidden layers of width 64 (ReLU)
// to 6 numbers, the Cholesky factor L of a 3x3 SPD block:
//   L = [[sp(o0), 0, 0], [o1, sp(o2), 0], [o3, o4, sp(o5)]],  sp(v) = softplus(v) + 1e-3,
//   block = L L^T.
// Weights are row-major fp32 [out][in] (W1 64x104, W2 64x64, W3 6x64) plus biases.
//
// Kernels:
//   k_mlp_infer_bf16  inference on MFMA (v_mfma_f32_32x32x8_bf16, FP32 accumulation): one
//                     wave per 32 atoms; activations are re-laid out through LDS between
//                     layers; writes the 9 floats of each block
//   k_mlp_fwd_f32     training forward in FP32, one lane per atom; keeps pre-activations
//   k_mlp_bwd_delta   per-atom backward deltas of the block loss sum ||L L^T - T||_F^2
//   k_mlp_wgrad       weight / bias gradients, one lane per parameter summing over atoms in
//                     a fixed order (deterministic, no atomics)
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "mmff_hip.h"

namespace {

constexpr int MLP_IN = 104;
constexpr int MLP_KPAD = 112;  // MLP_IN rounded up to the MFMA k-step of 8
constexpr int MLP_H = 64;
constexpr int MLP_OUT = 6;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ short to_bf16(float f) {
  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (short)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

__device__ __forceinline__ float softplus_eps(float v) {
  return (v > 20.f ? v : log1pf(expf(v))) + 1e-3f;
}
__device__ __forceinline__ float sigmoidf(float v) { return 1.f / (1.f + expf(-v)); }

// 3x3 block L L^T from the 6 outputs
__device__ __forceinline__ void chol_block(const float* o, float* blk) {
  const float l00 = softplus_eps(o[0]), l10 = o[1], l11 = softplus_eps(o[2]);
  const float l20 = o[3], l21 = o[4], l22 = softplus_eps(o[5]);
  const float b00 = l00 * l00, b01 = l00 * l10, b02 = l00 * l20;
  const float b11 = l10 * l10 + l11 * l11, b12 = l10 * l20 + l11 * l21;
  const float b22 = l20 * l20 + l21 * l21 + l22 * l22;
  blk[0] = b00; blk[1] = b01; blk[2] = b02;
  blk[3] = b01; blk[4] = b11; blk[5] = b12;
  blk[6] = b02; blk[7] = b12; blk[8] = b22;
}

struct MlpW {
  const float *w1, *b1, *w2, *b2, *w3, *b3;
};

// ------------------------------------------------------------------------------------
// inference on MFMA. LDS: W1 bf16 [64][112], W2 bf16 [64][64], W3 bf16 [32][64] (rows 6..31
// zero), plus per wave an activation tile bf16 [32][112] and an output tile f32 [32][8].
// v_mfma_f32_32x32x8_bf16 operand layout (wave64): lane l supplies A[l%32][4*(l/32)+0..3]
// and B[4*(l/32)+0..3][l%32]; it receives D[8*(r/4) + 4*(l/32) + r%4][l%32], r = 0..15.
constexpr int INF_WAVES = 4;
constexpr int INF_TILE_SHORTS = 32 * MLP_KPAD + 32 * 64 + 32 * 8 * 2;  // act, act2, out (f32)
constexpr size_t INF_LDS = (size_t)(64 * MLP_KPAD + 64 * 64 + 32 * 64) * 2 + (size_t)INF_WAVES * INF_TILE_SHORTS * 2;

__device__ __forceinline__ bf16x4 ld4(const short* p) {
  bf16x4 v;
  v.x = p[0];
  v.y = p[1];
  v.z = p[2];
  v.w = p[3];
  return v;
}

// D (32 x 32 block at column offset n0 of an [out][in] weight matrix Wt) += A (32 x K) * Wt^T
template <int K>
__device__ __forceinline__ f32x16 mfma_tile(const short* act, int act_ld, const short* wt, int w_ld, int n0, int lane) {
  f32x16 acc = {0.f};
  const int r = lane & 31, kg = (lane >> 5) * 4;
#pragma unroll
  for (int k = 0; k < K; k += 8) {
    const bf16x4 a = ld4(act + r * act_ld + k + kg);
    const bf16x4 b = ld4(wt + (n0 + r) * w_ld + k + kg);
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
  return acc;
}

// bias + ReLU, then store the 32 x 32 tile as bf16 activations at columns n0.. of act
__device__ __forceinline__ void store_relu(f32x16 acc, const float* bias, int n0, short* act, int act_ld, int lane) {
  const int col = lane & 31, rb = (lane >> 5) * 4;
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int row = 8 * (q >> 2) + rb + (q & 3);
    act[row * act_ld + n0 + col] = to_bf16(fmaxf(acc[q] + bias[n0 + col], 0.f));
  }
}

__global__ void __launch_bounds__(64 * INF_WAVES) k_mlp_infer_bf16(MlpW W, const float* __restrict__ feat, int n_atoms,
                                                                   float* __restrict__ blocks) {
  extern __shared__ __align__(16) unsigned char smem[];
  short* w1 = (short*)smem;
  short* w2 = w1 + 64 * MLP_KPAD;
  short* w3 = w2 + 64 * 64;
  short* tiles = w3 + 32 * 64;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int i = tid; i < 64 * MLP_KPAD; i += blockDim.x) {
    const int o = i / MLP_KPAD, k = i % MLP_KPAD;
    w1[i] = to_bf16(k < MLP_IN ? W.w1[o * MLP_IN + k] : 0.f);
  }
  for (int i = tid; i < 64 * 64; i += blockDim.x) w2[i] = to_bf16(W.w2[i]);
  for (int i = tid; i < 32 * 64; i += blockDim.x) w3[i] = to_bf16(i < MLP_OUT * 64 ? W.w3[i] : 0.f);
  __syncthreads();
  short* act = tiles + wv * INF_TILE_SHORTS;  // bf16 [32][112], later [32][64]
  short* act2 = act + 32 * MLP_KPAD;            // bf16 [32][64]
  float* outt = (float*)(act2 + 32 * 64);       // f32 [32][8]
  const int n_tiles = (n_atoms + 31) / 32;
  for (int t = blockIdx.x * INF_WAVES + wv; t < n_tiles; t += gridDim.x * INF_WAVES) {
    const int a0 = t * 32;
    // features of 32 atoms -> bf16 tile (rows past n_atoms are zero)
    for (int i = lane; i < 32 * MLP_KPAD; i += 64) {
      const int r = i / MLP_KPAD, k = i % MLP_KPAD;
      const int a = a0 + r;
      act[i] = to_bf16(a < n_atoms && k < MLP_IN ? feat[(size_t)a * MLP_IN + k] : 0.f);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    // layer 1: [32 x 112] x W1^T -> [32 x 64]
    f32x16 h0 = mfma_tile<MLP_KPAD>(act, MLP_KPAD, w1, MLP_KPAD, 0, lane);
    f32x16 h1 = mfma_tile<MLP_KPAD>(act, MLP_KPAD, w1, MLP_KPAD, 32, lane);
    __builtin_amdgcn_wave_barrier();
    store_relu(h0, W.b1, 0, act, 64, lane);  // reuse the tile as [32][64]
    store_relu(h1, W.b1, 32, act, 64, lane);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    // layer 2: [32 x 64] x W2^T -> [32 x 64]
    h0 = mfma_tile<64>(act, 64, w2, 64, 0, lane);
    h1 = mfma_tile<64>(act, 64, w2, 64, 32, lane);
    store_relu(h0, W.b2, 0, act2, 64, lane);
    store_relu(h1, W.b2, 32, act2, 64, lane);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    // layer 3: [32 x 64] x W3^T -> [32 x 32], columns 0..5 used
    const f32x16 o = mfma_tile<64>(act2, 64, w3, 64, 0, lane);
    {
      const int col = lane & 31, rb = (lane >> 5) * 4;
      if (col < MLP_OUT) {
#pragma unroll
        for (int q = 0; q < 16; q++) outt[(8 * (q >> 2) + rb + (q & 3)) * 8 + col] = o[q] + W.b3[col];
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if (lane < 32 && a0 + lane < n_atoms) {
      float blk[9];
      chol_block(outt + lane * 8, blk);
      float* dst = blocks + (size_t)(a0 + lane) * 9;
#pragma unroll
      for (int q = 0; q < 9; q++) dst[q] = blk[q];
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// ------------------------------------------------------------------------------------
// training forward (FP32), one lane per atom. act: per atom [z1 (64), z2 (64), o (6)]
// pre-activations; loss_out: per-atom ||L L^T - T||_F^2 (T = target blocks, 9 per atom).
constexpr int ACT_LD = 64 + 64 + 8;

__global__ void __launch_bounds__(256) k_mlp_fwd_f32(MlpW W, const float* __restrict__ feat, int n_atoms,
                                                     float* __restrict__ act, const float* __restrict__ target,
                                                     float* __restrict__ blocks, float* __restrict__ loss_out) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_atoms) return;
  const float* x = feat + (size_t)a * MLP_IN;
  float* z = act + (size_t)a * ACT_LD;
  float h1[64], h2[64], o[8];
  for (int j = 0; j < 64; j++) {
    float s = W.b1[j];
    for (int k = 0; k < MLP_IN; k++) s = fmaf(W.w1[j * MLP_IN + k], x[k], s);
    z[j] = s;
    h1[j] = fmaxf(s, 0.f);
  }
  for (int j = 0; j < 64; j++) {
    float s = W.b2[j];
    for (int k = 0; k < 64; k++) s = fmaf(W.w2[j * 64 + k], h1[k], s);
    z[64 + j] = s;
    h2[j] = fmaxf(s, 0.f);
  }
  for (int j = 0; j < MLP_OUT; j++) {
    float s = W.b3[j];
    for (int k = 0; k < 64; k++) s = fmaf(W.w3[j * 64 + k], h2[k], s);
    z[128 + j] = s;
    o[j] = s;
  }
  float blk[9];
  chol_block(o, blk);
  if (blocks)
    for (int q = 0; q < 9; q++) blocks[(size_t)a * 9 + q] = blk[q];
  if (loss_out && target) {
    float l = 0.f;
    for (int q = 0; q < 9; q++) {
      const float d = blk[q] - target[(size_t)a * 9 + q];
      l = fmaf(d, d, l);
    }
    loss_out[a] = l;
  }
}

// per-atom deltas: delta[a] = [dz1 (64), dz2 (64), do (6)] of sum_a w_a ||P_a - T_a||^2
__global__ void __launch_bounds__(256) k_mlp_bwd_delta(MlpW W, int n_atoms, const float* __restrict__ act,
                                                       const float* __restrict__ target,
                                                       const float* __restrict__ weight, float* __restrict__ delta) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_atoms) return;
  const float* z = act + (size_t)a * ACT_LD;
  float* dl = delta + (size_t)a * ACT_LD;
  const float* o = z + 128;
  float blk[9];
  chol_block(o, blk);
  const float w = weight ? weight[a] : 1.f;
  float G[9];  // dLoss/dP
  for (int q = 0; q < 9; q++) G[q] = 2.f * w * (blk[q] - target[(size_t)a * 9 + q]);
  // P = L L^T: dLoss/dL = (G + G^T) L
  const float L[9] = {softplus_eps(o[0]), 0.f, 0.f, o[1], softplus_eps(o[2]), 0.f, o[3], o[4], softplus_eps(o[5])};
  float S[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) S[i * 3 + j] = G[i * 3 + j] + G[j * 3 + i];
  float dL[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      float s = 0.f;
      for (int k = 0; k < 3; k++) s = fmaf(S[i * 3 + k], L[k * 3 + j], s);
      dL[i * 3 + j] = s;
    }
  float dout[6];
  dout[0] = dL[0] * sigmoidf(o[0]);
  dout[1] = dL[3];
  dout[2] = dL[4] * sigmoidf(o[2]);
  dout[3] = dL[6];
  dout[4] = dL[7];
  dout[5] = dL[8] * sigmoidf(o[5]);
  for (int j = 0; j < 6; j++) dl[128 + j] = dout[j];
  for (int k = 0; k < 64; k++) {
    float s = 0.f;
    for (int j = 0; j < 6; j++) s = fmaf(W.w3[j * 64 + k], dout[j], s);
    dl[64 + k] = z[64 + k] > 0.f ? s : 0.f;
  }
  for (int k = 0; k < 64; k++) {
    float s = 0.f;
    for (int j = 0; j < 64; j++) s = fmaf(W.w2[j * 64 + k], dl[64 + j], s);
    dl[k] = z[k] > 0.f ? s : 0.f;
  }
}

// parameter gradients: lane p owns one parameter and sums its contributions over atoms in
// atom order. Parameter order: w1 (64x104), b1, w2 (64x64), b2, w3 (6x64), b3.
constexpr int NP_W1 = 64 * MLP_IN, NP_B1 = 64, NP_W2 = 64 * 64, NP_B2 = 64, NP_W3 = MLP_OUT * 64, NP_B3 = MLP_OUT;
constexpr int MLP_NPARAM = NP_W1 + NP_B1 + NP_W2 + NP_B2 + NP_W3 + NP_B3;

__global__ void __launch_bounds__(256) k_mlp_wgrad(const float* __restrict__ feat, int n_atoms,
                                                   const float* __restrict__ act, const float* __restrict__ delta,
                                                   float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  int q = p;
  int dsel, dj, src, sk;  // delta column, input source (0 feat, 1 relu(z1), 2 relu(z2), 3 bias)
  if (q < NP_W1) {
    dsel = 0; dj = q / MLP_IN; src = 0; sk = q % MLP_IN;
  } else if ((q -= NP_W1) < NP_B1) {
    dsel = 0; dj = q; src = 3; sk = 0;
  } else if ((q -= NP_B1) < NP_W2) {
    dsel = 64; dj = q / 64; src = 1; sk = q % 64;
  } else if ((q -= NP_W2) < NP_B2) {
    dsel = 64; dj = q; src = 3; sk = 0;
  } else if ((q -= NP_B2) < NP_W3) {
    dsel = 128; dj = q / 64; src = 2; sk = q % 64;
  } else {
    q -= NP_W3;
    dsel = 128; dj = q; src = 3; sk = 0;
  }
  double s = 0.0;
  for (int a = 0; a < n_atoms; a++) {
    const float d = delta[(size_t)a * ACT_LD + dsel + dj];
    float v;
    if (src == 0) v = feat[(size_t)a * MLP_IN + sk];
    else if (src == 1) v = fmaxf(act[(size_t)a * ACT_LD + sk], 0.f);
    else if (src == 2) v = fmaxf(act[(size_t)a * ACT_LD + 64 + sk], 0.f);
    else v = 1.f;
    s += (double)d * v;
  }
  grad[p] = (float)s;
}

MlpW mlp_weights(const float* params) {
  MlpW W;
  W.w1 = params;
  W.b1 = W.w1 + NP_W1;
  W.w2 = W.b1 + NP_B1;
  W.b2 = W.w2 + NP_W2;
  W.w3 = W.b2 + NP_B2;
  W.b3 = W.w3 + NP_W3;
  return W;
}

}  // namespace

// defined with the rest of the error state in mmff_hip.hip
int mmffhip_set_error(int code, const char* msg);
#define perr mmffhip_set_error

extern "C" {

int mmffhip_mlp_dims(int32_t* n_in, int32_t* n_hidden, int32_t* n_out, int32_t* n_params) {
  if (n_in) *n_in = MLP_IN;
  if (n_hidden) *n_hidden = MLP_H;
  if (n_out) *n_out = MLP_OUT;
  if (n_params) *n_params = MLP_NPARAM;
  return MMFFHIP_OK;
}

int mmffhip_mlp_infer(const float* params, const float* feat, int32_t n_atoms, float* blocks, void* stream) {
  if (n_atoms < 0) return perr(MMFFHIP_ERR_ARG, "n_atoms < 0");
  if (n_atoms == 0) return MMFFHIP_OK;
  if (!params || !feat || !blocks) return perr(MMFFHIP_ERR_ARG, "params, feat and blocks are required");
  const MlpW W = mlp_weights(params);
  const int tiles = (n_atoms + 31) / 32;
  const int grid = (tiles + INF_WAVES - 1) / INF_WAVES < 1024 ? (tiles + INF_WAVES - 1) / INF_WAVES : 1024;
  if (INF_LDS > 65536 &&
      hipFuncSetAttribute((const void*)&k_mlp_infer_bf16, hipFuncAttributeMaxDynamicSharedMemorySize, (int)INF_LDS) !=
          hipSuccess)
    return perr(MMFFHIP_ERR_HIP, "hipFuncSetAttribute failed");
  hipLaunchKernelGGL(k_mlp_infer_bf16, dim3(grid), dim3(64 * INF_WAVES), INF_LDS, (hipStream_t)stream, W, feat,
                     n_atoms, blocks);
  return hipGetLastError() == hipSuccess ? MMFFHIP_OK : perr(MMFFHIP_ERR_HIP, "launch failed");
}

int mmffhip_mlp_workspace(int32_t n_atoms, size_t* bytes) {
  if (!bytes || n_atoms < 0) return perr(MMFFHIP_ERR_ARG, "bad arguments");
  *bytes = (size_t)n_atoms * ACT_LD * 4 * 2 + (size_t)n_atoms * 4 + 256;
  return MMFFHIP_OK;
}

int mmffhip_mlp_forward_backward(const float* params, const float* feat, int32_t n_atoms, const float* target,
                                 const float* weight, float* blocks, float* loss, float* grad, void* workspace,
                                 size_t workspace_bytes, void* stream) {
  if (n_atoms <= 0) return perr(MMFFHIP_ERR_ARG, "n_atoms must be positive");
  if (!params || !feat || !target || !workspace) return perr(MMFFHIP_ERR_ARG, "params, feat, target, workspace required");
  size_t need = 0;
  mmffhip_mlp_workspace(n_atoms, &need);
  if (workspace_bytes < need) return perr(MMFFHIP_ERR_SIZE, "workspace too small");
  const MlpW W = mlp_weights(params);
  float* act = (float*)workspace;
  float* delta = act + (size_t)n_atoms * ACT_LD;
  float* lbuf = delta + (size_t)n_atoms * ACT_LD;
  hipStream_t s = (hipStream_t)stream;
  const int g = (n_atoms + 255) / 256;
  hipLaunchKernelGGL(k_mlp_fwd_f32, dim3(g), dim3(256), 0, s, W, feat, n_atoms, act, target, blocks,
                     loss ? loss : lbuf);
  if (grad) {
    hipLaunchKernelGGL(k_mlp_bwd_delta, dim3(g), dim3(256), 0, s, W, n_atoms, act, target, weight, delta);
    hipLaunchKernelGGL(k_mlp_wgrad, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, feat, n_atoms, act, delta, grad);
  }
  return hipGetLastError() == hipSuccess ? MMFFHIP_OK : perr(MMFFHIP_ERR_HIP, "launch failed");
}

}  // extern "C"
