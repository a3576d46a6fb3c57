// Start geometries on the device for gfx950: seeded restraint embedding, one workgroup per
// conformer, so a batch's start geometries are made where the minimizer reads them.
//
// Stands in for the reference pipeline's device DG -> 4D->3D -> ETK stages
// (mlxmolkit/conformer_pipeline_v2.py:176-290), which are out of scope; the method is the
// package's host embedder (csrc/embed.c) moved onto the GPU:
//   * coordinates start uniformly in a 4D box of half-width 1.5 N^(1/3) + 1 A, drawn from a
//     counter-based splitmix64 stream per conformer (draw k is the generator's k-th output,
//     so lanes draw independently and the draws equal the host embedder's);
//   * restraint energy: equality springs 100 (d - d0)^2 for 1-2, 1-3 and ring targets, and
//     lower-bound walls 10 (d_min - d)^2 for every other pair closer than d_min;
//   * three relaxations by gradient descent with a 0.3 A per-atom step cap and an adaptive
//     step (x1.2 on an accepted step, x0.5 on a rejected one): 250 steps in 4D, 150 steps
//     with the fourth coordinate penalised (weight 2), then 200 steps in 3D after dropping it;
//   * N(0, noise) is added to every coordinate (Box-Muller on the same stream).
// Arithmetic is FP64. Gradients are summed in 64-bit fixed point (2^-32) in LDS and energies
// are reduced in a fixed order, so a conformer's result does not depend on thread timing.
// Coordinates and gradient sums sit in LDS as four planes (one per dimension, stride `ps`
// elements), so the lanes of a wave, which take consecutive restraints and hence mostly
// consecutive second atoms, read consecutive 8-byte words.
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

#include "mmff_hip.h"

__attribute__((visibility("hidden"))) int mmffhip_set_error(int code, const char* msg);

namespace {

constexpr double EFX_MAGIC = 1572864.0;  // 1.5 * 2^20: FP64 sums in [2^20, 2^21) step 2^-32
constexpr double EFX_CAP = 524000.0;     // |v| < 2^19
constexpr double EFX_UNIT = 2.3283064365386962890625e-10;  // 2^-32

// v rounded to the nearest multiple of 2^-32, as a two's-complement integer (exact negation)
__device__ __forceinline__ unsigned long long efx(double v) {
  const double c = (v == v) ? fmin(fmax(v, -EFX_CAP), EFX_CAP) : 0.0;
  return (unsigned long long)(__double_as_longlong(c + EFX_MAGIC) - __double_as_longlong(EFX_MAGIC));
}
__device__ __forceinline__ double efx_val(unsigned long long a) { return (double)(long long)a * EFX_UNIT; }

// k-th output (k = 0, 1, ...) of splitmix64 seeded with `seed`, as a double in [0, 1)
__device__ __forceinline__ double unif(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  constexpr int NW = NT / 64;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();  // earlier readers of red are done
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = red[0];
  for (int k = 1; k < NW; k++) s += red[k];
  return s;
}

struct EmbedMol {
  int n_atoms, r_off, n_restr;
};

// Restraint energy (this lane's part) and gradient (into g, zeroed by the caller) at x
// (coordinate c of atom a at [c * ps + a]; the first `dim` coordinates are used).
template <int NT>
__device__ double restraint_eg(const uint32_t* rij, const float* rd, const EmbedMol& m, const double* x,
                               unsigned long long* g, int ps, int dim, double w4) {
  double e = 0.0;
  for (int t = threadIdx.x; t < m.n_restr; t += NT) {
    const uint32_t ij = rij[t];
    const int i = (int)(ij & 0xffffu), j = (int)(ij >> 16);
    const float dd = rd[t];
    const bool wall = dd < 0.0f;
    const double d0 = fabs((double)dd);
    double dv[4] = {0.0, 0.0, 0.0, 0.0}, d2 = 0.0;
    for (int c = 0; c < dim; c++) {
      dv[c] = x[c * ps + i] - x[c * ps + j];
      d2 += dv[c] * dv[c];
    }
    if (wall && d2 + 1e-12 >= d0 * d0) continue;  // an inactive wall costs no square root
    const double d = sqrt(d2 + 1e-12);
    if (wall && d >= d0) continue;
    const double diff = d - d0, k = wall ? 10.0 : 100.0;
    e += k * diff * diff;
    const double f = 2.0 * k * diff / d;
    for (int c = 0; c < dim; c++) {
      const unsigned long long q = efx(f * dv[c]);
      atomicAdd(&g[c * ps + i], q);
      atomicSub(&g[c * ps + j], q);
    }
  }
  if (dim == 4 && w4 > 0.0) {
    for (int a = threadIdx.x; a < m.n_atoms; a += NT) {
      const double w = x[3 * ps + a];
      e += w4 * w * w;
      atomicAdd(&g[3 * ps + a], efx(2.0 * w4 * w));
    }
  }
  return e;
}

// Gradient descent with a per-atom step cap and an adaptive step; x/g and xt/gt swap on an
// accepted step (every lane takes the same branch: the sums are identical on all lanes).
template <int NT>
__device__ void relax(const uint32_t* rij, const float* rd, const EmbedMol& m, double*& x, unsigned long long*& g,
                      double*& xt, unsigned long long*& gt, double* red, int ps, int dim, double w4, int steps) {
  const int n4 = 4 * ps;
  double step = 0.002;
  for (int i = threadIdx.x; i < n4; i += NT) g[i] = 0ull;
  __syncthreads();
  double e = block_sum<NT>(restraint_eg<NT>(rij, rd, m, x, g, ps, dim, w4), red);
  for (int s = 0; s < steps; s++) {
    for (int a = threadIdx.x; a < m.n_atoms; a += NT) {
      double gv[4], gn = 0.0;
      for (int c = 0; c < dim; c++) {
        gv[c] = efx_val(g[c * ps + a]);
        gn += gv[c] * gv[c];
      }
      gn = sqrt(gn);
      double sc = step;
      if (gn * sc > 0.3) sc = 0.3 / gn;
      for (int c = 0; c < dim; c++) xt[c * ps + a] = x[c * ps + a] - sc * gv[c];
      for (int c = dim; c < 4; c++) xt[c * ps + a] = x[c * ps + a];
    }
    for (int i = threadIdx.x; i < n4; i += NT) gt[i] = 0ull;
    __syncthreads();
    const double et = block_sum<NT>(restraint_eg<NT>(rij, rd, m, xt, gt, ps, dim, w4), red);
    if (et <= e) {
      double* tx = x;
      x = xt;
      xt = tx;
      unsigned long long* tg = g;
      g = gt;
      gt = tg;
      e = et;
      step *= 1.2;
    } else {
      step = fmax(step * 0.5, 1e-7);
    }
    __syncthreads();
  }
}

template <int NT>
__global__ void __launch_bounds__(NT) k_embed(const int32_t* desc, const uint32_t* rij, const float* rd,
                                              const int32_t* conf_mol, const int32_t* conf_off,
                                              const uint64_t* seeds, int n_confs, int max_atoms, float noise,
                                              float* x_out) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int c = blockIdx.x;
  if (c >= n_confs) return;
  const int ps = max_atoms;  // plane stride
  const int mx4 = 4 * ps;
  double* red = (double*)smem;
  double* xa = red + 16;
  double* xb = xa + mx4;
  unsigned long long* ga = (unsigned long long*)(xb + mx4);
  unsigned long long* gb = ga + mx4;
  const int32_t* d = desc + 4 * conf_mol[c];
  EmbedMol m{d[0], d[1], d[2]};
  const uint32_t* r_ij = rij + m.r_off;
  const float* r_d = rd + m.r_off;
  const uint64_t seed = seeds[c];
  const int n = m.n_atoms;
  const double box = 1.5 * cbrt((double)n) + 1.0;
  // draw k = 4a + c is coordinate c of atom a (the host embedder's order)
  for (int i = threadIdx.x; i < 4 * n; i += NT) xa[(i & 3) * ps + (i >> 2)] = (unif(seed, (uint64_t)i) * 2.0 - 1.0) * box;
  __syncthreads();
  double *x = xa, *xt = xb;
  unsigned long long *g = ga, *gt = gb;
  relax<NT>(r_ij, r_d, m, x, g, xt, gt, red, ps, 4, 0.0, 250);
  relax<NT>(r_ij, r_d, m, x, g, xt, gt, red, ps, 4, 2.0, 150);
  for (int a = threadIdx.x; a < n; a += NT) x[3 * ps + a] = 0.0;  // drop the fourth coordinate
  __syncthreads();
  relax<NT>(r_ij, r_d, m, x, g, xt, gt, red, ps, 3, 0.0, 200);
  float* o = x_out + conf_off[c];
  const uint64_t base = 4 * (uint64_t)n;
  for (int q = threadIdx.x; q < 3 * n; q += NT) {
    double u1 = unif(seed, base + 2 * (uint64_t)q), u2 = unif(seed, base + 2 * (uint64_t)q + 1);
    if (u1 < 1e-300) u1 = 1e-300;
    const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    o[q] = (float)(x[(q % 3) * ps + q / 3] + (double)noise * z);
  }
}

size_t embed_lds(int max_atoms) { return 16 * 8 + (size_t)4 * max_atoms * (8 + 8 + 8 + 8); }

}  // namespace

extern "C" {

int mmffhip_embed(const int32_t* mol_desc, const uint32_t* restr_ij, const float* restr_d, const int32_t* conf_mol,
                  const int32_t* conf_off, const uint64_t* seeds, int32_t n_confs, int32_t max_atoms, float noise,
                  float* x_out, void* stream) {
  if (n_confs < 0 || max_atoms < 0) return mmffhip_set_error(MMFFHIP_ERR_ARG, "negative batch size");
  if (n_confs == 0) return MMFFHIP_OK;
  if (!mol_desc || !restr_ij || !restr_d || !conf_mol || !conf_off || !seeds || !x_out)
    return mmffhip_set_error(MMFFHIP_ERR_ARG, "mmffhip_embed: a required pointer is NULL");
  if (max_atoms < 1 || max_atoms > 65535)
    return mmffhip_set_error(MMFFHIP_ERR_SIZE, "mmffhip_embed: max_atoms must be in [1, 65535]");
  const size_t lds = embed_lds(max_atoms);
  int dev = 0;
  hipDeviceProp_t p;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess)
    return mmffhip_set_error(MMFFHIP_ERR_DEVICE, "mmffhip_embed: no HIP device");
  if (lds > (size_t)p.sharedMemPerBlock)
    return mmffhip_set_error(MMFFHIP_ERR_SIZE, "mmffhip_embed: molecule too large for one workgroup's LDS");
  if (lds > 65536 && hipFuncSetAttribute((const void*)&k_embed<128>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)lds) != hipSuccess)
    return mmffhip_set_error(MMFFHIP_ERR_HIP, "mmffhip_embed: cannot raise the LDS limit");
  hipLaunchKernelGGL((k_embed<128>), dim3(n_confs), dim3(128), lds, (hipStream_t)stream, mol_desc, restr_ij,
                     restr_d, conf_mol, conf_off, seeds, n_confs, max_atoms, noise, x_out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mmffhip_set_error(MMFFHIP_ERR_HIP, (std::string("mmffhip_embed: ") + hipGetErrorString(e)).c_str());
  return MMFFHIP_OK;
}

}  // extern "C"
