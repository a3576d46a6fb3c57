// Batched MMFF94 conformer minimization for gfx950 (MI355X): kernels + extern "C" ABI.
//
// Execution model (one conformer per workgroup, persistent grid with a work queue):
//   * positions, gradient accumulators and BFGS/L-BFGS vectors live in LDS;
//   * the dense inverse Hessian lives in a per-workgroup global slot as symmetric 4x4 blocks,
//     one per pair of row quads, each read once per iteration for both halves of H v (the
//     default), or in full storage read as coalesced 16-byte column quads, or in LDS as a
//     packed triangle with 16-byte aligned columns (n = 3N up to 264 at one workgroup per CU);
//   * term energies and gradients are FP32; each lane's energy is accumulated in FP64 and
//     reduced with wave shuffles + one LDS exchange, so the Armijo test and reported
//     energies are FP64 sums; gradients are summed exactly in 64-bit fixed point with LDS
//     atomics (order-independent, bit-reproducible).
// The control flow of each optimizer restates the reference kernel it replaces; the
// file:line citations are on each kernel.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "mmff_hip.h"
#include "mmff_terms.h"

using namespace mmff;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(MMFFHIP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// ------------------------------------------------------------------------------------
// device-side view of the batch
struct KBatch {
  const int32_t* mol_desc;
  const uint32_t* bond_ij;
  const float2* bond_p;
  const uint2* ang_idx;
  const float2* ang_p;
  const uint2* sb_idx;
  const float4* sb_p;
  const float* sb_t0;
  const uint2* oop_idx;
  const float* oop_k;
  const uint2* tor_idx;
  const float4* tor_p;
  const uint32_t* pair_ij;
  const float4* pair_p;
  const int32_t* conf_mol;
  const int32_t* conf_off;
  int n_mols, n_confs, max_atoms;
};

KBatch to_kbatch(const mmffhip_batch* b) {
  KBatch k;
  k.mol_desc = b->mol_desc;
  k.bond_ij = b->bond_ij;
  k.bond_p = (const float2*)b->bond_p;
  k.ang_idx = (const uint2*)b->ang_idx;
  k.ang_p = (const float2*)b->ang_p;
  k.sb_idx = (const uint2*)b->sb_idx;
  k.sb_p = (const float4*)b->sb_p;
  k.sb_t0 = b->sb_t0;
  k.oop_idx = (const uint2*)b->oop_idx;
  k.oop_k = b->oop_k;
  k.tor_idx = (const uint2*)b->tor_idx;
  k.tor_p = (const float4*)b->tor_p;
  k.pair_ij = b->pair_ij;
  k.pair_p = (const float4*)b->pair_p;
  k.conf_mol = b->conf_mol;
  k.conf_off = b->conf_off;
  k.n_mols = b->n_mols;
  k.n_confs = b->n_confs;
  k.max_atoms = b->max_atoms;
  return k;
}

struct MolDesc {
  int n_atoms, bond_off, nb, ang_off, na, sb_off, nsb, oop_off, noop, tor_off, ntor, pair_off, npair;
};

__device__ __forceinline__ MolDesc load_desc(const int32_t* md) {
  MolDesc d;
  d.n_atoms = md[0];
  d.bond_off = md[1];
  d.nb = md[2];
  d.ang_off = md[3];
  d.na = md[4];
  d.sb_off = md[5];
  d.nsb = md[6];
  d.oop_off = md[7];
  d.noop = md[8];
  d.tor_off = md[9];
  d.ntor = md[10];
  d.pair_off = md[11];
  d.npair = md[12];
  return d;
}

// ------------------------------------------------------------------------------------
// 64-bit fixed-point gradient accumulation in LDS (scale 2^28). Each FP32 contribution is
// rounded once to a multiple of 2^-28, so the accumulated sum is the exact sum of the
// rounded contributions, independent of arrival order.
//
// Conversion: for |v| < 2^23 the double FX_MAGIC + v, FX_MAGIC = 1.5 * 2^24, lies in
// [2^24, 2^25), whose spacing is 2^-28, so the addition rounds v to the nearest multiple of
// 2^-28 (ties to even) and the low bits of the sum hold that multiple: bits(FX_MAGIC + v) -
// bits(FX_MAGIC) = round(v * 2^28) as a two's-complement integer. Round-to-nearest-even is
// symmetric, so to_fx(-v) == -to_fx(v) exactly. One conversion, one FP64 add and one 64-bit
// subtract, against the sign/split/recombine sequence of a truncating conversion.
constexpr double FX_MAGIC = 25165824.0;   // 1.5 * 2^24
constexpr float FX_CAP = 8388000.0f;      // |v| is capped just below 2^23
constexpr double FX_UNIT = 3.7252902984619140625e-09;  // 2^-28

__device__ __forceinline__ unsigned long long to_fx(float v) {
  // fmaxf / fminf return the non-NaN operand (one v_med3_f32), so a NaN contribution maps to
  // -FX_CAP: finite, and only ever met where the energy is NaN too
  const float c = fminf(fmaxf(v, -FX_CAP), FX_CAP);
  const double y = (double)c + FX_MAGIC;
  return (unsigned long long)(__double_as_longlong(y) - __double_as_longlong(FX_MAGIC));
}

__device__ __forceinline__ float from_fx(unsigned long long a) {
  return (float)((double)(long long)a * FX_UNIT);
}

// The accumulators of a conformer with N atoms are three planes of N words (x, y, z), so
// lanes updating consecutive atoms touch consecutive 8-byte words (no bank conflicts); the
// gradient of coordinate i = 3a + c is plane c, word a. The -g side of a pair term is an LDS
// subtract (ds_sub_u64) of the same integer, so it needs no negation.
template <bool G>
struct LdsAcc {
  unsigned long long* g;
  int N;
  __device__ __forceinline__ void add(int a, float x, float y, float z) {
    if (G) {
      atomicAdd(&g[a], to_fx(x));
      atomicAdd(&g[N + a], to_fx(y));
      atomicAdd(&g[2 * N + a], to_fx(z));
    }
  }
  __device__ __forceinline__ static void sub(unsigned long long* p, unsigned long long v) {
    __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  // +g on atom a, -g on atom b: one conversion serves both
  __device__ __forceinline__ void add_pair(int a, int b, float x, float y, float z) {
    if (G) {
      const unsigned long long fx = to_fx(x), fy = to_fx(y), fz = to_fx(z);
      atomicAdd(&g[a], fx);
      atomicAdd(&g[N + a], fy);
      atomicAdd(&g[2 * N + a], fz);
      sub(&g[b], fx);
      sub(&g[N + b], fy);
      sub(&g[2 * N + b], fz);
    }
  }
};

// gradient of coordinate i (= 3a + c) from the accumulator planes of an N-atom conformer
__device__ __forceinline__ float gread(const unsigned long long* g, int i, int N) {
  const int a = i / 3;
  return from_fx(g[(i - 3 * a) * N + a]);
}

// ------------------------------------------------------------------------------------
// energy (+ gradient) of one conformer whose positions are in LDS, spread over NT lanes.
// Term tables are read from global memory (L2-resident: shared by the molecule's
// conformers); each lane prefetches its next term's indices and parameters before it
// evaluates the current one, so table latency overlaps the term arithmetic.
// Returns this lane's FP64 energy partial; per-term partials in et[7] when PT.
struct TermData {
  uint32_t a, b;  // packed atom indices
  float4 p;       // parameters
  float q;        // fifth parameter (stretch-bend theta0)
};

// Pointers to one molecule's term tables (offsets folded in). Built either over the
// global batch tables or over a copy staged in LDS.
struct TView {
  const uint32_t* bond_ij;
  const float2* bond_p;
  const uint2* ang_idx;
  const float2* ang_p;
  const uint2* sb_idx;
  const float4* sb_p;
  const float* sb_t0;
  const uint2* oop_idx;
  const float* oop_k;
  const uint2* tor_idx;
  const float4* tor_p;
  const uint32_t* pair_ij;
  const float4* pair_p;
};

__device__ __forceinline__ TView global_view(const KBatch& B, const MolDesc& m) {
  TView v;
  v.bond_ij = B.bond_ij + m.bond_off;
  v.bond_p = B.bond_p + m.bond_off;
  v.ang_idx = B.ang_idx + m.ang_off;
  v.ang_p = B.ang_p + m.ang_off;
  v.sb_idx = B.sb_idx + m.sb_off;
  v.sb_p = B.sb_p + m.sb_off;
  v.sb_t0 = B.sb_t0 + m.sb_off;
  v.oop_idx = B.oop_idx + m.oop_off;
  v.oop_k = B.oop_k + m.oop_off;
  v.tor_idx = B.tor_idx + m.tor_off;
  v.tor_p = B.tor_p + m.tor_off;
  v.pair_ij = B.pair_ij + m.pair_off;
  v.pair_p = B.pair_p + m.pair_off;
  return v;
}

// bytes of one molecule's tables in the LDS staging layout
__host__ __device__ inline size_t table_lds_bytes(int nb, int na, int nsb, int noop, int ntor, int npair) {
  return 16 * (size_t)(ntor + nsb + npair) + 8 * (size_t)(2 * na + nsb + noop + ntor + nb) +
         4 * (size_t)(nb + nsb + noop + npair);
}

template <class T>
__device__ __forceinline__ void stage_copy(T* dst, const T* src, int count, int tid, int nt) {
  const int words = count * (int)(sizeof(T) / 4);
  uint32_t* d = (uint32_t*)dst;
  const uint32_t* s = (const uint32_t*)src;
  for (int i = tid; i < words; i += nt) d[i] = s[i];
}

// Copy a molecule's tables into LDS at `base` (16-byte aligned) and return a view of them.
// Without `pairs` only the bonded tables are copied and the view reads pairs from global
// memory (their loop streams them with loads one step ahead).
__device__ __forceinline__ TView stage_tables(const KBatch& B, const MolDesc& m, unsigned char* base, int tid,
                                              int nt, bool pairs) {
  TView v;
  const int npair = pairs ? m.npair : 0;
  unsigned char* p = base;
  auto take = [&](size_t bytes) {
    unsigned char* q = p;
    p += bytes;
    return q;
  };
  float4* tor_p = (float4*)take(16 * (size_t)m.ntor);
  float4* sb_p = (float4*)take(16 * (size_t)m.nsb);
  float4* pair_p = (float4*)take(16 * (size_t)npair);
  uint2* ang_idx = (uint2*)take(8 * (size_t)m.na);
  float2* ang_p = (float2*)take(8 * (size_t)m.na);
  uint2* sb_idx = (uint2*)take(8 * (size_t)m.nsb);
  uint2* oop_idx = (uint2*)take(8 * (size_t)m.noop);
  uint2* tor_idx = (uint2*)take(8 * (size_t)m.ntor);
  float2* bond_p = (float2*)take(8 * (size_t)m.nb);
  uint32_t* bond_ij = (uint32_t*)take(4 * (size_t)m.nb);
  float* sb_t0 = (float*)take(4 * (size_t)m.nsb);
  float* oop_k = (float*)take(4 * (size_t)m.noop);
  uint32_t* pair_ij = (uint32_t*)take(4 * (size_t)npair);
  stage_copy(tor_p, B.tor_p + m.tor_off, m.ntor, tid, nt);
  stage_copy(sb_p, B.sb_p + m.sb_off, m.nsb, tid, nt);
  stage_copy(pair_p, B.pair_p + m.pair_off, npair, tid, nt);
  stage_copy(ang_idx, B.ang_idx + m.ang_off, m.na, tid, nt);
  stage_copy(ang_p, B.ang_p + m.ang_off, m.na, tid, nt);
  stage_copy(sb_idx, B.sb_idx + m.sb_off, m.nsb, tid, nt);
  stage_copy(oop_idx, B.oop_idx + m.oop_off, m.noop, tid, nt);
  stage_copy(tor_idx, B.tor_idx + m.tor_off, m.ntor, tid, nt);
  stage_copy(bond_p, B.bond_p + m.bond_off, m.nb, tid, nt);
  stage_copy(bond_ij, B.bond_ij + m.bond_off, m.nb, tid, nt);
  stage_copy(sb_t0, B.sb_t0 + m.sb_off, m.nsb, tid, nt);
  stage_copy(oop_k, B.oop_k + m.oop_off, m.noop, tid, nt);
  stage_copy(pair_ij, B.pair_ij + m.pair_off, npair, tid, nt);
  v.bond_ij = bond_ij;
  v.bond_p = bond_p;
  v.ang_idx = ang_idx;
  v.ang_p = ang_p;
  v.sb_idx = sb_idx;
  v.sb_p = sb_p;
  v.sb_t0 = sb_t0;
  v.oop_idx = oop_idx;
  v.oop_k = oop_k;
  v.tor_idx = tor_idx;
  v.tor_p = tor_p;
  v.pair_ij = pairs ? pair_ij : B.pair_ij + m.pair_off;
  v.pair_p = pairs ? (const float4*)pair_p : B.pair_p + m.pair_off;
  return v;
}

__device__ __forceinline__ TermData load_term(const TView& V, int t, int c1, int c2, int c3, int c4, int c5) {
  TermData d;
  d.q = 0.f;
  if (t < c1) {
    d.a = V.bond_ij[t];
    d.b = 0;
    const float2 q = V.bond_p[t];
    d.p = make_float4(q.x, q.y, 0.f, 0.f);
  } else if (t < c2) {
    const int u = t - c1;
    const uint2 ix = V.ang_idx[u];
    d.a = ix.x;
    d.b = ix.y;
    const float2 q = V.ang_p[u];
    d.p = make_float4(q.x, q.y, 0.f, 0.f);
  } else if (t < c3) {
    const int u = t - c2;
    const uint2 ix = V.sb_idx[u];
    d.a = ix.x;
    d.b = ix.y;
    d.p = V.sb_p[u];
    d.q = V.sb_t0[u];
  } else if (t < c4) {
    const int u = t - c3;
    const uint2 ix = V.oop_idx[u];
    d.a = ix.x;
    d.b = ix.y;
    d.p = make_float4(V.oop_k[u], 0.f, 0.f, 0.f);
  } else if (t < c5) {
    const int u = t - c4;
    const uint2 ix = V.tor_idx[u];
    d.a = ix.x;
    d.b = ix.y;
    d.p = V.tor_p[u];
  } else {
    const int u = t - c5;
    d.a = V.pair_ij[u];
    d.b = 0;
    d.p = V.pair_p[u];
  }
  return d;
}

// PREF: load each bonded term's tables one step ahead (tables in global memory); with the
// tables in LDS the loads are short and the prefetch only holds registers.
template <int FAM, bool G, bool PT, int NT, bool PREF = true>
__device__ __forceinline__ double eval_terms(const TView& TV, const MolDesc& m, const float* xs,
                                             unsigned long long* gacc, int diel, uint32_t mask, double* et) {
  LdsAcc<G> acc{gacc, m.n_atoms};
  double e = 0.0;
  // bonded terms (one generic loop, next term prefetched when PREF)
  const int c1 = m.nb, c2 = c1 + m.na, c3 = c2 + m.nsb, c4 = c3 + m.noop, c5 = c4 + m.ntor;
  int t = threadIdx.x;
  TermData cur;
  if (PREF && t < c5) cur = load_term(TV, t, c1, c2, c3, c4, c5);
  while (t < c5) {
    const int tn = t + NT;
    TermData nxt;
    if (!PREF)
      cur = load_term(TV, t, c1, c2, c3, c4, c5);
    else if (tn < c5)
      nxt = load_term(TV, tn, c1, c2, c3, c4, c5);
    const int i0 = (int)(cur.a & 0xffff), i1 = (int)(cur.a >> 16);
    const int i2 = (int)(cur.b & 0xffff), i3 = (int)(cur.b >> 16);
    if (t < c1) {
      if (mask & 1u) {
        const float v = t_bond<FAM, G>(xs, acc, i0, i1, cur.p.x, cur.p.y);
        e += v;
        if (PT) et[0] += v;
      }
    } else if (t < c2) {
      if (mask & 2u) {
        const float v = t_angle<FAM, G>(xs, acc, i0, i1, i2, cur.p.x, cur.p.y, i3 != 0);
        e += v;
        if (PT) et[1] += v;
      }
    } else if (t < c3) {
      if (mask & 4u) {
        const float v = t_strbend<FAM, G>(xs, acc, i0, i1, (int)cur.b, cur.p.x, cur.p.y, cur.p.z, cur.p.w, cur.q);
        e += v;
        if (PT) et[2] += v;
      }
    } else if (t < c4) {
      if (mask & 8u) {
        const float v = t_oop<FAM, G>(xs, acc, i0, i1, i2, i3, cur.p.x);
        e += v;
        if (PT) et[3] += v;
      }
    } else {
      if (mask & 16u) {
        const float v = t_torsion<FAM, G>(xs, acc, i0, i1, i2, i3, cur.p.x, cur.p.y, cur.p.z);
        e += v;
        if (PT) et[4] += v;
      }
    }
    if (PREF) cur = nxt;
    t = tn;
  }
  // non-bonded pairs: two pairs per lane per step (independent dependency chains), the
  // next two loaded while the current two are computed. The loads are unconditional (index
  // clamped to the last pair; such duplicates are never used), so the wait before the
  // current pairs' use leaves the next pairs' loads in flight.
  if ((mask & 96u) && m.npair > 0) {
    const float ke = (mask & 32u) ? 1.0f : 0.0f, kq = (mask & 64u) ? 1.0f : 0.0f;
    const int np = m.npair, last = np - 1;
    int u = threadIdx.x;
    uint32_t ija = TV.pair_ij[min(u, last)], ijb = TV.pair_ij[min(u + NT, last)];
    float4 pa = TV.pair_p[min(u, last)], pb = TV.pair_p[min(u + NT, last)];
    double ev_acc = 0.0, ee_acc = 0.0;
    while (u < np) {
      const int un = u + 2 * NT;
      const uint32_t nija = TV.pair_ij[min(un, last)], nijb = TV.pair_ij[min(un + NT, last)];
      const float4 npa = TV.pair_p[min(un, last)], npb = TV.pair_p[min(un + NT, last)];
      float eva, eea;
      t_pair<FAM, G>(xs, acc, (int)(ija & 0xffff), (int)(ija >> 16), pa.w, ke * pa.y, kq * pa.z, diel, eva, eea);
      // without per-term energies a pair's two parts are added in FP32 first (one rounding,
      // as small as the parts' own), then to the lane's FP64 sum
      if (PT) {
        ev_acc += eva;
        ee_acc += eea;
      } else {
        ev_acc += (double)(eva + eea);
      }
      if (u + NT < np) {
        float evb, eeb;
        t_pair<FAM, G>(xs, acc, (int)(ijb & 0xffff), (int)(ijb >> 16), pb.w, ke * pb.y, kq * pb.z, diel, evb, eeb);
        if (PT) {
          ev_acc += evb;
          ee_acc += eeb;
        } else {
          ev_acc += (double)(evb + eeb);
        }
      }
      ija = nija;
      ijb = nijb;
      pa = npa;
      pb = npb;
      u = un;
    }
    e += ev_acc + ee_acc;
    if (PT) {
      et[5] += ev_acc;
      et[6] += ee_acc;
    }
  }
  return e;
}

// ------------------------------------------------------------------------------------
// block reductions: wave shuffles, then one LDS exchange combined in fixed wave order
// Wave-level FP64 reductions with DPP: a butterfly inside each 16-lane row
// (quad xor 1, quad xor 2, half-row mirror, row mirror), then the four row results are
// read with readlane and combined in a fixed order, so every lane gets the same value.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}
constexpr int DPP_QUAD_XOR1 = 0xB1, DPP_QUAD_XOR2 = 0x4E, DPP_ROW_HALF_MIRROR = 0x141, DPP_ROW_MIRROR = 0x140;
__device__ __forceinline__ double wsum(double v) {
  v += dpp_d<DPP_QUAD_XOR1>(v);
  v += dpp_d<DPP_QUAD_XOR2>(v);
  v += dpp_d<DPP_ROW_HALF_MIRROR>(v);
  v += dpp_d<DPP_ROW_MIRROR>(v);
  return ((lane_d(v, 0) + lane_d(v, 16)) + lane_d(v, 32)) + lane_d(v, 48);
}
__device__ __forceinline__ double wmax(double v) {
  v = fmax(v, dpp_d<DPP_QUAD_XOR1>(v));
  v = fmax(v, dpp_d<DPP_QUAD_XOR2>(v));
  v = fmax(v, dpp_d<DPP_ROW_HALF_MIRROR>(v));
  v = fmax(v, dpp_d<DPP_ROW_MIRROR>(v));
  return fmax(fmax(lane_d(v, 0), lane_d(v, 16)), fmax(lane_d(v, 32), lane_d(v, 48)));
}

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// four consecutive elements (t, j..j+3) of a symmetric matrix in padded packed storage
// (column c at pcol(c), cb = pcol(t), pj = pcol(j)), with the matching gradient entries
__device__ __forceinline__ void hrow_load4(const float* H, const float2* VG, int t, int cb, int j, int pj,
                                           float (&h)[4], float (&w)[4]) {
  const int p1 = pj + ((j + 4) & ~3), p2 = p1 + ((j + 5) & ~3), p3 = p2 + ((j + 6) & ~3);
  h[0] = H[j <= t ? cb + j : pj + t];
  h[1] = H[j + 1 <= t ? cb + j + 1 : p1 + t];
  h[2] = H[j + 2 <= t ? cb + j + 2 : p2 + t];
  h[3] = H[j + 3 <= t ? cb + j + 3 : p3 + t];
  w[0] = VG[j].y;
  w[1] = VG[j + 1].y;
  w[2] = VG[j + 2].y;
  w[3] = VG[j + 3].y;
}

// one row of a column quad of the full Hessian gains a pending rank-2 update: column t's
// element gains ux[t] s + uy[t] h, with w = {s, h} of the row
__device__ __forceinline__ void hquad_upd(float4& h, float2 w, const float (&ux)[4], const float (&uy)[4]) {
  h.x = fmaf(uy[0], w.y, fmaf(ux[0], w.x, h.x));
  h.y = fmaf(uy[1], w.y, fmaf(ux[1], w.x, h.y));
  h.z = fmaf(uy[2], w.y, fmaf(ux[2], w.x, h.z));
  h.w = fmaf(uy[3], w.y, fmaf(ux[3], w.x, h.w));
}

// accumulate the row's gradient entry g into the quad's dot products
__device__ __forceinline__ void hquad_dot(const float4& h, float g, float (&a)[4]) {
  a[0] = fmaf(h.x, g, a[0]);
  a[1] = fmaf(h.y, g, a[1]);
  a[2] = fmaf(h.z, g, a[2]);
  a[3] = fmaf(h.w, g, a[3]);
}

// Rows [j, j1) of a column quad of the full Hessian (row r at base[r * nq]): accumulate
// H g into a, with H the stored rows plus the NU oldest pending rank-2 updates (update k:
// row vectors {s, h} in S[k], column coefficients cx[k], cy[k]). When all NPM slots are
// pending the updates are added to each element, oldest first, and the rows are written
// back. Otherwise the stored rows are only read, and the updates' part of H g is left as two
// dot products per update over these rows in lr (s.g, h.g), which the caller combines with
// the column coefficients after summing the quad's lanes. HQ rows are loaded together.
template <int HQ, int NU, int NPM>
__device__ __forceinline__ void hquad_rows(float4* base, int nq, int j, int j1, const float2* VG,
                                           const float2* const (&S)[NPM], const float (&cx)[NPM][4],
                                           const float (&cy)[NPM][4], float (&a)[4], float (&lr)[2][NPM]) {
  constexpr bool FLUSH = NU == NPM;
  for (; j + HQ - 1 < j1; j += HQ) {
    float4 h[HQ];
#pragma unroll
    for (int u = 0; u < HQ; u++) h[u] = base[(size_t)(j + u) * nq];
#pragma unroll
    for (int u = 0; u < HQ; u++) {
      const float g = VG[j + u].y;
#pragma unroll
      for (int k = 0; k < NU; k++) {
        const float2 w = S[k][j + u];
        if (FLUSH) {
          hquad_upd(h[u], w, cx[k], cy[k]);
        } else {
          lr[0][k] = fmaf(w.x, g, lr[0][k]);
          lr[1][k] = fmaf(w.y, g, lr[1][k]);
        }
      }
      hquad_dot(h[u], g, a);
    }
    if (FLUSH) {
#pragma unroll
      for (int u = 0; u < HQ; u++) base[(size_t)(j + u) * nq] = h[u];
    }
  }
  for (; j < j1; j++) {
    float4 h = base[(size_t)j * nq];
    const float g = VG[j].y;
#pragma unroll
    for (int k = 0; k < NU; k++) {
      const float2 w = S[k][j];
      if (FLUSH) {
        hquad_upd(h, w, cx[k], cy[k]);
      } else {
        lr[0][k] = fmaf(w.x, g, lr[0][k]);
        lr[1][k] = fmaf(w.y, g, lr[1][k]);
      }
    }
    hquad_dot(h, g, a);
    if (FLUSH) base[(size_t)j * nq] = h;
  }
}

// hquad_rows for the run-time count of pending updates (0..NPM)
template <int HQ, int NPM, int NU = NPM>
__device__ __forceinline__ void hquad_dispatch(int npend, float4* base, int nq, int j, int j1, const float2* VG,
                                               const float2* const (&S)[NPM], const float (&cx)[NPM][4],
                                               const float (&cy)[NPM][4], float (&a)[4], float (&lr)[2][NPM]) {
  if constexpr (NU == 0) {
    hquad_rows<HQ, 0, NPM>(base, nq, j, j1, VG, S, cx, cy, a, lr);
  } else {
    if (npend == NU)
      hquad_rows<HQ, NU, NPM>(base, nq, j, j1, VG, S, cx, cy, a, lr);
    else
      hquad_dispatch<HQ, NPM, NU - 1>(npend, base, nq, j, j1, VG, S, cx, cy, a, lr);
  }
}

// Row groups [ka, kb) (rows 4k..4k+3) of one column of the padded packed matrix gain
// alpha_i s + beta_i h (AB[i] = {alpha_i, beta_i}, u = {s, h} of the column); rows past the
// column's end are padding and never read as matrix elements.
__device__ __forceinline__ void hcol_update4(float* col, const float2* AB, float2 u, int ka, int kb) {
  float4* c4 = (float4*)col;
  const float4* ab4 = (const float4*)AB;
  for (int k = ka; k < kb; k++) {
    float4 h = c4[k];
    const float4 a01 = ab4[2 * k], a23 = ab4[2 * k + 1];
    h.x = fmaf(a01.y, u.y, fmaf(a01.x, u.x, h.x));
    h.y = fmaf(a01.w, u.y, fmaf(a01.z, u.x, h.y));
    h.z = fmaf(a23.y, u.y, fmaf(a23.x, u.x, h.z));
    h.w = fmaf(a23.w, u.y, fmaf(a23.z, u.x, h.w));
    c4[k] = h;
  }
}

template <int NT>
struct Reducer {
  static constexpr int NW = NT / 64;
  double* buf;  // RED_DOUBLES(NT) = 2 phases * NW * 8 doubles
  int phase;
  // K values; bit k of MAXMASK selects max instead of sum
  template <int K, int MAXMASK>
  __device__ __forceinline__ void run(double (&v)[K]) {
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = ((MAXMASK >> k) & 1) ? wmax(v[k]) : wsum(v[k]);
    static_assert(K <= 8, "at most 8 values per reduction");
    double* b = buf + phase * (NW * 8);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < K; k++) b[w * 8 + k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
      double a = b[k];
      for (int ww = 1; ww < NW; ww++) a = ((MAXMASK >> k) & 1) ? fmax(a, b[ww * 8 + k]) : a + b[ww * 8 + k];
      v[k] = a;
    }
    phase ^= 1;
  }
  __device__ __forceinline__ double sum(double a) {
    double v[1] = {a};
    run<1, 0>(v);
    return v[0];
  }
  __device__ __forceinline__ double max(double a) {
    double v[1] = {a};
    run<1, 1>(v);
    return v[0];
  }
};

// ------------------------------------------------------------------------------------
// Kernel: single-point energy / gradient / per-term energies, one workgroup per conformer.
// Replaces mmff_energy_grad_metal(_mega) (mmff_metal_kernel.py:713-775); the TG family is
// what the nk kernel evaluates (mmff_bfgs_header.metal).
template <int FAM, int NT>
__global__ void __launch_bounds__(NT) k_energy_grad(KBatch B, const float* __restrict__ x, double* e_out,
                                                    double* e_terms, float* g_out, uint32_t mask, int diel) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int c = blockIdx.x;
  if (c >= B.n_confs) return;
  constexpr int NW = NT / 64;
  double* red = (double*)smem;
  unsigned long long* gacc = (unsigned long long*)(red + 16 * NW);
  const int mx = 3 * B.max_atoms;
  float* xs = (float*)(gacc + mx);
  const MolDesc m = load_desc(B.mol_desc + 16 * B.conf_mol[c]);
  const int n = 3 * m.n_atoms;
  const int off = B.conf_off[c];
  for (int i = threadIdx.x; i < n; i += NT) xs[i] = x[off + i];
  for (int i = threadIdx.x; i < n; i += NT) gacc[i] = 0ull;
  __syncthreads();
  Reducer<NT> R{red, 0};
  double et[7] = {0, 0, 0, 0, 0, 0, 0};
  double e;
  if (g_out)
    e = eval_terms<FAM, true, true, NT>(global_view(B, m), m, xs, gacc, diel, mask, et);
  else
    e = eval_terms<FAM, false, true, NT>(global_view(B, m), m, xs, gacc, diel, mask, et);
  e = R.sum(e);  // includes the barrier that completes the gradient atomics
  if (e_terms) {
    double v4[4] = {et[0], et[1], et[2], et[3]};
    R.template run<4, 0>(v4);
    double v3[4] = {et[4], et[5], et[6], 0.0};
    R.template run<4, 0>(v3);
    if (threadIdx.x == 0) {
      for (int k = 0; k < 4; k++) e_terms[7 * c + k] = v4[k];
      for (int k = 0; k < 3; k++) e_terms[7 * c + 4 + k] = v3[k];
    }
  }
  if (threadIdx.x == 0) e_out[c] = e;
  if (g_out)
    for (int i = threadIdx.x; i < n; i += NT) g_out[off + i] = gread(gacc, i, m.n_atoms);
}

// ------------------------------------------------------------------------------------
// nk minimizer (in-kernel BFGS / L-BFGS), TG formula family.
// Control flow: mlxmolkit/mmff_bfgs_source_tg.metal:1-372 (BFGS) and
// mlxmolkit/mmff_lbfgs_source_tg.metal:1-298 (L-BFGS); constants mmff_bfgs_header.metal:3-10.
constexpr double TOLX = 1.2e-6;
constexpr double FUNCTOL = 1e-4;
constexpr double MOVETOL = 1e-6;
constexpr double EPS_GUARD = 3e-7;
constexpr double MAX_STEP_FACTOR = 100.0;
constexpr int MAX_LS_ITERS = 1000;
constexpr float GRAD_SCALE_INIT = 0.1f;
constexpr double GRAD_CAP = 10.0;
constexpr int LBFGS_MAX_M = 16;
// pending rank-2 updates the global-Hessian BFGS pass applies on the fly before writing H back
constexpr int HPEND_MAX = 4;
// the same for the 512-thread kernel, which has the registers for more (one workgroup per CU,
// 233 VGPRs without spills at six): H is then written every sixth iteration. Ragged workload,
// n > 360 class: 1692.5 ms against 1726.2 ms with four and 1716.2 ms with eight
constexpr int HPEND_LARGE = 6;
__host__ __device__ constexpr int hpend_for(int nt) { return nt >= 512 ? HPEND_LARGE : HPEND_MAX; }

struct NkParams {
  const float* x_in;
  float* x_out;
  double* e_out;
  int32_t* status;
  int32_t* counters;
  const int32_t* order;
  int32_t* queue;     // work counters, zeroed before launch: one, or one per XCD
  float* hglob;       // per-workgroup global slots (H or L-BFGS history)
  size_t slot_floats;
  const float* precond;
  int max_iters;
  float grad_tol;
  int diel;
  int m;
  int max_n;          // 3 * max_atoms rounded up to a multiple of 4
  int ls_grad;
  int tbytes;         // LDS bytes reserved for staged term tables (TL kernels)
  int xcd_on;         // order is split into 8 per-XCD queues at xcd_off
  int32_t xcd_off[9];
};

// Packed upper-triangle storage of a symmetric n x n matrix with every column padded to a
// multiple of 4 floats (16 bytes): column c holds rows 0..c from offset pcol(c) =
// sum_{k<c} round4(k + 1), so every column starts 16-byte aligned and can be walked in float4
// steps. packed_floats(n) = pcol(n) floats hold the matrix.
__host__ __device__ inline int pcol(int c) {
  const int m = c >> 2, r = c & 3;
  return 8 * m * (m + 1) + 4 * r * (m + 1);
}
__host__ __device__ inline size_t packed_floats(int n) { return (size_t)pcol(n); }

// Symmetric blocked storage in HBM: the matrix is cut into 4x4 blocks over the nq = n/4 row
// quads, and only one block per unordered quad pair {L, M} is kept. Block (s, L) holds rows
// 4L..4L+3 and columns 4M..4M+3 with M = (L + s) mod nq, for the cyclic distances
// s = 0 .. nq/2 (for even nq the pairs at distance nq/2 are kept once, by L < nq/2); its row r
// is float4 number (4 s + r) nq + L, so the lanes that walk consecutive L of one distance read
// consecutive 16-byte words. sym_slot_floats(max_n) floats hold the largest matrix.
__host__ __device__ inline int sym_dists(int nq) { return (nq >> 1) + 1; }
__host__ __device__ inline size_t sym_slot_floats(int max_n) {
  const int nq = max_n >> 2;
  return (size_t)16 * sym_dists(nq) * nq;
}
// LDS floats of one wave's private H v accumulators: up to 64/nq copies of the 4 nq entries
// when nq <= 64 (at most 256 floats), one copy otherwise
__host__ __device__ inline int sym_acc_floats(int max_n) { return max_n > 256 ? max_n : 256; }

// LDS footprint (bytes) of the nk kernel
// hm: Hessian placement (0 packed LDS, 1 full HBM, 2 packed HBM, 3 symmetric blocked HBM);
// only the full and symmetric HBM BFGS kernels keep the extra pending-update slots, and the
// symmetric one adds its per-wave accumulators
__host__ __device__ inline size_t nk_lds_bytes(int algo, int hm, int nt, int max_n, int m, size_t tbytes = 0) {
  const bool hlds = hm == 0;
  size_t nw = (size_t)nt / 64;
  size_t b = 16 * nw * sizeof(double) + 64 + 2 * (size_t)LBFGS_MAX_M * sizeof(double);
  b += (size_t)max_n * 8;  // gradient accumulators
  b += (size_t)max_n * 4 * 4;   // xs, d, xo, hd
  b += (size_t)max_n * 16;      // V = {dg, g, pending xi, pending hd}
  b += (size_t)max_n * 8;       // per-lane coefficients of the pending update
  if (algo == MMFFHIP_BFGS && (hm == 1 || hm == 3)) b += (size_t)(hpend_for(nt) - 1) * max_n * 16;  // more pending updates
  if (algo == MMFFHIP_BFGS && hm == 3) b += nw * (size_t)sym_acc_floats(max_n) * 4;
  if (algo == MMFFHIP_LBFGS) {
    if (hlds) b += 2 * (size_t)m * max_n * 4;
  } else if (hlds) {
    b += ((packed_floats(max_n) * 4) + 15) & ~(size_t)15;
  }
  b += (tbytes + 15) & ~(size_t)15;
  return b;
}

// initial inverse Hessian element: the identity, or the symmetric part of the learned
// per-atom 3x3 block times psc (blocks approximate (0.1 * Hessian block)^-1; psc = 0.1 /
// gradient scale rescales them to the scaled gradient the minimizer works with). BFGS keeps
// H symmetric, and every placement reads one orientation of each element, so the start is
// made symmetric rather than taken from one triangle of the block.
__device__ __forceinline__ float h_init(const float* pc, float psc, int i, int j) {
  if (pc) {
    if ((i / 3) != (j / 3)) return 0.0f;
    const float* b = pc + (j / 3) * 9;
    return psc * (0.5f * (b[(i % 3) * 3 + (j % 3)] + b[(j % 3) * 3 + (i % 3)]));
  }
  return i == j ? 1.0f : 0.0f;
}

// HM: where the dense inverse Hessian lives -- 0 packed in LDS, 1 full in HBM, 2 packed in HBM
// TL: term tables read from global memory (0), staged in LDS (1), or bonded tables staged and
// pairs read from global memory (2). Global-Hessian kernels of up to 256 threads are compiled
// for three waves per SIMD (at most 168 VGPRs), so three workgroups share a CU: their LDS
// use is small, and the extra resident conformers hide memory latency better than the
// registers the compiler would otherwise use (medium workload: 172.6 against 178.5 ms).
template <int ALGO, int HM, int TL, int NT>
__global__ void __launch_bounds__(NT, ((HM == 1 || HM == 3) && NT <= 256) ? 3 : 1) k_minimize_nk(KBatch B, NkParams P) {
  constexpr bool HLDS = HM == 0;    // LDS-resident (also selects the LDS L-BFGS history)
  constexpr bool PACKED = HM == 0 || HM == 2;  // packed upper triangle, select read + column-pair write-back
  constexpr bool SYM = HM == 3 && ALGO == MMFFHIP_BFGS;  // symmetric 4x4 blocks in HBM
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x;
  const int mn = P.max_n;
  double* red = (double*)smem;
  int* s_int = (int*)(red + 16 * NW);
  double* lb_sc = (double*)(s_int + 16);  // rho[16], alpha[16]
  unsigned long long* gacc = (unsigned long long*)(lb_sc + 2 * LBFGS_MAX_M);
  float2* VG = (float2*)(gacc + mn);  // {dg, g}
  float2* SH = VG + mn;               // {s, h} of the last rank-2 update (s = x - x_old, h = H dg)
  float* xs = (float*)(SH + mn);
  float* d = xs + mn;   // direction (also y2 = H g scratch)
  float* xo = d + mn;   // previous x, then xi = x - x_old
  float* hd = xo + mn;  // H dg (BFGS) / two-loop q (L-BFGS)
  float2* AB = (float2*)(hd + mn);  // pending update U(t, j) = AB[t].x xi_j + AB[t].y hd_j
  // global-Hessian BFGS keeps up to NPM pending updates (slot 0 in SH/AB, slot k >= 1 in XS)
  // and writes H back once every NPM iterations
  constexpr bool LAZY = (HM == 1 || HM == 3) && ALGO == MMFFHIP_BFGS;
  constexpr int NPM = LAZY ? hpend_for(NT) : 1;
  float2* XS = AB + mn;
  auto sh_slot = [&](int k) { return k == 0 ? SH : XS + (size_t)(k - 1) * 2 * mn; };
  auto ab_slot = [&](int k) { return k == 0 ? AB : XS + (size_t)(k - 1) * 2 * mn + mn; };
  float* big = (float*)(XS + (size_t)(NPM - 1) * 2 * mn);
  float* H = HLDS ? big : P.hglob + (size_t)blockIdx.x * P.slot_floats;
  size_t big_floats = 0;
  if (HLDS) big_floats = ALGO == MMFFHIP_BFGS ? packed_floats(mn) : 2 * (size_t)P.m * mn;
  // symmetric storage: per-wave H v accumulators (kept zero between passes)
  const int sacc_w = sym_acc_floats(mn);
  if (SYM) {
    big_floats = (size_t)NW * sacc_w;
    for (int i = tid; i < NW * sacc_w; i += NT) big[i] = 0.f;
  }
  unsigned char* tstage = (unsigned char*)(big + big_floats);
  const int M = P.m;
  float* S = H;
  float* Y = H + (size_t)M * mn;
  double* rho = lb_sc;
  double* alf = lb_sc + LBFGS_MAX_M;

  Reducer<NT> R{red, 0};

  for (;;) {
    if (tid == 0) {
      int kk;
      if (P.xcd_on) {
        // the queue of this workgroup's XCD first (workgroups are dispatched to the 8 XCDs
        // round-robin), then the others' until every queue is drained
        const int x0 = blockIdx.x & 7;
        kk = B.n_confs;
        for (int t = 0; t < 8; t++) {
          const int x = (x0 + t) & 7;
          const int cnt = P.xcd_off[x + 1] - P.xcd_off[x];
          if (cnt <= 0) continue;
          const int v = atomicAdd(P.queue + x, 1);
          if (v < cnt) {
            kk = P.xcd_off[x] + v;
            break;
          }
        }
      } else {
        kk = atomicAdd(P.queue, 1);
      }
      s_int[0] = kk;
    }
    __syncthreads();
    const int k = s_int[0];
    __syncthreads();
    if (k >= B.n_confs) break;
    const int c = P.order ? P.order[k] : k;
    const MolDesc m = load_desc(B.mol_desc + 16 * B.conf_mol[c]);
    const int n = 3 * m.n_atoms;
    const int off = B.conf_off[c];
    // packed-Hessian passes: 2^lsh lanes per row in the read pass, 2^lsh2 lanes per column
    // pair in the write-back (at most 8 lanes each)
    const int np = (n + 1) >> 1;
    int lsh = 0, lsh2 = 0;
    while (lsh < 3 && (n << (lsh + 1)) <= NT) lsh++;
    while (lsh2 < 3 && (np << (lsh2 + 1)) <= NT) lsh2++;
    const float* pc = P.precond ? P.precond + (size_t)(off / 3) * 9 : nullptr;
    TView TV;
    if (TL) {
      if (table_lds_bytes(m.nb, m.na, m.nsb, m.noop, m.ntor, TL == 1 ? m.npair : 0) > (size_t)P.tbytes) {
        // tables larger than the reserved LDS: report and skip (the planner never does this)
        for (int i = tid; i < n; i += NT) P.x_out[off + i] = P.x_in[off + i];
        if (tid == 0) {
          P.e_out[c] = __longlong_as_double(0x7ff8000000000000ll);
          P.status[c] = 2;
          if (P.counters) P.counters[3 * c] = P.counters[3 * c + 1] = P.counters[3 * c + 2] = 0;
        }
        continue;
      }
      TV = stage_tables(B, m, tstage, tid, NT, TL == 1);
    } else {
      TV = global_view(B, m);
    }

    for (int i = tid; i < n; i += NT) xs[i] = P.x_in[off + i];
    for (int i = tid; i < n; i += NT) gacc[i] = 0ull;
    if (SYM) {
      // the padding rows of the last quad: zero gradient and update vectors, so that the
      // padding elements of H stay zero and add nothing to H v
      for (int i = n + tid; i < ((n + 3) & ~3); i += NT) {
        VG[i] = make_float2(0.f, 0.f);
#pragma unroll
        for (int k = 0; k < NPM; k++) sh_slot(k)[i] = ab_slot(k)[i] = make_float2(0.f, 0.f);
      }
    }
    float psc = 1.0f;
    auto init_h = [&]() {
      if (ALGO != MMFFHIP_BFGS) return;
      if (PACKED) {
        for (int col = tid; col < n; col += NT) {
          float* cp = H + pcol(col);
          for (int i = 0; i < ((col + 4) & ~3); i++) cp[i] = i <= col ? h_init(pc, psc, i, col) : 0.0f;
        }
      } else if (SYM) {
        const int nq = (n + 3) >> 2;
        float4* H4 = (float4*)H;
        for (int f = tid; f < 4 * sym_dists(nq) * nq; f += NT) {
          const int sr = f / nq, L = f - sr * nq, r = sr & 3, s = sr >> 2;
          const int M = L + s < nq ? L + s : L + s - nq;
          const int jr = 4 * L + r;
          float v[4];
#pragma unroll
          for (int c = 0; c < 4; c++) {
            const int t = 4 * M + c;
            v[c] = (jr < n && t < n) ? h_init(pc, psc, jr, t) : 0.0f;
          }
          H4[f] = make_float4(v[0], v[1], v[2], v[3]);
        }
      } else {
        const int ldh = (n + 3) & ~3;  // rows padded to 16 bytes; padding columns stay zero
        for (int col = tid; col < ldh; col += NT)
          for (int j = 0; j < n; j++) H[(size_t)j * ldh + col] = col < n ? h_init(pc, psc, j, col) : 0.0f;
      }
    };
    __syncthreads();

    int n_grads = 1, n_evals = 0, iters = 0;
    // initial energy + gradient (mmff_bfgs_source_tg.metal:115-135)
    double E = R.sum(eval_terms<FAM_TG, true, false, NT, TL == 0>(TV, m, xs, gacc, P.diel, 0x7fu, nullptr));
    float gscale = GRAD_SCALE_INIT;
    {
      double mxg = 0.0;
      for (int i = tid; i < n; i += NT) mxg = fmax(mxg, (double)fabsf(gread(gacc, i, m.n_atoms) * gscale));
      mxg = R.max(mxg);
      while (mxg > GRAD_CAP) {
        gscale *= 0.5f;
        mxg *= 0.5;
      }
      for (int i = tid; i < n; i += NT) {
        VG[i] = make_float2(0.f, gread(gacc, i, m.n_atoms) * gscale);
        SH[i] = make_float2(0.f, 0.f);
      }
    }
    psc = GRAD_SCALE_INIT / gscale;
    init_h();  // H0 needs the gradient scale when it comes from the learned blocks
    __syncthreads();
    // initial direction: -g, or -H0 g with the learned preconditioner
    for (int i = tid; i < n; i += NT) {
      if (pc && ALGO == MMFFHIP_BFGS) {
        const int a = i / 3;
        float sacc = 0.f;
        for (int q = 0; q < 3; q++) sacc += h_init(pc, psc, i, 3 * a + q) * VG[3 * a + q].y;
        d[i] = -sacc;
      } else {
        d[i] = -VG[i].y;
      }
    }
    double ss = 0.0;
    for (int i = tid; i < n; i += NT) ss += (double)xs[i] * xs[i];
    ss = R.sum(ss);
    const double max_step = MAX_STEP_FACTOR * fmax(sqrt(ss), (double)n);  // :141-145
    int status = 1;
    int hist_count = 0, hist_idx = 0;
    bool fallback_used = (pc == nullptr);
    int npend = 0;  // rank-2 updates pending (applied lazily in later H passes)

    for (int it = 0; it < P.max_iters; it++) {
      iters = it + 1;
      const double E_old = E;
      double v3[3] = {0.0, 0.0, 0.0};
      for (int i = tid; i < n; i += NT) {
        xo[i] = xs[i];
        v3[0] += (double)d[i] * d[i];
        v3[1] += (double)d[i] * VG[i].y;
        v3[2] = fmax(v3[2], (double)(fabsf(d[i]) / fmaxf(fabsf(xs[i]), 1.0f)));
      }
      R.template run<3, 4>(v3);
      double slope = v3[1], test = v3[2];
      const double dn = sqrt(v3[0]);
      // a step longer than max_step is scaled down (:157-166); the scale goes into the
      // step length instead of into d, so that d stays -H g, which the next H pass uses
      double dsc = 1.0;
      if (dn > max_step) {
        dsc = (double)(float)(max_step / dn);
        slope *= dsc;
        test *= dsc;
      }
      if (slope >= 0.0) {  // ascent bail-out :170-174
        status = 0;
        break;
      }
      const double lam_min = MOVETOL / fmax(test, 1e-30);
      // backtracking line search :192-256
      double lam = 1.0, prev_lam = 1.0, prev_e = E_old;
      bool accepted = false, have_grad = false;
      for (int ls = 0; ls < MAX_LS_ITERS; ls++) {
        if (lam < lam_min) break;
        const float lf = (float)(lam * dsc);
        for (int i = tid; i < n; i += NT) {
          xs[i] = xo[i] + lf * d[i];
          if (P.ls_grad) gacc[i] = 0ull;
        }
        __syncthreads();
        double te;
        if (P.ls_grad)
          te = R.sum(eval_terms<FAM_TG, true, false, NT, TL == 0>(TV, m, xs, gacc, P.diel, 0x7fu, nullptr));
        else
          te = R.sum(eval_terms<FAM_TG, false, false, NT, TL == 0>(TV, m, xs, gacc, P.diel, 0x7fu, nullptr));
        n_evals++;
        if (te - E_old <= FUNCTOL * lam * slope) {
          E = te;
          accepted = true;
          have_grad = P.ls_grad != 0;
          break;
        }
        double tmp;
        if (ls == 0) {
          tmp = -slope / (2.0 * (te - E_old - slope));
        } else {
          const double rhs1 = te - E_old - lam * slope;
          const double rhs2 = prev_e - E_old - prev_lam * slope;
          const double l2 = lam * lam, pl2 = prev_lam * prev_lam, dv = lam - prev_lam;
          if (fabs(dv) < 1e-30) {
            tmp = 0.5 * lam;
          } else {
            const double a = (rhs1 / l2 - rhs2 / pl2) / dv;
            const double b = (-prev_lam * rhs1 / l2 + lam * rhs2 / pl2) / dv;
            if (fabs(a) < 1e-30) {
              tmp = fabs(b) > 1e-30 ? -slope / (2.0 * b) : 0.5 * lam;
            } else {
              const double disc = b * b - 3.0 * a * slope;
              if (disc < 0.0)
                tmp = 0.5 * lam;
              else if (b <= 0.0)
                tmp = (-b + sqrt(disc)) / (3.0 * a);
              else
                tmp = -slope / (b + sqrt(disc));
            }
          }
        }
        tmp = fmin(tmp, 0.5 * lam);
        tmp = fmax(tmp, 0.1 * lam);
        prev_lam = lam;
        prev_e = te;
        lam = tmp;
      }
      if (!accepted) {
        for (int i = tid; i < n; i += NT) xs[i] = xo[i];
        if (!fallback_used) {
          // learned preconditioner fallback: restart as identity-scaled BFGS
          fallback_used = true;
          pc = nullptr;
          npend = 0;
          init_h();
          for (int i = tid; i < n; i += NT) d[i] = -VG[i].y;
          __syncthreads();
          continue;
        }
      }
      // xi = x - x_old and TOLX :258-279
      double tx = 0.0;
      for (int i = tid; i < n; i += NT) {
        const float xi = xs[i] - xo[i];
        xo[i] = xi;
        tx = fmax(tx, (double)(fabsf(xi) / fmaxf(fabsf(xs[i]), 1.0f)));
      }
      tx = R.max(tx);
      if (tx < TOLX) {
        status = 0;
        break;
      }
      // new gradient :285-298
      if (!have_grad) {
        for (int i = tid; i < n; i += NT) gacc[i] = 0ull;
        __syncthreads();
        eval_terms<FAM_TG, true, false, NT, TL == 0>(TV, m, xs, gacc, P.diel, 0x7fu, nullptr);
        __syncthreads();
      }
      n_grads++;
      double v2[2] = {0.0, 0.0};
      for (int i = tid; i < n; i += NT) {
        const float sg = fabsf(gread(gacc, i, m.n_atoms) * gscale);
        v2[0] = fmax(v2[0], (double)sg);
        v2[1] = fmax(v2[1], (double)sg * fmax((double)fabsf(xs[i]), 1.0));
      }
      R.template run<2, 3>(v2);
      while (v2[0] > GRAD_CAP) {  // scale_grad_serial, mmff_bfgs_header.metal:435-454
        gscale *= 0.5f;
        v2[0] *= 0.5;
        v2[1] *= 0.5;
      }
      for (int i = tid; i < n; i += NT) {
        const float gn = gread(gacc, i, m.n_atoms) * gscale;
        VG[i] = make_float2(gn - VG[i].y, gn);
      }
      if (v2[1] / fmax(E * (double)gscale, 1.0) < (double)P.grad_tol) {  // :300-316
        status = 0;
        break;
      }
      __syncthreads();  // V visible to every lane

      if (ALGO == MMFFHIP_BFGS) {
        // One pass over H computes hd = H_k dg and y2 = H_k g, where H_k = H_stored +
        // pending rank-2 update; the pending update is written back in the same pass.
        // The update U = fi s s^T - fei h h^T + fae a a^T (a = fi s - fei h) is kept in the
        // factored form U(t, j) = alpha_t s_j + beta_t h_j (2 FMAs per element).
        // Reference: H dg (:318-324), dots (:326-345), rank-2 update (:347-358), d = -H g (:360-366).
        if (PACKED) {
          // Packed upper triangle (column j holds rows 0..j), kept fully updated. 2^lsh
          // consecutive lanes share row t and split its n columns into equal segments;
          // element (t, j) is read from column t when j <= t and from column j otherwise, so
          // every lane reads the same number of elements and H is read-only here. The
          // partial sums of a row are combined with DPP in a fixed order. Loads of the next
          // four columns are issued before the current four are consumed.
            const int lpr = 1 << lsh, seg = (n + lpr - 1) >> lsh;
            for (int r = tid; r < (n << lsh); r += NT) {
              const int t = r >> lsh, q = r & (lpr - 1);
              const int j0 = q * seg, j1 = min(n, j0 + seg);
              const int cb = pcol(t);
              int pj = pcol(j0);
              float a2[4] = {0.f, 0.f, 0.f, 0.f};
              int j = j0;
              if (j + 3 < j1) {
                float h[4], w[4];
                hrow_load4(H, VG, t, cb, j, pj, h, w);
                while (j + 7 < j1) {
                  float hn[4], wn[4];
                  const int pn = pcol(j + 4);
                  hrow_load4(H, VG, t, cb, j + 4, pn, hn, wn);
  #pragma unroll
                  for (int u = 0; u < 4; u++) {
                    a2[u] += h[u] * w[u];
                    h[u] = hn[u];
                    w[u] = wn[u];
                  }
                  j += 4;
                  pj = pn;
                }
  #pragma unroll
                for (int u = 0; u < 4; u++) a2[u] += h[u] * w[u];
                j += 4;
                pj = pcol(j);
              }
              for (; j < j1; j++) {
                const float hv = H[j <= t ? cb + j : pj + t];
                a2[0] += hv * VG[j].y;
                pj += (j + 4) & ~3;
              }
              float s2 = (a2[0] + a2[1]) + (a2[2] + a2[3]);
              if (lsh >= 1) s2 += dpp_f<DPP_QUAD_XOR1>(s2);
              if (lsh >= 2) s2 += dpp_f<DPP_QUAD_XOR2>(s2);
              if (lsh >= 3) s2 += dpp_f<DPP_ROW_HALF_MIRROR>(s2);
              if (q == 0) {
                // H dg = H g_new - H g_old = H g_new + d_old
                hd[t] = s2 + d[t];
                d[t] = s2;
              }
            }
            __syncthreads();
        } else if (SYM) {
          // Symmetric blocked storage (sym_slot_floats): each kept 4x4 block B of the quad pair
          // (L, M) is read once and serves both halves of H v: rows 4L.. gain B v_M and, off
          // the diagonal, rows 4M.. gain B^T v_L. Wave w takes the distances s = w, w + NW, ...
          // and its lanes walk (distance, L) items: for nq <= 64, 64/nq groups of nq lanes take
          // one distance each; for nq >= 64 + NW the items run on consecutively, so that one
          // instruction spans at most two distances NW apart, whose rows 4M.. are disjoint;
          // in between, 64 lanes and then nq - 64 take each distance. The products go into the
          // wave's private LDS accumulators (one copy per lane group, a float4 per quad) by a
          // read, add and write of the quad: no two lanes of one instruction touch the same
          // quad and a wave's LDS operations complete in order, so each sum is formed in a
          // fixed order and results are deterministic. (LDS float atomics give the same sums
          // but ran the medium workload at 302.6 ms against 119.0 ms: waves waited on LDS for
          // 42% of their cycles.) When all NPM update slots are pending they are added to each
          // element as it is read, oldest first, and the block is written back; otherwise
          // the pending updates enter H v through their factored form, from dot products
          // taken after the pass.
          const int nq = (n + 3) >> 2, D = sym_dists(nq), n4 = nq << 2;
          const int lane = tid & 63, w = tid >> 6;
          const int Jw = w < D ? (D - 1 - w) / NW + 1 : 0;  // distances of this wave
          const int G = nq <= 64 ? 64 / nq : 1;             // lane groups (accumulator copies)
          float* acc = big + (size_t)w * sacc_w;
          int j, L, cp = 0;
          if (nq <= 64) {
            cp = lane / nq;
            L = lane - cp * nq;
            j = cp < G ? cp : Jw;  // lanes past the last group stay idle
          } else {
            j = 0;
            L = lane;
          }
          const int steps = nq <= 64 ? (Jw + G - 1) / G : nq < 64 + NW ? 2 * Jw : (Jw * nq + 63) >> 6;
          const bool flush = npend == NPM;
          const float4* VG4 = (const float4*)VG;  // VG4[i] = {VG[2i], VG[2i + 1]}
          float4* H4 = (float4*)H;
          for (int kk = 0; kk < steps; kk++) {
            if (j < Jw && L < nq) {
              const int s = w + j * NW;
              const int M = L + s < nq ? L + s : L + s - nq;
              if (!((nq & 1) == 0 && s == D - 1 && s != 0 && L >= (nq >> 1))) {
                float4* bp = H4 + (size_t)(4 * s) * nq + L;
                float4 b0 = bp[0], b1 = bp[nq], b2 = bp[2 * nq], b3 = bp[3 * nq];
                if (flush) {
#pragma unroll
                  for (int k = 0; k < NPM; k++) {
                    const float4* sk = (const float4*)(sh_slot(k) + 4 * L);  // {s, h} of rows 4L..
                    const float4* ak = (const float4*)(ab_slot(k) + 4 * M);  // {alpha, beta} of columns 4M..
                    const float4 s01 = sk[0], s23 = sk[1], a01 = ak[0], a23 = ak[1];
                    const float ux[4] = {a01.x, a01.z, a23.x, a23.z}, uy[4] = {a01.y, a01.w, a23.y, a23.w};
                    hquad_upd(b0, make_float2(s01.x, s01.y), ux, uy);
                    hquad_upd(b1, make_float2(s01.z, s01.w), ux, uy);
                    hquad_upd(b2, make_float2(s23.x, s23.y), ux, uy);
                    hquad_upd(b3, make_float2(s23.z, s23.w), ux, uy);
                  }
                  bp[0] = b0;
                  bp[nq] = b1;
                  bp[2 * nq] = b2;
                  bp[3 * nq] = b3;
                }
                // rows 4L.. gain B v_M; then, off the diagonal, rows 4M.. gain B^T v_L. Another
                // lane of this wave may touch quad M in the first update or quad L in the second,
                // so the compiler barriers keep each read behind the previous write (the
                // hardware keeps a wave's LDS operations in order). (Keeping the first update in
                // registers while a lane's L stays fixed measured no faster: 119.9 against
                // 118.9 ms on the medium workload; loading the next block's rows while this one
                // is used spilled 76-244 bytes per lane and measured 125.1 against 119.5 ms.)
                const float4 vm01 = VG4[2 * M], vm23 = VG4[2 * M + 1];
                float4* a4 = (float4*)(acc + cp * n4);
                {
                  float4 v = a4[L];
                  v.x += fmaf(b0.w, vm23.w, fmaf(b0.z, vm23.y, fmaf(b0.y, vm01.w, b0.x * vm01.y)));
                  v.y += fmaf(b1.w, vm23.w, fmaf(b1.z, vm23.y, fmaf(b1.y, vm01.w, b1.x * vm01.y)));
                  v.z += fmaf(b2.w, vm23.w, fmaf(b2.z, vm23.y, fmaf(b2.y, vm01.w, b2.x * vm01.y)));
                  v.w += fmaf(b3.w, vm23.w, fmaf(b3.z, vm23.y, fmaf(b3.y, vm01.w, b3.x * vm01.y)));
                  a4[L] = v;
                }
                asm volatile("" ::: "memory");
                if (s != 0) {
                  const float4 vl01 = VG4[2 * L], vl23 = VG4[2 * L + 1];
                  float4 v = a4[M];
                  v.x += fmaf(b3.x, vl23.w, fmaf(b2.x, vl23.y, fmaf(b1.x, vl01.w, b0.x * vl01.y)));
                  v.y += fmaf(b3.y, vl23.w, fmaf(b2.y, vl23.y, fmaf(b1.y, vl01.w, b0.y * vl01.y)));
                  v.z += fmaf(b3.z, vl23.w, fmaf(b2.z, vl23.y, fmaf(b1.z, vl01.w, b0.z * vl01.y)));
                  v.w += fmaf(b3.w, vl23.w, fmaf(b2.w, vl23.y, fmaf(b1.w, vl01.w, b0.w * vl01.y)));
                  a4[M] = v;
                }
                asm volatile("" ::: "memory");
              }
            }
            if (nq <= 64) {
              j += G;
            } else if (nq < 64 + NW) {
              if (L < 64) {
                L += 64;
              } else {
                L -= 64;
                j++;
              }
            } else {
              L += 64;
              if (L >= nq) {
                L -= nq;
                j++;
              }
            }
          }
          // the pending updates' part of H g when they were not added element by element:
          // y_t += alpha_t (s.g) + beta_t (h.g), with the dot products reduced in FP64
          const bool fact = npend > 0 && npend < NPM;
          float fs[NPM], fh[NPM];
          if (fact) {
            double ds[NPM], dh[NPM];
#pragma unroll
            for (int k = 0; k < NPM; k++) ds[k] = dh[k] = 0.0;
            for (int i = tid; i < n; i += NT) {
              const double g = VG[i].y;
#pragma unroll
              for (int k = 0; k < NPM; k++) {
                if (k < npend) {
                  const float2 u = sh_slot(k)[i];
                  ds[k] += (double)u.x * g;
                  dh[k] += (double)u.y * g;
                }
              }
            }
            if constexpr (2 * NPM <= 8) {
              double v[2 * NPM];
#pragma unroll
              for (int k = 0; k < NPM; k++) {
                v[k] = ds[k];
                v[NPM + k] = dh[k];
              }
              R.template run<2 * NPM, 0>(v);
#pragma unroll
              for (int k = 0; k < NPM; k++) {
                ds[k] = v[k];
                dh[k] = v[NPM + k];
              }
            } else {
              R.template run<NPM, 0>(ds);
              R.template run<NPM, 0>(dh);
            }
#pragma unroll
            for (int k = 0; k < NPM; k++) {
              fs[k] = (float)ds[k];
              fh[k] = (float)dh[k];
            }
          } else {
            __syncthreads();  // every wave's accumulators are complete
          }
          // sum the waves' accumulators in a fixed order and clear them for the next pass
          for (int t = tid; t < n4; t += NT) {
            const int idx = t;
            float y = 0.f;
            for (int ww = 0; ww < NW; ww++)
              for (int c = 0; c < G; c++) {
                float* a = big + (size_t)ww * sacc_w + c * n4 + idx;
                y += *a;
                *a = 0.f;
              }
            if (t < n) {
              if (fact) {
#pragma unroll
                for (int k = 0; k < NPM; k++)
                  if (k < npend) {
                    const float2 ab = ab_slot(k)[t];
                    y = fmaf(ab.y, fh[k], fmaf(ab.x, fs[k], y));
                  }
              }
              // H dg = H g_new - H g_old = H g_new + d_old
              hd[t] = y + d[t];
              d[t] = y;
            }
          }
          __syncthreads();
          if (npend == NPM) npend = 0;  // every pending update is in H now
        } else {
          // full storage in HBM with rows padded to ldh = round4(n) floats (16-byte aligned).
          // 2^lq consecutive lanes share the column quad 4tq..4tq+3 and split its rows into
          // equal segments; a row of a quad is one 16-byte load and consecutive quads are
          // consecutive addresses (coalesced). H is written once every NPM iterations: when all
          // NPM update slots are pending, they are added to each element as it is read, oldest
          // first (which rounds each stored element exactly as a write-back after every update
          // would), and the rows are written back by their only owner. In the other iterations
          // the stored rows are only read and the pending updates enter H v through their
          // factored form, alpha_t (s.v) + beta_t (h.v), from dot products the same pass takes.
          // HQ rows per lane are loaded together: four fits the three-waves-per-SIMD register
          // budget (with two waves per SIMD, four and eight rows measured 178.5 and 181.1 ms on
          // the medium workload).
          constexpr int HQ = 4;
          const int nq = (n + 3) >> 2;
          int lq = 0;
          while (lq < 3 && (nq << (lq + 1)) <= NT) lq++;
          const int lpr = 1 << lq, seg = (n + lpr - 1) >> lq;
          for (int r = tid; r < (nq << lq); r += NT) {
            const int tq = r >> lq, q = r & (lpr - 1);
            const int j0 = q * seg, j1 = min(n, j0 + seg);
            float cx[NPM][4], cy[NPM][4];
            const float2* S[NPM];
#pragma unroll
            for (int k = 0; k < NPM; k++) {
              S[k] = sh_slot(k);
              const float2* ab = ab_slot(k);
#pragma unroll
              for (int u = 0; u < 4; u++) {
                const int t = 4 * tq + u;
                const float2 c = (npend > k && t < n) ? ab[t] : make_float2(0.f, 0.f);
                cx[k][u] = c.x;
                cy[k][u] = c.y;
              }
            }
            float a2[4] = {0.f, 0.f, 0.f, 0.f};
            float4* base = (float4*)H + tq;  // row j of the quad at base[j * nq]
            float lr[2][NPM];
#pragma unroll
            for (int k = 0; k < NPM; k++) lr[0][k] = lr[1][k] = 0.f;
            hquad_dispatch<HQ, NPM>(npend, base, nq, j0, j1, VG, S, cx, cy, a2, lr);
            // the pending updates' part of H g when they were not added element by element:
            // sum the dot products over the quad's lanes, then y_t += alpha_t (s.g) + beta_t (h.g)
            if (npend > 0 && npend < NPM) {
#pragma unroll
              for (int k = 0; k < NPM; k++)
#pragma unroll
                for (int c = 0; c < 2; c++) {
                  float x = lr[c][k];
                  if (lq >= 1) x += dpp_f<DPP_QUAD_XOR1>(x);
                  if (lq >= 2) x += dpp_f<DPP_QUAD_XOR2>(x);
                  if (lq >= 3) x += dpp_f<DPP_ROW_HALF_MIRROR>(x);
                  lr[c][k] = x;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
              float s2 = a2[u];
              if (lq >= 1) s2 += dpp_f<DPP_QUAD_XOR1>(s2);
              if (lq >= 2) s2 += dpp_f<DPP_QUAD_XOR2>(s2);
              if (lq >= 3) s2 += dpp_f<DPP_ROW_HALF_MIRROR>(s2);
              if (npend > 0 && npend < NPM) {
#pragma unroll
                for (int k = 0; k < NPM; k++) s2 = fmaf(cy[k][u], lr[1][k], fmaf(cx[k][u], lr[0][k], s2));
              }
              const int t = 4 * tq + u;
              if (q == 0 && t < n) {
                // H dg = H g_new - H g_old = H g_new + d_old
                hd[t] = s2 + d[t];
                d[t] = s2;
              }
            }
          }
          __syncthreads();
          if (npend == NPM) npend = 0;  // every pending update is in H now
        }
        double v6[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int i = tid; i < n; i += NT) {
          const float2 v = VG[i];
          const double xi = xo[i], h = hd[i];
          v6[0] += (double)v.x * xi;
          v6[1] += (double)v.x * h;
          v6[2] += (double)v.x * v.x;
          v6[3] += xi * xi;
          v6[4] += xi * v.y;
          v6[5] += h * v.y;
        }
        R.template run<6, 0>(v6);
        const double fac = v6[0], fae = v6[1];
        const bool upd = fac * fac > EPS_GUARD * v6[2] * v6[3] && fac > 0.0;  // :347
        if (upd) {
          // U = A s s^T + Bc (s h^T + h s^T) + Cc h h^T, the expansion of the a a^T term
          const double fi = 1.0 / fac, fei = 1.0 / fae;
          const double A = fi + fae * fi * fi, Bc = -fae * fi * fei, Cc = fae * fei * fei - fei;
          const double xig = v6[4], hdg = v6[5];
          // the new update goes behind the pending ones (slot 0 when H is kept updated)
          float2* sh_new = sh_slot(LAZY ? npend : 0);
          float2* ab_new = ab_slot(LAZY ? npend : 0);
          for (int t = tid; t < n; t += NT) {
            const double xi = xo[t], h = hd[t];
            const double al = A * xi + Bc * h, be = Bc * xi + Cc * h;
            d[t] = -(float)((double)d[t] + al * xig + be * hdg);
            sh_new[t] = make_float2((float)xi, (float)h);
            ab_new[t] = make_float2((float)al, (float)be);
          }
          npend++;
        } else {
          for (int t = tid; t < n; t += NT) d[t] = -d[t];
        }
        __syncthreads();
        if (PACKED && npend) {
          // write the update into the packed triangle: column c gains alpha_i s_c + beta_i h_c
          // in rows i <= c. Columns c and n-1-c form a pair and 2^lsh2 consecutive lanes split
          // the pair's 4-row groups evenly; each lane walks its groups in float4 steps with
          // s_c, h_c in registers (columns are 16-byte aligned).
          const int lp = 1 << lsh2;
          for (int r = tid; r < (np << lsh2); r += NT) {
            const int g = r >> lsh2, q = r & (lp - 1);
            const int c1 = g, c2 = n - 1 - g;
            const bool two = c2 != c1;
            const int l1 = (c1 + 4) >> 2, lb = l1 + (two ? (c2 + 4) >> 2 : 0);  // 4-row groups
            const int sb = (lb + lp - 1) >> lsh2, f0 = q * sb, f1 = min(lb, f0 + sb);
            if (f0 < l1) hcol_update4(H + pcol(c1), AB, SH[c1], f0, min(f1, l1));
            if (two && f1 > l1) hcol_update4(H + pcol(c2), AB, SH[c2], max(f0, l1) - l1, f1 - l1);
          }
          npend = 0;
          __syncthreads();
        }
      } else {
        // L-BFGS history and two-loop recursion (mmff_lbfgs_source_tg.metal:249-292)
        double ys = 0.0;
        for (int i = tid; i < n; i += NT) ys += (double)VG[i].x * xo[i];
        ys = R.sum(ys);
        if (ys > 1e-10) {
          const int slot = hist_idx % M;
          for (int i = tid; i < n; i += NT) {
            S[(size_t)slot * mn + i] = xo[i];
            Y[(size_t)slot * mn + i] = VG[i].x;
          }
          if (tid == 0) rho[slot] = 1.0 / ys;
          hist_idx++;
          if (hist_count < M) hist_count++;
        }
        for (int i = tid; i < n; i += NT) hd[i] = VG[i].y;
        __syncthreads();  // rho visible
        // The two-loop recursion with each axpy fused into the next dot product's pass (and
        // gamma's dots into the first pass, the gamma scaling into the last axpy of the first
        // loop): the same FP32 updates of q and the same FP64 dot products, element by
        // element and lane by lane, so results are bit-identical to running the passes one
        // by one, in 2m + 1 passes over the history instead of 4m + 3.
        const int h = hist_count;
        if (h == 0) {
          for (int i = tid; i < n; i += NT) d[i] = -hd[i];
        } else {
          int sl = (hist_idx - 1) % M;  // newest entry (j = h - 1)
          if (sl < 0) sl += M;
          double v3[3] = {0.0, 0.0, 0.0};  // s_{h-1}.q, and s.y, y.y of the newest pair for gamma
          for (int i = tid; i < n; i += NT) {
            const double s_ = S[(size_t)sl * mn + i], y_ = Y[(size_t)sl * mn + i];
            v3[0] += s_ * hd[i];
            v3[1] += s_ * y_;
            v3[2] += y_ * y_;
          }
          R.template run<3, 0>(v3);
          const float gam = (float)(v3[1] / fmax(v3[2], 1e-30));
          double aj = rho[sl] * v3[0];
          if (tid == 0) alf[h - 1] = aj;
          for (int j = h - 1; j >= 1; j--) {  // q -= alpha_j y_j, then s_{j-1}.q
            const int sn = sl == 0 ? M - 1 : sl - 1;
            const float af = (float)aj;
            double sq = 0.0;
            for (int i = tid; i < n; i += NT) {
              const float q = hd[i] - af * Y[(size_t)sl * mn + i];
              hd[i] = q;
              sq += (double)S[(size_t)sn * mn + i] * q;
            }
            sq = R.sum(sq);
            aj = rho[sn] * sq;
            if (tid == 0) alf[j - 1] = aj;
            sl = sn;
          }
          // q -= alpha_0 y_0, r = gamma q, then y_0.r (sl is entry 0's slot)
          double yq = 0.0;
          {
            const float af = (float)aj;
            for (int i = tid; i < n; i += NT) {
              const float y_ = Y[(size_t)sl * mn + i];
              float q = hd[i] - af * y_;
              q *= gam;
              hd[i] = q;
              yq += (double)y_ * q;
            }
          }
          yq = R.sum(yq);  // its barrier also makes alf visible
          for (int j = 0; j < h; j++) {  // r += (alpha_j - beta_j) s_j, then y_{j+1}.r
            const float cf = (float)(alf[j] - rho[sl] * yq);
            const int sn = sl + 1 == M ? 0 : sl + 1;
            if (j + 1 < h) {
              double y2 = 0.0;
              for (int i = tid; i < n; i += NT) {
                const float r = hd[i] + cf * S[(size_t)sl * mn + i];
                hd[i] = r;
                y2 += (double)Y[(size_t)sn * mn + i] * r;
              }
              yq = R.sum(y2);
            } else {
              for (int i = tid; i < n; i += NT) {
                const float r = hd[i] + cf * S[(size_t)sl * mn + i];
                hd[i] = r;
                d[i] = -r;
              }
            }
            sl = sn;
          }
        }
        __syncthreads();
      }
    }
    for (int i = tid; i < n; i += NT) P.x_out[off + i] = xs[i];
    if (tid == 0) {
      P.e_out[c] = E;
      P.status[c] = status;
      if (P.counters) {
        P.counters[3 * c] = iters;
        P.counters[3 * c + 1] = n_grads;
        P.counters[3 * c + 2] = n_evals;
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------
// BB fused steepest descent, MEGA family: mmff_metal_kernel.py:845-1075.
template <int NT>
__global__ void __launch_bounds__(NT) k_bb_fused(KBatch B, const float* __restrict__ x_in, float* x_out,
                                                 double* e_out, int max_it, float max_step, float grad_lr) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NW = NT / 64;
  const int c = blockIdx.x;
  if (c >= B.n_confs) return;
  const int tid = threadIdx.x;
  const int mx = 3 * B.max_atoms;
  double* red = (double*)smem;
  unsigned long long* gacc = (unsigned long long*)(red + 16 * NW);
  float* xs = (float*)(gacc + mx);
  float* pp = xs + mx;
  float* pg = pp + mx;
  float* g = pg + mx;
  const MolDesc m = load_desc(B.mol_desc + 16 * B.conf_mol[c]);
  const int n = 3 * m.n_atoms;
  const int off = B.conf_off[c];
  for (int i = tid; i < n; i += NT) {
    xs[i] = x_in[off + i];
    pp[i] = xs[i];
    pg[i] = 0.f;
  }
  Reducer<NT> R{red, 0};
  float alpha = grad_lr;
  double E_prev = 1e30, E = 0.0;
  for (int it = 0; it < max_it; it++) {
    for (int i = tid; i < n; i += NT) gacc[i] = 0ull;
    __syncthreads();
    E = R.sum(eval_terms<FAM_MEGA, true, false, NT>(global_view(B, m), m, xs, gacc, 1, 0x7fu, nullptr));
    if (it > 0 && E > E_prev + 1.0) {  // energy guard :1031-1038
      for (int i = tid; i < n; i += NT) xs[i] = pp[i];
      alpha = fmaxf(alpha * 0.5f, 1e-7f);
      E = E_prev;
      continue;
    }
    for (int i = tid; i < n; i += NT) g[i] = gread(gacc, i, m.n_atoms);
    if (it > 0) {  // BB step :1041-1052
      double v2[2] = {0.0, 0.0};
      for (int i = tid; i < n; i += NT) {
        const float si = xs[i] - pp[i], yi = g[i] - pg[i];
        v2[0] += (double)si * si;
        v2[1] += (double)si * yi;
      }
      R.template run<2, 0>(v2);
      if (v2[1] > 1e-20) alpha = fminf(fmaxf((float)(v2[0] / v2[1]), 1e-6f), 0.3f);
    }
    E_prev = E;
    for (int i = tid; i < n; i += NT) {
      pp[i] = xs[i];
      pg[i] = g[i];
    }
    __syncthreads();
    for (int a = tid; a < m.n_atoms; a += NT) {  // per-atom capped step :1063-1071
      const float gx = g[3 * a], gy = g[3 * a + 1], gz = g[3 * a + 2];
      const float gn = sqrtf(gx * gx + gy * gy + gz * gz + 1e-30f);
      const float s = fminf(max_step / gn, alpha);
      xs[3 * a] -= s * gx;
      xs[3 * a + 1] -= s * gy;
      xs[3 * a + 2] -= s * gz;
    }
    __syncthreads();
  }
  __syncthreads();
  for (int i = tid; i < n; i += NT) x_out[off + i] = xs[i];
  if (tid == 0) e_out[c] = E;
}

// ------------------------------------------------------------------------------------
// multi_mol lock-step L-BFGS (MEGA family): mmff_optimize_metal_multi_mol
// (mmff_metal_optimizer.py:217-535). State per conformer in a global workspace; each
// launch runs `steps` iterations; launches that end at a sync point apply the best-x guard
// and convergence test and count unconverged conformers in a flag that the next launch
// reads to skip the rest (the reference's break), so no host round trip is needed.
struct MMState {
  float* x;       // [total]
  float* g;       // [total] scaled gradient
  float* bx;      // [total] best x
  float* S;       // [m][total]
  float* Y;       // [m][total]
  double* rho;    // [m][n_confs]
  double* best_e; // [n_confs]
  double* e_new;  // [n_confs]
  int32_t* ring;  // [2*n_confs] len, head
  int32_t* conv;  // [n_confs]
  int32_t* pci;   // [n_confs]
  int32_t* flags; // [n_flags]
  size_t total;
};

template <int NT>
__device__ void mm_eval(const KBatch& B, const MolDesc& m, float* xs, unsigned long long* gacc, Reducer<NT>& R,
                        double& E, int n) {
  for (int i = threadIdx.x; i < n; i += NT) gacc[i] = 0ull;
  __syncthreads();
  E = R.sum(eval_terms<FAM_MEGA, true, false, NT>(global_view(B, m), m, xs, gacc, 1, 0x7fu, nullptr));
}

// mode 0: initial (x = x_in), mode 1: polish restart (x = best_x); both compute E+G, scale g,
// reset history; mode 0 also initialises best_e/best_x/conv/pci.
template <int NT>
__global__ void __launch_bounds__(NT) k_mm_init(KBatch B, MMState S, const float* x_in, float gsc, int mode,
                                                int max_iters) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NW = NT / 64;
  const int c = blockIdx.x;
  if (c >= B.n_confs) return;
  const int mx = 3 * B.max_atoms;
  double* red = (double*)smem;
  unsigned long long* gacc = (unsigned long long*)(red + 16 * NW);
  float* xs = (float*)(gacc + mx);
  const MolDesc m = load_desc(B.mol_desc + 16 * B.conf_mol[c]);
  const int n = 3 * m.n_atoms, off = B.conf_off[c];
  for (int i = threadIdx.x; i < n; i += NT) xs[i] = mode == 0 ? x_in[off + i] : S.bx[off + i];
  Reducer<NT> R{red, 0};
  double E;
  mm_eval<NT>(B, m, xs, gacc, R, E, n);
  for (int i = threadIdx.x; i < n; i += NT) {
    S.x[off + i] = xs[i];
    S.g[off + i] = gread(gacc, i, m.n_atoms) * gsc;
    if (mode == 0) S.bx[off + i] = xs[i];
  }
  if (threadIdx.x == 0) {
    S.ring[2 * c] = 0;
    S.ring[2 * c + 1] = 0;
    if (mode == 0) {
      S.best_e[c] = E;
      S.conv[c] = 0;
      S.pci[c] = max_iters;
    }
    S.e_new[c] = E;
  }
}

template <int NT>
__global__ void __launch_bounds__(NT) k_mm_chunk(KBatch B, MMState S, int steps, int it_end, int do_sync,
                                                 int phase, const int32_t* prev_flag, int32_t* my_flag, int M,
                                                 float max_step, float gsc, float tol) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NW = NT / 64;
  const int c = blockIdx.x;
  if (c >= B.n_confs) return;
  if (prev_flag && *prev_flag == 0) return;  // every conformer converged at the last sync
  const int tid = threadIdx.x;
  const int mx = 3 * B.max_atoms;
  double* red = (double*)smem;
  unsigned long long* gacc = (unsigned long long*)(red + 16 * NW);
  float* xs = (float*)(gacc + mx);
  float* gs = xs + mx;
  float* q = gs + mx;
  float* xn = q + mx;
  double* alf = (double*)(xn + mx);  // [M]
  const MolDesc m = load_desc(B.mol_desc + 16 * B.conf_mol[c]);
  const int n = 3 * m.n_atoms, off = B.conf_off[c];
  const size_t T = S.total;
  for (int i = tid; i < n; i += NT) {
    xs[i] = S.x[off + i];
    gs[i] = S.g[off + i];
  }
  int len = S.ring[2 * c], head = S.ring[2 * c + 1];
  Reducer<NT> R{red, 0};
  double E = S.e_new[c];
  __syncthreads();
  for (int st = 0; st < steps; st++) {
    // two-loop direction, mmff_metal_optimizer.py:47-74 (history oldest -> newest), with each
    // axpy fused into the next dot product's pass, gamma's dots into the first pass and the
    // gamma scaling into the first loop's last axpy: the same FP32 updates of q and FP64 dot
    // products as running the passes one by one (bit-identical), in 2 len + 1 passes.
    for (int i = tid; i < n; i += NT) q[i] = gs[i];
    if (len > 0) {
      int sl = (head + len - 1) % M;  // newest entry
      double v3[3] = {0.0, 0.0, 0.0};
      for (int i = tid; i < n; i += NT) {
        const double s = S.S[sl * T + off + i], y = S.Y[sl * T + off + i];
        v3[0] += s * q[i];
        v3[1] += s * y;
        v3[2] += y * y;
      }
      R.template run<3, 0>(v3);
      const float gf = (float)(v3[1] / (v3[2] + 1e-30));
      double a = S.rho[(size_t)sl * B.n_confs + c] * v3[0];
      if (tid == 0) alf[len - 1] = a;
      for (int h = len - 1; h >= 1; h--) {  // q -= alpha_h y_h, then s_{h-1}.q
        const int sn = (head + h - 1) % M;
        const float af = (float)a;
        double sq = 0.0;
        for (int i = tid; i < n; i += NT) {
          const float qv = q[i] - af * S.Y[sl * T + off + i];
          q[i] = qv;
          sq += (double)S.S[sn * T + off + i] * qv;
        }
        sq = R.sum(sq);
        a = S.rho[(size_t)sn * B.n_confs + c] * sq;
        if (tid == 0) alf[h - 1] = a;
        sl = sn;
      }
      double yr = 0.0;  // q -= alpha_0 y_0, q *= gamma, then y_0.q
      {
        const float af = (float)a;
        for (int i = tid; i < n; i += NT) {
          const float y = S.Y[sl * T + off + i];
          float qv = q[i] - af * y;
          qv *= gf;
          q[i] = qv;
          yr += (double)y * qv;
        }
      }
      yr = R.sum(yr);  // its barrier also makes alf visible
      for (int h = 0; h < len; h++) {  // q += (alpha_h - beta_h) s_h, then y_{h+1}.q
        const float cf = (float)(alf[h] - S.rho[(size_t)sl * B.n_confs + c] * yr);
        const int sn = (head + h + 1) % M;
        if (h + 1 < len) {
          double y2 = 0.0;
          for (int i = tid; i < n; i += NT) {
            const float r = q[i] + cf * S.S[sl * T + off + i];
            q[i] = r;
            y2 += (double)S.Y[sn * T + off + i] * r;
          }
          yr = R.sum(y2);
        } else {
          for (int i = tid; i < n; i += NT) q[i] += cf * S.S[sl * T + off + i];
        }
        sl = sn;
      }
    }
    __syncthreads();
    // d = -q with the per-atom cap of max_step (:365-373); x_new = x + d
    for (int a = tid; a < m.n_atoms; a += NT) {
      const float dx = -q[3 * a], dy = -q[3 * a + 1], dz = -q[3 * a + 2];
      const float disp = sqrtf(dx * dx + dy * dy + dz * dz + 1e-30f);
      const float f = fminf(max_step / disp, 1.0f);
      xn[3 * a] = xs[3 * a] + dx * f;
      xn[3 * a + 1] = xs[3 * a + 1] + dy * f;
      xn[3 * a + 2] = xs[3 * a + 2] + dz * f;
    }
    __syncthreads();
    mm_eval<NT>(B, m, xn, gacc, R, E, n);
    // push (s, y, rho) into the ring (:385-399); rho = 0 on bad curvature
    int sl;
    if (len >= M) {
      sl = head;
      head = (head + 1) % M;
    } else {
      sl = (head + len) % M;
      len++;
    }
    double sy = 0.0;
    for (int i = tid; i < n; i += NT) {
      const float gn = gread(gacc, i, m.n_atoms) * gsc;
      const float s = xn[i] - xs[i], y = gn - gs[i];
      S.S[sl * T + off + i] = s;
      S.Y[sl * T + off + i] = y;
      sy += (double)s * y;
      xs[i] = xn[i];
      gs[i] = gn;
    }
    sy = R.sum(sy);
    if (tid == 0) S.rho[(size_t)sl * B.n_confs + c] = sy > 1e-10 ? 1.0 / (sy + 1e-30) : 0.0;
    __syncthreads();
  }
  if (do_sync) {
    // best-x guard (:408-412) and convergence (:354-359 main, :477-481 polish)
    const bool improved = E < S.best_e[c];
    double gmax = 0.0;
    for (int a = tid; a < m.n_atoms; a += NT) {
      const float gx = gs[3 * a], gy = gs[3 * a + 1], gz = gs[3 * a + 2];
      gmax = fmax(gmax, (double)sqrtf(gx * gx + gy * gy + gz * gz));
    }
    gmax = R.max(gmax);
    if (improved)
      for (int i = tid; i < n; i += NT) S.bx[off + i] = xs[i];
    if (tid == 0) {
      if (improved) S.best_e[c] = E;
      const bool under = gmax <= (double)tol;
      if (phase == 0) {
        if (!S.conv[c] && under) {
          S.conv[c] = 1;
          S.pci[c] = it_end;
        }
        if (!S.conv[c]) atomicAdd(my_flag, 1);
      } else if (!under) {
        atomicAdd(my_flag, 1);
      }
    }
  }
  for (int i = tid; i < n; i += NT) {
    S.x[off + i] = xs[i];
    S.g[off + i] = gs[i];
  }
  if (tid == 0) {
    S.ring[2 * c] = len;
    S.ring[2 * c + 1] = head;
    S.e_new[c] = E;
  }
}

// l2_norm: grad_norm = ||g_sc||_2 of the last iterate (mmff_optimize_metal_batch,
// mmff_metal_optimizer.py:186) instead of its largest per-atom norm (multi_mol, :504-525)
template <int NT>
__global__ void __launch_bounds__(NT) k_mm_final(KBatch B, MMState S, float* x_best, double* e_final,
                                                 float* grad_norm, int32_t* pci, int32_t* conv, int32_t* iters_used,
                                                 int n_main_flags, int eval_freq, int max_iters, int l2_norm) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NW = NT / 64;
  const int c = blockIdx.x;
  if (c >= B.n_confs) return;
  const int mx = 3 * B.max_atoms;
  double* red = (double*)smem;
  unsigned long long* gacc = (unsigned long long*)(red + 16 * NW);
  float* xs = (float*)(gacc + mx);
  const MolDesc m = load_desc(B.mol_desc + 16 * B.conf_mol[c]);
  const int n = 3 * m.n_atoms, off = B.conf_off[c];
  for (int i = threadIdx.x; i < n; i += NT) xs[i] = S.bx[off + i];
  __syncthreads();
  Reducer<NT> R{red, 0};
  double E = R.sum(eval_terms<FAM_MEGA, false, false, NT>(global_view(B, m), m, xs, gacc, 1, 0x7fu, nullptr));
  double gmax = 0.0;
  if (l2_norm) {
    for (int i = threadIdx.x; i < n; i += NT) gmax += (double)S.g[off + i] * S.g[off + i];
    gmax = sqrt(R.sum(gmax));
  } else {
    for (int a = threadIdx.x; a < m.n_atoms; a += NT) {
      const float gx = S.g[off + 3 * a], gy = S.g[off + 3 * a + 1], gz = S.g[off + 3 * a + 2];
      gmax = fmax(gmax, (double)sqrtf(gx * gx + gy * gy + gz * gz));
    }
    gmax = R.max(gmax);
  }
  for (int i = threadIdx.x; i < n; i += NT) x_best[off + i] = xs[i];
  if (threadIdx.x == 0) {
    e_final[c] = E;
    grad_norm[c] = (float)gmax;
    pci[c] = S.pci[c];
    conv[c] = S.conv[c];
    if (c == 0) {
      int used = max_iters;
      for (int k = 0; k < n_main_flags; k++)
        if (S.flags[k] == 0) {
          used = (k + 1) * eval_freq;
          break;
        }
      *iters_used = used;
    }
  }
}

__global__ void k_fill_i32(int32_t* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------------------
// host helpers
struct DevInfo {
  int n_cu = 0, lds = 0, ok = 0;
};

DevInfo dev_info() {
  static thread_local DevInfo cache[16];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return DevInfo{};
  if (dev < 0 || dev >= 16) return DevInfo{};
  DevInfo& di = cache[dev];
  if (!di.ok) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) {
      di.n_cu = p.multiProcessorCount;
      di.lds = (int)p.sharedMemPerBlock;
      di.ok = 1;
    }
  }
  return di;
}

// allow dynamic LDS above the 64 KB default for a kernel (gfx950 has 160 KB per workgroup)
void allow_lds(const void* kernel, size_t bytes) {
  if (bytes > 65536) hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

int check_batch(const mmffhip_batch* b) {
  if (!b) return fail(MMFFHIP_ERR_ARG, "batch is NULL");
  if (b->n_confs < 0 || b->n_mols < 0) return fail(MMFFHIP_ERR_ARG, "negative batch size");
  if (b->n_confs > 0 && (!b->mol_desc || !b->conf_mol || !b->conf_off))
    return fail(MMFFHIP_ERR_ARG, "batch descriptor pointers are NULL");
  if (b->max_atoms <= 0 && b->n_confs > 0) return fail(MMFFHIP_ERR_ARG, "max_atoms must be positive");
  if (b->max_atoms > 65535) return fail(MMFFHIP_ERR_SIZE, "molecules above 65535 atoms are not supported");
  return MMFFHIP_OK;
}

size_t eg_lds(int nt, int max_atoms) {
  return 16 * (nt / 64) * 8 + (size_t)3 * max_atoms * (8 + 4) + 16;
}

template <int FAM>
int launch_eg(const KBatch& kb, const float* x, double* e, double* et, float* g, uint32_t mask, int diel,
              hipStream_t s) {
  const size_t lds = eg_lds(256, kb.max_atoms);
  if ((int)lds > dev_info().lds) return fail(MMFFHIP_ERR_SIZE, "molecule too large for the LDS of one workgroup");
  allow_lds((const void*)&k_energy_grad<FAM, 256>, lds);
  hipLaunchKernelGGL((k_energy_grad<FAM, 256>), dim3(kb.n_confs), dim3(256), lds, s, kb, x, e, et, g, mask, diel);
  HIP_TRY(hipGetLastError());
  return MMFFHIP_OK;
}

int round4(int v) { return (v + 3) & ~3; }

struct NkPlan {
  int nt = 256;
  int hm = 0;  // Hessian placement of k_minimize_nk (0 packed LDS, 1 full HBM, 2 packed HBM, 3 symmetric HBM)
  int tl = 0;  // term tables: 0 global, 1 all staged in LDS, 2 bonded staged
  int tbytes = 0;
  size_t lds = 0;
  int grid = 0;
  size_t slot_floats = 0;
  int max_n = 0;
};

template <int ALGO, int HM, int TL, int NT>
const void* nk_kernel_ptr() {
  return (const void*)&k_minimize_nk<ALGO, HM, TL, NT>;
}

#define NK_FOR_EACH(X)                                                                                   \
  X(0, 0, 1, 64) X(0, 0, 1, 128) X(0, 0, 1, 192) X(0, 0, 1, 256) X(0, 0, 1, 384) X(0, 0, 1, 512)        \
  X(0, 0, 0, 64) X(0, 0, 0, 128) X(0, 0, 0, 192) X(0, 0, 0, 256) X(0, 0, 0, 320) X(0, 0, 0, 384)        \
  X(0, 0, 0, 512) X(0, 0, 0, 640) X(0, 0, 0, 768) X(0, 0, 0, 1024)                                     \
  X(0, 1, 0, 64) X(0, 1, 0, 128) X(0, 1, 0, 192) X(0, 1, 0, 256) X(0, 1, 0, 320) X(0, 1, 0, 512)        \
  X(0, 1, 0, 640) X(0, 1, 0, 768) X(0, 1, 0, 1024)                                                     \
  X(0, 1, 2, 64) X(0, 1, 2, 128) X(0, 1, 2, 192) X(0, 1, 2, 256) X(0, 1, 2, 320) X(0, 1, 2, 512)       \
  X(0, 2, 0, 256) X(0, 2, 0, 512) X(0, 2, 0, 640) X(0, 2, 0, 1024)                                     \
  X(0, 3, 0, 64) X(0, 3, 0, 128) X(0, 3, 0, 256) X(0, 3, 0, 512)                                        \
  X(0, 3, 2, 64) X(0, 3, 2, 128) X(0, 3, 2, 256) X(0, 3, 2, 512)                                        \
  X(1, 0, 1, 64) X(1, 0, 1, 128) X(1, 0, 1, 256)                                                       \
  X(1, 0, 0, 128) X(1, 0, 0, 256) X(1, 0, 0, 512)                                                      \
  X(1, 1, 0, 128) X(1, 1, 0, 192) X(1, 1, 0, 256) X(1, 1, 0, 512) X(1, 1, 2, 256)

const void* nk_kernel(int algo, int hm, int tl, int nt) {
#define NKSEL(A, H, TL, T) \
  if (algo == A && hm == H && tl == TL && nt == T) return nk_kernel_ptr<A, H, TL, T>();
  NK_FOR_EACH(NKSEL)
#undef NKSEL
  return nullptr;
}

int make_nk_plan(const mmffhip_batch* b, const mmffhip_nk_opts* o, NkPlan& pl) {
  if (!o) return fail(MMFFHIP_ERR_ARG, "options are NULL");
  if (o->algo != MMFFHIP_BFGS && o->algo != MMFFHIP_LBFGS) return fail(MMFFHIP_ERR_ARG, "unknown algo");
  if (o->algo == MMFFHIP_LBFGS && (o->lbfgs_m < 1 || o->lbfgs_m > LBFGS_MAX_M))
    return fail(MMFFHIP_ERR_ARG, "lbfgs_m must be in [1, 16]");
  if (o->diel != 1 && o->diel != 2) return fail(MMFFHIP_ERR_ARG, "diel must be 1 or 2");
  if (o->max_iters < 0) return fail(MMFFHIP_ERR_ARG, "max_iters must be >= 0");
  DevInfo di = dev_info();
  if (!di.ok) return fail(MMFFHIP_ERR_DEVICE, "no HIP device");
  pl.nt = o->block > 0 ? o->block : (3 * b->max_atoms > 256 ? 512 : 256);
  if (pl.nt < 64 || pl.nt > 1024 || pl.nt % 64 != 0)
    return fail(MMFFHIP_ERR_ARG, "block must be a multiple of 64 in [64, 1024]");
  pl.max_n = round4(3 * std::max(b->max_atoms, 1));
  const int m = o->algo == MMFFHIP_LBFGS ? o->lbfgs_m : 0;
  const size_t lds_in = nk_lds_bytes(o->algo, 0, pl.nt, pl.max_n, m);
  if (o->hess_place == MMFFHIP_HESS_LDS) {
    pl.hm = 0;
  } else if (o->hess_place == MMFFHIP_HESS_GLOBAL) {
    pl.hm = 1;
  } else if (o->hess_place == MMFFHIP_HESS_GLOBAL_PACKED) {
    pl.hm = o->algo == MMFFHIP_BFGS ? 2 : 1;
  } else if (o->hess_place == MMFFHIP_HESS_GLOBAL_SYM) {
    pl.hm = o->algo == MMFFHIP_BFGS ? 3 : 1;
  } else if (o->hess_place == MMFFHIP_HESS_AUTO) {
    // dense BFGS: the symmetric blocks in HBM (they beat the LDS triangle even where it
    // fits); L-BFGS: the history in LDS when it fits, else in HBM
    pl.hm = o->algo == MMFFHIP_BFGS ? 3 : lds_in <= (size_t)di.lds ? 0 : 1;
  } else {
    return fail(MMFFHIP_ERR_ARG, "unknown hess_place");
  }
  pl.lds = nk_lds_bytes(o->algo, pl.hm, pl.nt, pl.max_n, m);
  if (pl.lds > (size_t)di.lds) return fail(MMFFHIP_ERR_SIZE, "plan exceeds the LDS of one workgroup");
  // Stage term tables in LDS when asked and when they fit: all of them next to an LDS
  // Hessian; only the bonded ones beside a global Hessian, and only when that keeps the
  // unstaged kernel's workgroups per CU (there, occupancy hides the pair loads).
  if (o->table_lds_bytes > 0) {
    const int tl = pl.hm == 0 ? 1 : 2;
    const size_t with = nk_lds_bytes(o->algo, pl.hm, pl.nt, pl.max_n, m, (size_t)o->table_lds_bytes);
    const void* kt = nk_kernel(o->algo, pl.hm, tl, pl.nt);
    bool take = with <= (size_t)di.lds && kt != nullptr;
    if (take && tl == 2) {
      const void* k0 = nk_kernel(o->algo, pl.hm, 0, pl.nt);
      int pc0 = 0, pc1 = 0;
      allow_lds(kt, with);
      if (k0) allow_lds(k0, pl.lds);
      take = k0 && hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc0, k0, pl.nt, pl.lds) == hipSuccess &&
             hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc1, kt, pl.nt, with) == hipSuccess && pc1 >= pc0;
    }
    if (take) {
      pl.tl = tl;
      pl.tbytes = ((o->table_lds_bytes + 15) / 16) * 16;
      pl.lds = with;
    }
  }
  const void* kp = nk_kernel(o->algo, pl.hm, pl.tl, pl.nt);
  if (!kp) return fail(MMFFHIP_ERR_ARG, "no kernel instantiated for this (algo, placement, block) combination");
  allow_lds(kp, pl.lds);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kp, pl.nt, pl.lds) != hipSuccess || per_cu < 1)
    per_cu = 1;
  pl.grid = std::max(1, std::min(b->n_confs, di.n_cu * per_cu));
  pl.slot_floats = 0;
  if (pl.hm != 0) {
    pl.slot_floats = o->algo == MMFFHIP_LBFGS ? (size_t)2 * m * pl.max_n
                     : pl.hm == 2             ? packed_floats(pl.max_n)
                     : pl.hm == 3             ? sym_slot_floats(pl.max_n)
                                              : (size_t)pl.max_n * pl.max_n;
    pl.slot_floats = (pl.slot_floats + 63) & ~(size_t)63;
  }
  return MMFFHIP_OK;
}

size_t nk_ws_bytes(const NkPlan& pl) { return 256 + (size_t)pl.grid * pl.slot_floats * 4; }

template <int ALGO, int HM, int TL, int NT>
void launch_nk_t(const KBatch& kb, const NkParams& P, const NkPlan& pl, hipStream_t s) {
  hipLaunchKernelGGL((k_minimize_nk<ALGO, HM, TL, NT>), dim3(pl.grid), dim3(NT), pl.lds, s, kb, P);
}

void launch_nk(int algo, const KBatch& kb, const NkParams& P, const NkPlan& pl, hipStream_t s) {
#define NKL(A, H, TL, T)                                          \
  if (algo == A && pl.hm == H && pl.tl == TL && pl.nt == T) {     \
    launch_nk_t<A, H, TL, T>(kb, P, pl, s);                       \
    return;                                                       \
  }
  NK_FOR_EACH(NKL)
#undef NKL
}

// Clear the work queue at the front of `workspace` with a fill kernel (graph-capture safe)
// and launch the plan's minimize kernel; the global slots follow the 256-byte queue header.
int run_nk(const mmffhip_batch* b, const mmffhip_nk_opts* o, const NkPlan& pl, const float* x_in, float* x_out,
           double* e_out, int32_t* status_out, int32_t* counters, const int32_t* order, void* workspace,
           hipStream_t s) {
  int32_t* queue = (int32_t*)workspace;
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, queue, 8, 0);
  HIP_TRY(hipGetLastError());
  NkParams P;
  P.x_in = x_in;
  P.x_out = x_out;
  P.e_out = e_out;
  P.status = status_out;
  P.counters = counters;
  P.order = order;
  P.queue = queue;
  P.hglob = (float*)((char*)workspace + 256);
  P.slot_floats = pl.slot_floats;
  P.precond = o->precond;
  P.max_iters = o->max_iters;
  P.grad_tol = o->grad_tol;
  P.diel = o->diel;
  P.m = o->algo == MMFFHIP_LBFGS ? o->lbfgs_m : 1;
  P.max_n = pl.max_n;
  P.ls_grad = o->ls_grad;
  P.tbytes = pl.tbytes;
  P.xcd_on = 0;
  for (int x = 0; x < 9; x++) P.xcd_off[x] = 0;
  if (o->xcd_offsets) {
    if (!order) return fail(MMFFHIP_ERR_ARG, "xcd_offsets need an order");
    int prev = 0;
    for (int x = 0; x < 9; x++) {
      const int v = o->xcd_offsets[x];
      if (v < prev || v > b->n_confs || (x == 0 && v != 0) || (x == 8 && v != b->n_confs))
        return fail(MMFFHIP_ERR_ARG, "xcd_offsets must rise from 0 to n_confs");
      P.xcd_off[x] = v;
      prev = v;
    }
    P.xcd_on = 1;
  }
  KBatch kb = to_kbatch(b);
  launch_nk(o->algo, kb, P, pl, s);
  HIP_TRY(hipGetLastError());
  return MMFFHIP_OK;
}

// Scratch of the allocation-free minimize entry point: the work-queue header and the
// per-workgroup global slots live in this file-scope device array (one copy per GPU, held
// until the process ends). Host code never passes `scratch` itself to a kernel: its address
// on each device comes from hipGetSymbolAddress, taken once per device on that device's
// first call.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
void* g_scratch_ptr[MAX_DEVICES];
std::mutex g_scratch_mu;

int scratch_ptr(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail(MMFFHIP_ERR_DEVICE, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(MMFFHIP_ERR_HIP, "hipGetSymbolAddress returned NULL for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return MMFFHIP_OK;
}

}  // namespace

// error reporting shared with the library's other translation units (not exported)
__attribute__((visibility("hidden"))) int mmffhip_set_error(int code, const char* msg) { return fail(code, msg); }

// ====================================================================================
// extern "C" ABI
extern "C" {

int mmffhip_abi_version(void) { return MMFFHIP_ABI_VERSION; }

const char* mmffhip_last_error(void) { return g_last_error.c_str(); }

int mmffhip_device_info(int32_t* n_cu, int32_t* n_xcd, int32_t* lds_bytes) {
  DevInfo di = dev_info();
  if (!di.ok) return fail(MMFFHIP_ERR_DEVICE, "no HIP device");
  if (n_cu) *n_cu = di.n_cu;
  if (n_xcd) *n_xcd = 8;
  if (lds_bytes) *lds_bytes = di.lds;
  return MMFFHIP_OK;
}

int mmffhip_energy_grad(const mmffhip_batch* b, int family, int diel, const float* x, double* e_out,
                        double* e_terms, float* g_out, uint32_t term_mask, void* stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (b->n_confs == 0) return MMFFHIP_OK;
  if (!x || !e_out) return fail(MMFFHIP_ERR_ARG, "x and e_out are required");
  if (diel != 1 && diel != 2) return fail(MMFFHIP_ERR_ARG, "diel must be 1 or 2");
  if (family == MMFFHIP_FAMILY_MEGA && diel != 1)
    return fail(MMFFHIP_ERR_ARG, "the MEGA family has a constant dielectric only");
  KBatch kb = to_kbatch(b);
  hipStream_t s = (hipStream_t)stream;
  if (family == MMFFHIP_FAMILY_TG) return launch_eg<FAM_TG>(kb, x, e_out, e_terms, g_out, term_mask, diel, s);
  if (family == MMFFHIP_FAMILY_MEGA) return launch_eg<FAM_MEGA>(kb, x, e_out, e_terms, g_out, term_mask, 1, s);
  return fail(MMFFHIP_ERR_ARG, "unknown family");
}

int mmffhip_minimize_nk_workspace(const mmffhip_batch* b, const mmffhip_nk_opts* o, size_t* bytes) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!bytes) return fail(MMFFHIP_ERR_ARG, "bytes is NULL");
  if (b->n_confs == 0) {
    *bytes = 256;
    return MMFFHIP_OK;
  }
  NkPlan pl;
  rc = make_nk_plan(b, o, pl);
  if (rc) return rc;
  *bytes = nk_ws_bytes(pl);
  return MMFFHIP_OK;
}

int mmffhip_minimize_nk_plan(const mmffhip_batch* b, const mmffhip_nk_opts* o, int32_t* grid, int32_t* block,
                             int32_t* lds_bytes, int32_t* hess_place) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!grid) return fail(MMFFHIP_ERR_ARG, "grid is NULL");
  NkPlan pl;
  rc = make_nk_plan(b, o, pl);
  if (rc) return rc;
  *grid = b->n_confs == 0 ? 0 : pl.grid;
  if (block) *block = pl.nt;
  if (lds_bytes) *lds_bytes = (int32_t)pl.lds;
  if (hess_place)
    *hess_place = pl.hm == 0   ? MMFFHIP_HESS_LDS
                  : pl.hm == 2 ? MMFFHIP_HESS_GLOBAL_PACKED
                  : pl.hm == 3 ? MMFFHIP_HESS_GLOBAL_SYM
                               : MMFFHIP_HESS_GLOBAL;
  return MMFFHIP_OK;
}

int mmffhip_minimize_nk(const mmffhip_batch* b, const mmffhip_nk_opts* o, const float* x_in, float* x_out,
                        double* e_out, int32_t* status_out, int32_t* counters, const int32_t* order,
                        void* workspace, size_t workspace_bytes, void* stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (b->n_confs == 0) return MMFFHIP_OK;
  if (!x_in || !x_out || !e_out || !status_out || !workspace)
    return fail(MMFFHIP_ERR_ARG, "x_in, x_out, e_out, status_out and workspace are required");
  NkPlan pl;
  rc = make_nk_plan(b, o, pl);
  if (rc) return rc;
  if (workspace_bytes < nk_ws_bytes(pl)) return fail(MMFFHIP_ERR_SIZE, "workspace too small");
  if ((uintptr_t)workspace & 255) return fail(MMFFHIP_ERR_ARG, "workspace must be 256-byte aligned");
  return run_nk(b, o, pl, x_in, x_out, e_out, status_out, counters, order, workspace, (hipStream_t)stream);
}

int mmffhip_minimize_nk_noalloc(const mmffhip_batch* b, const mmffhip_nk_opts* o, const float* x_in, float* x_out,
                                double* e_out, int32_t* status_out, int32_t* counters, const int32_t* order,
                                void* stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (b->n_confs == 0) return MMFFHIP_OK;
  if (!x_in || !x_out || !e_out || !status_out)
    return fail(MMFFHIP_ERR_ARG, "x_in, x_out, e_out and status_out are required");
  NkPlan pl;
  rc = make_nk_plan(b, o, pl);
  if (rc) return rc;
  // The persistent grid is capped to the global slots that fit beside the queue header; a
  // plan needs at least one slot per compute unit (fewer would leave CUs idle), else the
  // batch's molecules are too large for the array and the call is refused.
  if (pl.slot_floats > 0) {
    const size_t slot_bytes = pl.slot_floats * 4;
    const size_t fit = (NOALLOC_SCRATCH_BYTES - 256) / slot_bytes;
    const size_t need = (size_t)std::min(pl.grid, dev_info().n_cu);
    if (fit < need)
      return fail(MMFFHIP_ERR_SIZE, "mmffhip_minimize_nk_noalloc: " + std::to_string(need) + " slots of " +
                                        std::to_string(slot_bytes) + " bytes (one per compute unit) exceed its 64 MiB "
                                        "scratch; use mmffhip_minimize_nk with a workspace, or a smaller largest molecule");
    pl.grid = (int)std::min<size_t>((size_t)pl.grid, fit);
  }
  void* ws = nullptr;
  rc = scratch_ptr(&ws);
  if (rc) return rc;
  return run_nk(b, o, pl, x_in, x_out, e_out, status_out, counters, order, ws, (hipStream_t)stream);
}

int mmffhip_bb_fused(const mmffhip_batch* b, const float* x_in, float* x_out, double* e_out, int32_t max_iters,
                     float max_step, float grad_lr, void* stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (b->n_confs == 0) return MMFFHIP_OK;
  if (!x_in || !x_out || !e_out) return fail(MMFFHIP_ERR_ARG, "x_in, x_out and e_out are required");
  if (max_iters < 0) return fail(MMFFHIP_ERR_ARG, "max_iters must be >= 0");
  const size_t lds = 16 * 4 * 8 + (size_t)3 * b->max_atoms * (8 + 4 * 4) + 16;
  if ((int)lds > dev_info().lds) return fail(MMFFHIP_ERR_SIZE, "molecule too large for the LDS of one workgroup");
  KBatch kb = to_kbatch(b);
  allow_lds((const void*)&k_bb_fused<256>, lds);
  hipLaunchKernelGGL((k_bb_fused<256>), dim3(b->n_confs), dim3(256), lds, (hipStream_t)stream, kb, x_in, x_out,
                     e_out, (int)max_iters, max_step, grad_lr);
  HIP_TRY(hipGetLastError());
  return MMFFHIP_OK;
}

int mmffhip_multimol_workspace(const mmffhip_batch* b, int32_t lbfgs_m, size_t* bytes) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!bytes) return fail(MMFFHIP_ERR_ARG, "bytes is NULL");
  if (lbfgs_m < 1 || lbfgs_m > 64) return fail(MMFFHIP_ERR_ARG, "lbfgs_m must be in [1, 64]");
  // the total coordinate count is not known on the host without a device read; the
  // workspace is sized with n_confs * 3 * max_atoms as an upper bound
  const size_t T = (size_t)b->n_confs * 3 * b->max_atoms;
  const size_t C = (size_t)b->n_confs;
  auto r256 = [](size_t v) { return (v + 255) & ~(size_t)255; };
  size_t by = 0;
  by += 3 * r256(T * 4) + 2 * r256(T * 4 * (size_t)lbfgs_m);
  by += r256(C * 8 * (size_t)lbfgs_m) + 2 * r256(C * 8) + r256(C * 8) + 2 * r256(C * 4);
  by += r256(4096 * 4);
  *bytes = by + 1024;
  return MMFFHIP_OK;
}

int mmffhip_multimol_run(const mmffhip_batch* b, const float* x_in, float* x_best, double* e_final, float* grad_norm,
                         int32_t* per_conf_iters, int32_t* converged, int32_t* iters_used, int32_t max_iters,
                         int32_t eval_freq, float max_step_per_atom, int32_t lbfgs_m, float grad_scale,
                         float grad_tol, void* workspace, size_t workspace_bytes, void* stream) {
  return mmffhip_multimol_run_ex(b, x_in, x_best, e_final, grad_norm, per_conf_iters, converged, iters_used, max_iters,
                                 eval_freq, max_step_per_atom, lbfgs_m, grad_scale, grad_tol, 0, workspace,
                                 workspace_bytes, stream);
}

int mmffhip_multimol_run_ex(const mmffhip_batch* b, const float* x_in, float* x_best, double* e_final,
                            float* grad_norm, int32_t* per_conf_iters, int32_t* converged, int32_t* iters_used,
                            int32_t max_iters, int32_t eval_freq, float max_step_per_atom, int32_t lbfgs_m,
                            float grad_scale, float grad_tol, int32_t flags, void* workspace, size_t workspace_bytes,
                            void* stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (b->n_confs == 0) return MMFFHIP_OK;
  if (!x_in || !x_best || !e_final || !grad_norm || !per_conf_iters || !converged || !iters_used || !workspace)
    return fail(MMFFHIP_ERR_ARG, "NULL output or workspace");
  if (max_iters < 0 || eval_freq < 1) return fail(MMFFHIP_ERR_ARG, "max_iters >= 0 and eval_freq >= 1 required");
  size_t need = 0;
  rc = mmffhip_multimol_workspace(b, lbfgs_m, &need);
  if (rc) return rc;
  if (workspace_bytes < need) return fail(MMFFHIP_ERR_SIZE, "workspace too small");
  if (flags & ~(int32_t)MMFFHIP_MM_BATCH) return fail(MMFFHIP_ERR_ARG, "unknown multimol flags");
  const bool batch = (flags & MMFFHIP_MM_BATCH) != 0;
  const int n_main = (max_iters + eval_freq - 1) / eval_freq;
  const int polish = batch ? 0 : max_iters / 4;
  const int n_pol = (polish + eval_freq - 1) / eval_freq;
  if (n_main + n_pol > 4096) return fail(MMFFHIP_ERR_ARG, "too many sync points (max_iters / eval_freq > 4096)");
  const size_t T = (size_t)b->n_confs * 3 * b->max_atoms;
  const size_t C = (size_t)b->n_confs;
  char* p = (char*)workspace;
  auto take = [&](size_t bytes) {
    char* q = p;
    p += (bytes + 255) & ~(size_t)255;
    return (void*)q;
  };
  MMState S;
  S.total = T;
  S.x = (float*)take(T * 4);
  S.g = (float*)take(T * 4);
  S.bx = (float*)take(T * 4);
  S.S = (float*)take(T * 4 * lbfgs_m);
  S.Y = (float*)take(T * 4 * lbfgs_m);
  S.rho = (double*)take(C * 8 * lbfgs_m);
  S.best_e = (double*)take(C * 8);
  S.e_new = (double*)take(C * 8);
  S.ring = (int32_t*)take(C * 8);
  S.conv = (int32_t*)take(C * 4);
  S.pci = (int32_t*)take(C * 4);
  S.flags = (int32_t*)take(4096 * 4);
  if ((size_t)(p - (char*)workspace) > workspace_bytes) return fail(MMFFHIP_ERR_SIZE, "workspace too small");
  hipStream_t s = (hipStream_t)stream;
  KBatch kb = to_kbatch(b);
  constexpr int NT = 256;
  const int mx = 3 * b->max_atoms;
  const size_t lds_init = 16 * (NT / 64) * 8 + (size_t)mx * (8 + 4) + 16;
  const size_t lds_chunk = 16 * (NT / 64) * 8 + (size_t)mx * (8 + 16) + 8 * (size_t)lbfgs_m + 32;
  if ((int)lds_chunk > dev_info().lds) return fail(MMFFHIP_ERR_SIZE, "molecule too large for one workgroup");
  const int nfl = n_main + n_pol;
  allow_lds((const void*)&k_mm_init<NT>, lds_init);
  allow_lds((const void*)&k_mm_chunk<NT>, lds_chunk);
  allow_lds((const void*)&k_mm_final<NT>, lds_init);
  hipLaunchKernelGGL(k_fill_i32, dim3((4096 + 255) / 256), dim3(256), 0, s, S.flags, 4096, 0);
  hipLaunchKernelGGL((k_mm_init<NT>), dim3(b->n_confs), dim3(NT), lds_init, s, kb, S, x_in, grad_scale, 0,
                     (int)max_iters);
  HIP_TRY(hipGetLastError());
  int it = 0, k = 0;
  while (it < max_iters) {
    int steps = std::min(eval_freq, max_iters - it);
    it += steps;
    const bool sync = (it % eval_freq) == 0;
    // the batch form has no convergence exit: no conformer ever passes the test (tolerance
    // -1) and no launch reads the previous sync point's count
    hipLaunchKernelGGL((k_mm_chunk<NT>), dim3(b->n_confs), dim3(NT), lds_chunk, s, kb, S, steps, it, sync ? 1 : 0,
                       0, (k > 0 && !batch) ? S.flags + (k - 1) : (const int32_t*)nullptr, S.flags + k,
                       (int)lbfgs_m, max_step_per_atom, grad_scale, batch ? -1.0f : grad_tol);
    if (sync) k++;
  }
  HIP_TRY(hipGetLastError());
  if (!batch)
    hipLaunchKernelGGL((k_mm_init<NT>), dim3(b->n_confs), dim3(NT), lds_init, s, kb, S, x_in, grad_scale, 1,
                       (int)max_iters);
  it = 0;
  int kp = n_main;
  while (it < polish) {
    int steps = std::min(eval_freq, polish - it);
    it += steps;
    const bool sync = (it % eval_freq) == 0;
    hipLaunchKernelGGL((k_mm_chunk<NT>), dim3(b->n_confs), dim3(NT), lds_chunk, s, kb, S, steps, it, sync ? 1 : 0,
                       1, kp > n_main ? S.flags + (kp - 1) : (const int32_t*)nullptr, S.flags + kp, (int)lbfgs_m,
                       max_step_per_atom, grad_scale, grad_tol);
    if (sync) kp++;
  }
  HIP_TRY(hipGetLastError());
  (void)nfl;
  hipLaunchKernelGGL((k_mm_final<NT>), dim3(b->n_confs), dim3(NT), lds_init, s, kb, S, x_best, e_final, grad_norm,
                     per_conf_iters, converged, iters_used, batch ? 0 : (int)(max_iters / eval_freq),
                     (int)eval_freq, (int)max_iters, batch ? 1 : 0);
  HIP_TRY(hipGetLastError());
  return MMFFHIP_OK;
}

}  // extern "C"
