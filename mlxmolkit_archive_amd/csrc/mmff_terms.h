// MMFF94 term energies and gradients for gfx950, FP32 arithmetic.
//
// Two formula families, selected at compile time, because the reference ships two:
//   FAM_TG   -- the run_mmff / mmff_minimize_nk path, mlxmolkit/mmff_bfgs_header.metal:14-402
//               (theta in degrees, cubic bend cb = -0.4*pi/180, linear-angle branch,
//               Chebyshev torsion, distance-dependent electrostatics option)
//   FAM_MEGA -- the multi_mol / fused path, mlxmolkit/mmff_metal_kernel.py:274-505
//               (theta0 converted to radians, CB = -0.007, no linear branch, KEPS = 1e-12,
//               atan2 torsion, constant dielectric)
//
// Gradient contributions go to an accumulator object with add(atom, gx, gy, gz) and
// add_pair(a, b, g) (+g on a, -g on b). The kernels use LDS 64-bit fixed-point
// accumulators, so the sum over terms is exact and independent of the order in which
// lanes arrive (bit-reproducible runs).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FAM_TG 0
#define FAM_MEGA 1

namespace mmff {

constexpr float MDYNE = 143.9325f;
constexpr float PI_F = 3.14159265358979323846f;
constexpr float DEG2RAD = PI_F / 180.0f;
constexpr float RAD2DEG = 180.0f / PI_F;
// mega family constants, mmff_metal_kernel.py:275-284
constexpr float M_CS = -2.0f;
constexpr float M_CB = -0.007f;
constexpr float M_ABC = 0.043844f;
constexpr float M_SBC = 2.51210f;
constexpr float M_ELC = 332.0716f;
constexpr float M_ELD = 0.05f;
constexpr float M_R2D = 57.295779513082323f;
constexpr float M_KEPS = 1e-12f;

struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

__device__ __forceinline__ f3 ld(const float* x, int a) { return {x[3 * a], x[3 * a + 1], x[3 * a + 2]}; }

// 1/x with one v_rcp_f32 (within 1 ulp) for x positive and far from the denormal range, as
// every argument below is (a buffered or floored denominator); the generic division adds
// a range reduction around the same instruction.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// ----------------------------------------------------------------------------------
// bond stretch: TG mmff_bfgs_header.metal:15-46, MEGA mmff_metal_kernel.py:334-345
template <int FAM, bool G, class Acc>
__device__ __forceinline__ float t_bond(const float* x, Acc& acc, int i, int j, float kb, float r0) {
  f3 d = ld(x, i) - ld(x, j);
  float d2 = dot(d, d);
  if (FAM == FAM_TG) {
    float r = sqrtf(d2);
    float dr = r - r0;
    float e = (MDYNE * 0.5f) * kb * dr * dr * (1.0f - 2.0f * dr + (7.0f / 3.0f) * dr * dr);
    if (G) {
      float rg = sqrtf(fmaxf(d2, 1e-16f));
      float inv = rg > 1e-8f ? 1.0f / rg : 0.0f;
      float drg = rg - r0;
      float f = MDYNE * kb * drg * (1.0f - 3.0f * drg + (14.0f / 3.0f) * drg * drg) * inv;
      acc.add_pair(i, j, f * d.x, f * d.y, f * d.z);
    }
    return e;
  } else {
    float r = sqrtf(d2) + M_KEPS;
    float dr = r - r0;
    float e = MDYNE * kb * 0.5f * dr * dr * (1.0f + M_CS * dr + (7.0f / 12.0f) * M_CS * M_CS * dr * dr);
    if (G) {
      float dE = MDYNE * kb * dr * (1.0f + 1.5f * M_CS * dr + (7.0f / 6.0f) * M_CS * M_CS * dr * dr);
      float f = dE / r;
      acc.add_pair(i, j, f * d.x, f * d.y, f * d.z);
    }
    return e;
  }
}

// ----------------------------------------------------------------------------------
// angle bend: TG :49-125 (theta0 degrees, linear branch), MEGA :350-368 (theta0 degrees in
// the table, converted to radians here as the mega packer does, :283)
template <int FAM, bool G, class Acc>
__device__ __forceinline__ float t_angle(const float* x, Acc& acc, int i, int j, int k, float ka, float t0,
                                         bool lin) {
  f3 pj = ld(x, j);
  f3 r1 = ld(x, i) - pj, r2 = ld(x, k) - pj;
  if (FAM == FAM_TG) {
    float id1 = rsqrtf(fmaxf(dot(r1, r1), 1e-16f));
    float id2 = rsqrtf(fmaxf(dot(r2, r2), 1e-16f));
    float c = clampf(dot(r1, r2) * id1 * id2, -1.0f, 1.0f);
    float th = RAD2DEG * acosf(c);
    float dt = th - t0;
    const float cb = -0.4f * DEG2RAD;
    float e = lin ? MDYNE * ka * (1.0f + c) : 0.5f * MDYNE * DEG2RAD * DEG2RAD * ka * dt * dt * (1.0f + cb * dt);
    if (G) {
      float s2 = 1.0f - c * c;
      if (s2 >= 1e-16f) {
        float ins = -rsqrtf(fmaxf(s2, 1e-16f));
        float de = lin ? -MDYNE * ka * sqrtf(fmaxf(s2, 0.0f))
                       : (MDYNE * DEG2RAD) * ka * dt * (1.0f + (cb * 1.5f) * dt);
        float cf = de * ins;
        f3 h1 = id1 * r1, h2 = id2 * r2;
        f3 in1 = id1 * (h2 - c * h1), in3 = id2 * (h1 - c * h2);
        f3 g1 = cf * in1, g3 = cf * in3;
        acc.add(i, g1.x, g1.y, g1.z);
        acc.add(k, g3.x, g3.y, g3.z);
        acc.add(j, -(g1.x + g3.x), -(g1.y + g3.y), -(g1.z + g3.z));
      }
    }
    return e;
  } else {
    float t0r = t0 * 0.017453292519943295f;
    float rr1 = sqrtf(dot(r1, r1)) + M_KEPS, rr2 = sqrtf(dot(r2, r2)) + M_KEPS;
    float ct = clampf(dot(r1, r2) / (rr1 * rr2), -1.0f + 1e-7f, 1.0f - 1e-7f);
    float th = acosf(ct);
    float st = fmaxf(sinf(th), M_KEPS);
    float dt = (th - t0r) * M_R2D;
    float e = M_ABC * ka * 0.5f * dt * dt * (1.0f + M_CB * dt);
    if (G) {
      float dE = M_ABC * ka * dt * (1.0f + 1.5f * M_CB * dt) * M_R2D;
      f3 gi = (dE / (rr1 * st)) * ((ct / rr1) * r1 - (1.0f / rr2) * r2);
      f3 gk = (dE / (rr2 * st)) * ((ct / rr2) * r2 - (1.0f / rr1) * r1);
      acc.add(i, gi.x, gi.y, gi.z);
      acc.add(j, -gi.x - gk.x, -gi.y - gk.y, -gi.z - gk.z);
      acc.add(k, gk.x, gk.y, gk.z);
    }
    return e;
  }
}

// ----------------------------------------------------------------------------------
// stretch-bend: TG :128-167, MEGA :373-398
template <int FAM, bool G, class Acc>
__device__ __forceinline__ float t_strbend(const float* x, Acc& acc, int i, int j, int k, float kij, float kkj,
                                           float r0ij, float r0kj, float t0) {
  f3 pj = ld(x, j);
  f3 r1 = ld(x, i) - pj, r2 = ld(x, k) - pj;
  if (FAM == FAM_TG) {
    float id1 = rsqrtf(fmaxf(dot(r1, r1), 1e-16f));
    float id2 = rsqrtf(fmaxf(dot(r2, r2), 1e-16f));
    float c = clampf(dot(r1, r2) * id1 * id2, -1.0f, 1.0f);
    float th = RAD2DEG * acosf(c);
    float d1 = 1.0f / fmaxf(id1, 1e-8f), d2 = 1.0f / fmaxf(id2, 1e-8f);
    float dt = th - t0;
    float dr1 = d1 - r0ij, dr2 = d2 - r0kj;
    float e = 2.51210f * dt * (dr1 * kij + dr2 * kkj);
    if (G) {
      float s2 = 1.0f - c * c;
      if (s2 >= 1e-16f) {
        float ins = fminf(rsqrtf(fmaxf(s2, 1e-16f)), 1e8f);
        const float pf = MDYNE * DEG2RAD;
        float beis = RAD2DEG * (kij * dr1 + kkj * dr2) * ins;
        f3 h1 = id1 * r1, h2 = id2 * r2;
        f3 in1 = id1 * (h2 - c * h1), in3 = id2 * (h1 - c * h2);
        f3 g1 = pf * ((dt * kij) * h1 - beis * in1);
        f3 g3 = pf * ((dt * kkj) * h2 - beis * in3);
        f3 g2 = pf * ((-dt) * (kij * h1 + kkj * h2) + beis * (in1 + in3));
        acc.add(i, g1.x, g1.y, g1.z);
        acc.add(k, g3.x, g3.y, g3.z);
        acc.add(j, g2.x, g2.y, g2.z);
      }
    }
    return e;
  } else {
    float t0r = t0 * 0.017453292519943295f;
    float rij = sqrtf(dot(r1, r1)) + M_KEPS, rkj = sqrtf(dot(r2, r2)) + M_KEPS;
    float ct = clampf(dot(r1, r2) / (rij * rkj), -1.0f + 1e-7f, 1.0f - 1e-7f);
    float th = acosf(ct);
    float st = fmaxf(sinf(th), M_KEPS);
    float drij = rij - r0ij, drkj = rkj - r0kj;
    float dtd = (th - t0r) * M_R2D;
    float e = M_SBC * (kij * drij + kkj * drkj) * dtd;
    if (G) {
      float dEdth = M_SBC * (kij * drij + kkj * drkj) * M_R2D;
      float dEij = M_SBC * kij * dtd, dEkj = M_SBC * kkj * dtd;
      f3 dti = (1.0f / (rij * st)) * ((ct / rij) * r1 - (1.0f / rkj) * r2);
      f3 dtk = (1.0f / (rkj * st)) * ((ct / rkj) * r2 - (1.0f / rij) * r1);
      f3 gi = dEdth * dti + (dEij / rij) * r1;
      f3 gk = dEdth * dtk + (dEkj / rkj) * r2;
      acc.add(i, gi.x, gi.y, gi.z);
      acc.add(j, -gi.x - gk.x, -gi.y - gk.y, -gi.z - gk.z);
      acc.add(k, gk.x, gk.y, gk.z);
    }
    return e;
  }
}

// ----------------------------------------------------------------------------------
// out-of-plane: TG :170-249, MEGA :403-430. j is the centre, l the out-of-plane atom.
template <int FAM, bool G, class Acc>
__device__ __forceinline__ float t_oop(const float* x, Acc& acc, int i, int j, int k, int l, float koop) {
  f3 pj = ld(x, j);
  f3 u = ld(x, i) - pj, v = ld(x, k) - pj, w = ld(x, l) - pj;
  if (FAM == FAM_TG) {
    float iI = rsqrtf(fmaxf(dot(u, u), 1e-16f));
    float iK = rsqrtf(fmaxf(dot(v, v), 1e-16f));
    float iL = rsqrtf(fmaxf(dot(w, w), 1e-16f));
    f3 dI = iI * u, dK = iK * v, dL = iL * w;
    f3 n = cross(-dI, dK);
    float inl = rsqrtf(fmaxf(dot(n, n), 1e-16f));
    n = inl * n;
    float sc = clampf(dot(dL, n), -1.0f, 1.0f);
    float chi = RAD2DEG * asinf(sc);
    float e = 0.5f * MDYNE * DEG2RAD * DEG2RAD * koop * chi * chi;
    if (G) {
      float c2 = 1.0f - sc * sc;
      float icc = c2 > 0.0f ? rsqrtf(fmaxf(c2, 1e-16f)) : 1e8f;
      float ct = clampf(dot(dI, dK), -1.0f, 1.0f);
      float ist = rsqrtf(fmaxf(1.0f - ct * ct, 1e-8f));
      float dE = MDYNE * DEG2RAD * koop * chi;
      float tm1 = icc * ist, tm2 = sc * icc * ist * ist;
      f3 t1 = cross(dL, dK), t2 = cross(dI, dL), t3 = cross(dK, dI);
      f3 g1 = (dE * iI) * (tm1 * t1 - tm2 * (dI - ct * dK));
      f3 g3 = (dE * iK) * (tm1 * t2 - tm2 * (dK - ct * dI));
      f3 g4 = (dE * iL) * (tm1 * t3 - (sc * icc) * dL);
      acc.add(i, g1.x, g1.y, g1.z);
      acc.add(k, g3.x, g3.y, g3.z);
      acc.add(l, g4.x, g4.y, g4.z);
      acc.add(j, -(g1.x + g3.x + g4.x), -(g1.y + g3.y + g4.y), -(g1.z + g3.z + g4.z));
    }
    return e;
  } else {
    f3 nv = cross(u, v);
    float A = sqrtf(dot(nv, nv)) + M_KEPS, B = sqrtf(dot(w, w)) + M_KEPS;
    float AB = A * B;
    float sinc = clampf(dot(nv, w) / AB, -1.0f, 1.0f);
    float chi = asinf(sinc);
    float chid = chi * M_R2D;
    float e = M_ABC * koop * 0.5f * chid * chid;
    if (G) {
      float cosc = fmaxf(cosf(chi), M_KEPS);
      float dE = M_ABC * koop * chid * M_R2D / cosc;
      f3 nh = (1.0f / A) * nv, wh = (1.0f / B) * w;
      f3 vxw = cross(v, w), wxu = cross(w, u), vxn = cross(v, nh), nxu = cross(nh, u);
      f3 ds1 = (1.0f / AB) * vxw - (sinc / A) * vxn;
      f3 ds3 = (1.0f / AB) * wxu - (sinc / A) * nxu;
      f3 ds4 = (1.0f / B) * (nh - sinc * wh);
      f3 ds2 = -(ds1 + ds3 + ds4);
      acc.add(i, dE * ds1.x, dE * ds1.y, dE * ds1.z);
      acc.add(j, dE * ds2.x, dE * ds2.y, dE * ds2.z);
      acc.add(k, dE * ds3.x, dE * ds3.y, dE * ds3.z);
      acc.add(l, dE * ds4.x, dE * ds4.y, dE * ds4.z);
    }
    return e;
  }
}

// ----------------------------------------------------------------------------------
// torsion: TG :252-326 (Chebyshev in cos phi), MEGA :435-462 (atan2)
template <int FAM, bool G, class Acc>
__device__ __forceinline__ float t_torsion(const float* x, Acc& acc, int i1, int i2, int i3, int i4, float V1,
                                           float V2, float V3) {
  f3 p1 = ld(x, i1), p2 = ld(x, i2), p3 = ld(x, i3), p4 = ld(x, i4);
  if (FAM == FAM_TG) {
    f3 dx1 = p1 - p2, dx2 = p3 - p2, dx4 = p4 - p3;
    f3 c1 = cross(dx1, dx2), c2 = cross(-dx2, dx4);
    float n1 = dot(c1, c1), n2 = dot(c2, c2);
    float comb = n1 * n2;
    float e = 0.0f;
    if (comb >= 1e-16f) {
      float cp = clampf(dot(c1, c2) * rsqrtf(comb), -1.0f, 1.0f);
      float cos2 = 2.0f * cp * cp - 1.0f, cos3 = cp * (4.0f * cp * cp - 3.0f);
      e = 0.5f * (V1 * (1.0f + cp) + V2 * (1.0f - cos2) + V3 * (1.0f + cos3));
    }
    if (G) {
      if (n1 >= 1e-30f && n2 >= 1e-30f) {
        float in1 = fminf(rsqrtf(n1), 1e5f), in2 = fminf(rsqrtf(n2), 1e5f);
        f3 u1 = in1 * c1, u2 = in2 * c2;
        float cp = clampf(dot(u1, u2), -1.0f, 1.0f);
        float s2 = 1.0f - cp * cp;
        float st = 0.0f;
        if (s2 > 0.0f) st = 0.5f * (V1 - 2.0f * V2 * (2.0f * cp) + 3.0f * V3 * (3.0f - 4.0f * s2));
        f3 t0 = in1 * (u2 - cp * u1), t1 = in2 * (u1 - cp * u2);
        acc.add(i1, st * (t0.z * dx2.y - t0.y * dx2.z), st * (t0.x * dx2.z - t0.z * dx2.x),
                st * (t0.y * dx2.x - t0.x * dx2.y));
        acc.add(i4, st * (t1.z * dx2.y - t1.y * dx2.z), st * (t1.x * dx2.z - t1.z * dx2.x),
                st * (t1.y * dx2.x - t1.x * dx2.y));
        acc.add(i2,
                st * (t0.y * (dx2.z - dx1.z) + t0.z * (dx1.y - dx2.y) + t1.y * (-dx4.z) + t1.z * dx4.y),
                st * (t0.x * (dx1.z - dx2.z) + t0.z * (dx2.x - dx1.x) + t1.x * dx4.z + t1.z * (-dx4.x)),
                st * (t0.x * (dx2.y - dx1.y) + t0.y * (dx1.x - dx2.x) + t1.x * (-dx4.y) + t1.y * dx4.x));
        acc.add(i3,
                st * (t0.y * dx1.z + t0.z * (-dx1.y) + t1.y * (dx4.z + dx2.z) + t1.z * (-dx4.y - dx2.y)),
                st * (t0.x * (-dx1.z) + t0.z * dx1.x + t1.x * (-dx4.z - dx2.z) + t1.z * (dx4.x + dx2.x)),
                st * (t0.x * dx1.y + t0.y * (-dx1.x) + t1.x * (dx4.y + dx2.y) + t1.y * (-dx4.x - dx2.x)));
      }
    }
    return e;
  } else {
    f3 b1 = p2 - p1, b2 = p3 - p2, b3 = p4 - p3;
    f3 c1 = cross(b1, b2), c2 = cross(b2, b3);
    float b2n = sqrtf(dot(b2, b2)) + M_KEPS;
    f3 mv = cross(c1, (1.0f / b2n) * b2);
    float n1n2 = sqrtf(dot(c1, c1) * dot(c2, c2) + M_KEPS);
    float sw = dot(mv, c2) / n1n2, cw = dot(c1, c2) / n1n2;
    float om = atan2f(sw, cw);
    float e = 0.5f * (V1 * (1.0f + cosf(om)) + V2 * (1.0f - cosf(2.0f * om)) + V3 * (1.0f + cosf(3.0f * om)));
    if (G) {
      float dE = 0.5f * (-V1 * sinf(om) + 2.0f * V2 * sinf(2.0f * om) - 3.0f * V3 * sinf(3.0f * om));
      f3 rij = p1 - p2, rkj = p3 - p2, rkl = p3 - p4;
      f3 m = cross(rij, rkj), n = cross(rkj, rkl);
      float msq = dot(m, m) + M_KEPS, nsq = dot(n, n) + M_KEPS;
      float rkjn = sqrtf(dot(rkj, rkj)) + M_KEPS, rkjn2 = rkjn * rkjn;
      f3 fi = ((-rkjn / msq) * dE) * m;
      f3 fl = ((rkjn / nsq) * dE) * n;
      float pij = dot(rij, rkj) / rkjn2, pkl = dot(rkl, rkj) / rkjn2;
      f3 fj = -fi + pij * fi - pkl * fl;
      f3 fk = -fl - pij * fi + pkl * fl;
      acc.add(i1, fi.x, fi.y, fi.z);
      acc.add(i2, fj.x, fj.y, fj.z);
      acc.add(i3, fk.x, fk.y, fk.z);
      acc.add(i4, fl.x, fl.y, fl.z);
    }
    return e;
  }
}

// ----------------------------------------------------------------------------------
// non-bonded pair: buffered 14-7 vdW (TG :329-365, MEGA :467-485) fused with
// electrostatics (TG :368-402, MEGA :490-502). qqs = q_i q_j * (0.75 for 1-4 pairs, else 1);
// eps = 0 / qqs = 0 switch a half off. diel = 1 (constant) or 2 (distance dependent, TG only).
// Returns vdW energy in ev and ele energy in ee.
// iR = 1 / R* (1 / max(R*, 1e-8) for the TG family), precomputed in the pair table.
template <int FAM, bool G, class Acc>
__device__ __forceinline__ void t_pair(const float* x, Acc& acc, int i, int j, float iR, float eps, float qqs,
                                       int diel, float& ev, float& ee) {
  f3 d = ld(x, i) - ld(x, j);
  float d2 = dot(d, d);
  if (FAM == FAM_TG) {
    // four reciprocal-type operations per pair: 1/r (r = d2 / r), 1/(q + 0.07),
    // 1/(q^7 + 0.12) and 1/(r + 0.05), each shared by the energy and the gradient
    const float d2c = fmaxf(d2, 1e-16f);
    const float ir = rsqrtf(d2c);
    const float r = d2c * ir;
    const float q = r * iR;
    const float q2 = q * q, q6 = q2 * q2 * q2, q7 = q6 * q;
    const float rq = rcp(q + 0.07f);
    const float t = 1.07f * rq;
    const float t2 = t * t, t7 = t2 * t2 * t2 * t;
    const float r7 = rcp(q7 + 0.12f);  // q >= 0: q7 + 0.12 >= 0.12
    ev = eps * t7 * (1.12f * r7 - 2.0f);
    const float dpd = r + 0.05f;
    const float rd = rcp(dpd);  // dpd >= 0.05
    const float qe = 332.0716f * qqs;
    ee = diel == 2 ? qe * rd * rd : qe * rd;
    if (G) {
      const float inv = r > 1e-8f ? ir : 0.0f;
      const float dv = eps * iR * t7 * (-7.84f * q6 * r7 * r7 + (-7.84f * r7 + 14.0f) * rq);
      const float de = diel == 2 ? -2.0f * qe * rd * rd * rd : -qe * rd * rd;
      const float f = (dv + de) * inv;
      acc.add_pair(i, j, f * d.x, f * d.y, f * d.z);
    }
  } else {
    // shared reciprocals: 1/(rho + 0.07), 1/(rho^7 + 0.12), 1/(r + 0.05), 1/r
    const float r = sqrtf(d2) + M_KEPS;
    const float rho = r * iR;
    const float rho2 = rho * rho, rho6 = rho2 * rho2 * rho2, rho7 = rho6 * rho;
    const float rq = rcp(rho + 0.07f);
    const float r7 = rcp(rho7 + 0.12f);
    const float t1 = 1.07f * rq;
    const float t2 = 1.12f * r7 - 2.0f;
    const float t1s = t1 * t1, t16 = t1s * t1s * t1s, t17 = t16 * t1;
    ev = eps * t17 * t2;
    const float rd = rcp(r + M_ELD);
    ee = M_ELC * qqs * rd;
    if (G) {
      // d t1 / d rho = -1.07 rq^2 = -t1 rq; d t2 / d rho = -7.84 rho^6 r7^2
      const float dv = eps * iR * (-7.0f * t17 * rq * t2 - 7.84f * t17 * rho6 * r7 * r7);
      const float de = -M_ELC * qqs * rd * rd;
      const float f = (dv + de) * rcp(r);
      acc.add_pair(i, j, f * d.x, f * d.y, f * d.z);
    }
  }
}

}  // namespace mmff
