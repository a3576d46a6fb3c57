/*
 * FP64 CPU oracle for batched MMFF94 minimization -- TEST INFRASTRUCTURE ONLY.
 *
 * A double-precision restatement of the reference's algorithms for the batched
 * MMFF94 minimizer, used only to check the HIP product path (tests/, smoke(),
 * bench.py cpu_baseline). Each function cites the reference file:line it follows.
 *
 * Pinning status (which check fixes each function; see tests/test_oracle_*.py):
 *   term energies, both families ....... closed forms on hand-built 2-4 atom systems
 *   term gradients, both families ...... central finite differences of the energy,
 *                                        translation/rotation invariance
 *   family agreement ................... TG vs MEGA <= 1e-3 kcal/mol near equilibrium
 *                                        (restates tests/test_mmff_minimize_nk.py:25-40)
 *   BFGS / L-BFGS control flow ......... Rosenbrock-free: quadratic bowls with known
 *                                        minima, scipy BFGS agreement, max_iters=0 and
 *                                        grad_tol=0 contracts, monotone energy
 *   BB fused ........................... hand-unrolled first iterations on a 2-atom bond
 *   multi_mol lock-step L-BFGS ......... contract tests of
 *                                        tests/test_metal_multi_mol_convergence.py:39-93
 */
#include "mmff_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define MDYNE 143.9325
#define PI_D 3.14159265358979323846
#define DEG2RAD (PI_D / 180.0)
#define RAD2DEG (180.0 / PI_D)

/* constants of mlxmolkit/mmff_bfgs_header.metal:3-10 */
#define TOLX 1.2e-6
#define FUNCTOL 1e-4
#define MOVETOL 1e-6
#define EPS_GUARD 3e-7
#define MAX_STEP_FACTOR 100.0
#define MAX_LS_ITERS 1000
#define GRAD_SCALE_INIT 0.1
#define GRAD_CAP 10.0

/* constants of mlxmolkit/mmff_metal_kernel.py:275-284 (mega family) */
#define MEGA_CS (-2.0)
#define MEGA_CB (-0.007)
#define MEGA_ABC 0.043844
#define MEGA_SBC 2.51210
#define MEGA_ELC 332.0716
#define MEGA_ELD 0.05
#define MEGA_KEPS 1e-12

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double dclamp(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline void sub3(const double *x, int i, int j, double *o) {
    o[0] = x[3 * i] - x[3 * j];
    o[1] = x[3 * i + 1] - x[3 * j + 1];
    o[2] = x[3 * i + 2] - x[3 * j + 2];
}

/* ======================= TG family (mmff_bfgs_header.metal) ======================= */

/* bond stretch, mmff_bfgs_header.metal:15-46 */
static double tg_bond(const double *x, double *g, int i, int j, double kb, double r0) {
    double d[3];
    sub3(x, i, j, d);
    double d2 = dot3(d, d);
    double r = sqrt(d2);
    double dr = r - r0;
    double e = 0.5 * MDYNE * kb * dr * dr * (1.0 - 2.0 * dr + (7.0 / 3.0) * dr * dr);
    if (g) {
        double rg = sqrt(dmax(d2, 1e-16));
        double inv = rg > 1e-8 ? 1.0 / rg : 0.0;
        double drg = rg - r0;
        double de = MDYNE * kb * drg * (1.0 - 3.0 * drg + (14.0 / 3.0) * drg * drg);
        double f = de * inv;
        for (int c = 0; c < 3; c++) {
            g[3 * i + c] += f * d[c];
            g[3 * j + c] -= f * d[c];
        }
    }
    return e;
}

typedef struct {
    double cosT, theta_deg, inv_d1, inv_d2, r1[3], r2[3];
} tg_angle_data;

/* mmff_bfgs_header.metal:58-74 */
static void tg_angle_internals(const double *x, int i, int j, int k, tg_angle_data *a) {
    sub3(x, i, j, a->r1);
    sub3(x, k, j, a->r2);
    a->inv_d1 = 1.0 / sqrt(dmax(dot3(a->r1, a->r1), 1e-16));
    a->inv_d2 = 1.0 / sqrt(dmax(dot3(a->r2, a->r2), 1e-16));
    a->cosT = dclamp(dot3(a->r1, a->r2) * a->inv_d1 * a->inv_d2, -1.0, 1.0);
    a->theta_deg = RAD2DEG * acos(a->cosT);
}

/* angle bend, mmff_bfgs_header.metal:77-125 (linear branch :82-84, :103-105) */
static double tg_angle(const double *x, double *g, int i, int j, int k, double ka, double t0, int lin) {
    tg_angle_data a;
    tg_angle_internals(x, i, j, k, &a);
    const double cb = -0.4 * DEG2RAD;
    double dt = a.theta_deg - t0;
    double e = lin ? MDYNE * ka * (1.0 + a.cosT)
                   : 0.5 * MDYNE * DEG2RAD * DEG2RAD * ka * dt * dt * (1.0 + cb * dt);
    if (g) {
        double s2 = 1.0 - a.cosT * a.cosT;
        if (s2 < 1e-16) return e;
        double inv_neg_sin = -1.0 / sqrt(dmax(s2, 1e-16));
        double de = lin ? -MDYNE * ka * sqrt(dmax(s2, 0.0))
                        : MDYNE * DEG2RAD * ka * dt * (1.0 + 1.5 * cb * dt);
        double cf = de * inv_neg_sin;
        for (int c = 0; c < 3; c++) {
            double r1h = a.r1[c] * a.inv_d1, r2h = a.r2[c] * a.inv_d2;
            double in1 = a.inv_d1 * (r2h - a.cosT * r1h);
            double in3 = a.inv_d2 * (r1h - a.cosT * r2h);
            g[3 * i + c] += cf * in1;
            g[3 * k + c] += cf * in3;
            g[3 * j + c] -= cf * (in1 + in3);
        }
    }
    return e;
}

/* stretch-bend, mmff_bfgs_header.metal:128-167 */
static double tg_strbend(const double *x, double *g, int i, int j, int k, double r0ij, double r0kj,
                         double t0, double kij, double kkj) {
    tg_angle_data a;
    tg_angle_internals(x, i, j, k, &a);
    double d1 = 1.0 / dmax(a.inv_d1, 1e-8), d2 = 1.0 / dmax(a.inv_d2, 1e-8);
    double dt = a.theta_deg - t0;
    double drij = d1 - r0ij, drkj = d2 - r0kj;
    /* the reference writes the literal 2.51210 here but 143.9325*pi/180 in the gradient
     * (mmff_bfgs_header.metal:136 vs :154); both are kept */
    double e = 2.51210 * dt * (drij * kij + drkj * kkj);
    if (g) {
        double s2 = 1.0 - a.cosT * a.cosT;
        if (s2 < 1e-16) return e;
        double inv_sin = dmin(1.0 / sqrt(dmax(s2, 1e-16)), 1e8);
        double pf = MDYNE * DEG2RAD;
        double beis = RAD2DEG * (kij * drij + kkj * drkj) * inv_sin;
        for (int c = 0; c < 3; c++) {
            double r1h = a.r1[c] * a.inv_d1, r2h = a.r2[c] * a.inv_d2;
            double in1 = a.inv_d1 * (r2h - a.cosT * r1h);
            double in3 = a.inv_d2 * (r1h - a.cosT * r2h);
            g[3 * i + c] += pf * (dt * r1h * kij - in1 * beis);
            g[3 * k + c] += pf * (dt * r2h * kkj - in3 * beis);
            g[3 * j + c] += pf * (-dt * (r1h * kij + r2h * kkj) + (in1 + in3) * beis);
        }
    }
    return e;
}

/* out-of-plane (Wilson), mmff_bfgs_header.metal:170-249. Atom j is the centre,
 * l the atom whose angle to the (i,j,k) plane is chi. */
static double tg_oop(const double *x, double *g, int i, int j, int k, int l, double koop) {
    double rJI[3], rJK[3], rJL[3], dJI[3], dJK[3], dJL[3], n[3];
    sub3(x, i, j, rJI);
    sub3(x, k, j, rJK);
    sub3(x, l, j, rJL);
    double iJI = 1.0 / sqrt(dmax(dot3(rJI, rJI), 1e-16));
    double iJK = 1.0 / sqrt(dmax(dot3(rJK, rJK), 1e-16));
    double iJL = 1.0 / sqrt(dmax(dot3(rJL, rJL), 1e-16));
    for (int c = 0; c < 3; c++) {
        dJI[c] = rJI[c] * iJI;
        dJK[c] = rJK[c] * iJK;
        dJL[c] = rJL[c] * iJL;
    }
    double mI[3] = {-dJI[0], -dJI[1], -dJI[2]};
    cross3(mI, dJK, n);
    double inl = 1.0 / sqrt(dmax(dot3(n, n), 1e-16));
    for (int c = 0; c < 3; c++) n[c] *= inl;
    double sinChi = dclamp(dot3(dJL, n), -1.0, 1.0);
    double chi = RAD2DEG * asin(sinChi);
    double e = 0.5 * MDYNE * DEG2RAD * DEG2RAD * koop * chi * chi;
    if (g) {
        double cos2 = 1.0 - sinChi * sinChi;
        double invCos = cos2 > 0 ? 1.0 / sqrt(dmax(cos2, 1e-16)) : 1e8;
        double cosT = dclamp(dot3(dJI, dJK), -1.0, 1.0);
        double invSinT = 1.0 / sqrt(dmax(1.0 - cosT * cosT, 1e-8));
        double dE = MDYNE * DEG2RAD * koop * chi;
        double term1 = invCos * invSinT;
        double term2 = sinChi * invCos * invSinT * invSinT;
        double t1[3], t2[3], t3[3];
        cross3(dJL, dJK, t1);
        cross3(dJI, dJL, t2);
        cross3(dJK, dJI, t3);
        for (int c = 0; c < 3; c++) {
            double g1 = dE * (t1[c] * term1 - (dJI[c] - dJK[c] * cosT) * term2) * iJI;
            double g3 = dE * (t2[c] * term1 - (dJK[c] - dJI[c] * cosT) * term2) * iJK;
            double g4 = dE * (t3[c] * term1 - dJL[c] * sinChi * invCos) * iJL;
            g[3 * i + c] += g1;
            g[3 * k + c] += g3;
            g[3 * l + c] += g4;
            g[3 * j + c] -= (g1 + g3 + g4);
        }
    }
    return e;
}

/* torsion, mmff_bfgs_header.metal:252-326 (Chebyshev form in cos(phi)) */
static double tg_torsion(const double *x, double *g, int i1, int i2, int i3, int i4, double V1,
                         double V2, double V3) {
    double dx1[3], dx2[3], dx4[3], c1[3], c2[3], mdx2[3];
    sub3(x, i1, i2, dx1);
    sub3(x, i3, i2, dx2);
    sub3(x, i4, i3, dx4);
    cross3(dx1, dx2, c1);
    for (int c = 0; c < 3; c++) mdx2[c] = -dx2[c];
    cross3(mdx2, dx4, c2);
    double n1 = dot3(c1, c1), n2 = dot3(c2, c2);
    double e = 0.0;
    double comb = n1 * n2;
    if (comb >= 1e-16) {
        double cp = dclamp(dot3(c1, c2) / sqrt(comb), -1.0, 1.0);
        double cos2 = 2.0 * cp * cp - 1.0, cos3 = cp * (4.0 * cp * cp - 3.0);
        e = 0.5 * (V1 * (1.0 + cp) + V2 * (1.0 - cos2) + V3 * (1.0 + cos3));
    }
    if (g) {
        if (n1 < 1e-30 || n2 < 1e-30) return e;
        double in1 = dmin(1.0 / sqrt(n1), 1e5), in2 = dmin(1.0 / sqrt(n2), 1e5);
        double u1[3], u2[3];
        for (int c = 0; c < 3; c++) {
            u1[c] = c1[c] * in1;
            u2[c] = c2[c] * in2;
        }
        double cp = dclamp(dot3(u1, u2), -1.0, 1.0);
        double s2 = 1.0 - cp * cp;
        double st = 0.0;
        if (s2 > 0.0) st = 0.5 * (V1 - 2.0 * V2 * (2.0 * cp) + 3.0 * V3 * (3.0 - 4.0 * s2));
        double t0[3], t1[3];
        for (int c = 0; c < 3; c++) {
            t0[c] = in1 * (u2[c] - cp * u1[c]);
            t1[c] = in2 * (u1[c] - cp * u2[c]);
        }
        g[3 * i1 + 0] += st * (t0[2] * dx2[1] - t0[1] * dx2[2]);
        g[3 * i1 + 1] += st * (t0[0] * dx2[2] - t0[2] * dx2[0]);
        g[3 * i1 + 2] += st * (t0[1] * dx2[0] - t0[0] * dx2[1]);
        g[3 * i4 + 0] += st * (t1[1] * (-dx2[2]) - t1[2] * (-dx2[1]));
        g[3 * i4 + 1] += st * (t1[2] * (-dx2[0]) - t1[0] * (-dx2[2]));
        g[3 * i4 + 2] += st * (t1[0] * (-dx2[1]) - t1[1] * (-dx2[0]));
        g[3 * i2 + 0] += st * (t0[1] * (dx2[2] - dx1[2]) + t0[2] * (dx1[1] - dx2[1]) + t1[1] * (-dx4[2]) + t1[2] * dx4[1]);
        g[3 * i2 + 1] += st * (t0[0] * (dx1[2] - dx2[2]) + t0[2] * (dx2[0] - dx1[0]) + t1[0] * dx4[2] + t1[2] * (-dx4[0]));
        g[3 * i2 + 2] += st * (t0[0] * (dx2[1] - dx1[1]) + t0[1] * (dx1[0] - dx2[0]) + t1[0] * (-dx4[1]) + t1[1] * dx4[0]);
        g[3 * i3 + 0] += st * (t0[1] * dx1[2] + t0[2] * (-dx1[1]) + t1[1] * (dx4[2] + dx2[2]) + t1[2] * (-dx4[1] - dx2[1]));
        g[3 * i3 + 1] += st * (t0[0] * (-dx1[2]) + t0[2] * dx1[0] + t1[0] * (-dx4[2] - dx2[2]) + t1[2] * (dx4[0] + dx2[0]));
        g[3 * i3 + 2] += st * (t0[0] * dx1[1] + t0[1] * (-dx1[0]) + t1[0] * (dx4[1] + dx2[1]) + t1[1] * (-dx4[0] - dx2[0]));
    }
    return e;
}

/* buffered 14-7 vdW, mmff_bfgs_header.metal:329-365 */
static double tg_vdw(const double *x, double *g, int i, int j, double R, double eps) {
    double d[3];
    sub3(x, i, j, d);
    double d2 = dot3(d, d);
    double r = sqrt(dmax(d2, 1e-16));
    double iR = 1.0 / dmax(R, 1e-8);
    double q = r * iR;
    double q2 = q * q, q6 = q2 * q2 * q2, q7 = q6 * q;
    double t = 1.07 / (q + 0.07);
    double t2 = t * t, t7 = t2 * t2 * t2 * t;
    double e = eps * t7 * (1.12 / (q7 + 0.12) - 2.0);
    if (g) {
        double inv = r > 1e-8 ? 1.0 / r : 0.0;
        double iq7p = 1.0 / dmax(q7 + 0.12, 1e-30);
        double dEdr = eps * iR * t7 * (-7.84 * q6 * iq7p * iq7p + (-7.84 * iq7p + 14.0) / (q + 0.07));
        double f = dEdr * inv;
        for (int c = 0; c < 3; c++) {
            g[3 * i + c] += f * d[c];
            g[3 * j + c] -= f * d[c];
        }
    }
    return e;
}

/* electrostatics, mmff_bfgs_header.metal:368-402; diel n = 1 or 2 */
static double tg_ele(const double *x, double *g, int i, int j, double qq, int diel, int is14) {
    double d[3];
    sub3(x, i, j, d);
    double r = sqrt(dmax(dot3(d, d), 1e-16));
    double dpd = r + 0.05;
    double sc = is14 ? 0.75 : 1.0;
    double e = 332.0716 * qq * sc / (diel == 2 ? dpd * dpd : dpd);
    if (g) {
        double inv = r > 1e-8 ? 1.0 / r : 0.0;
        double dn1 = diel == 2 ? dpd * dpd * dpd : dpd * dpd;
        double dE = -332.0716 * (double)diel * qq * sc / dmax(dn1, 1e-30);
        double f = dE * inv;
        for (int c = 0; c < 3; c++) {
            g[3 * i + c] += f * d[c];
            g[3 * j + c] -= f * d[c];
        }
    }
    return e;
}

/* ======================= MEGA family (mmff_metal_kernel.py:274-505) ======================= */

static inline double len3(const double *a) { return sqrt(dot3(a, a)); }

/* :334-345 */
static double mg_bond(const double *x, double *g, int i, int j, double kb, double r0) {
    double d[3];
    sub3(x, i, j, d);
    double r = len3(d) + MEGA_KEPS;
    double dr = r - r0;
    double e = MDYNE * kb * 0.5 * dr * dr * (1.0 + MEGA_CS * dr + (7.0 / 12.0) * MEGA_CS * MEGA_CS * dr * dr);
    if (g) {
        double dE = MDYNE * kb * dr * (1.0 + 1.5 * MEGA_CS * dr + (7.0 / 6.0) * MEGA_CS * MEGA_CS * dr * dr);
        for (int c = 0; c < 3; c++) {
            g[3 * i + c] += dE / r * d[c];
            g[3 * j + c] -= dE / r * d[c];
        }
    }
    return e;
}

/* :350-368 ; t0 in radians */
static double mg_angle(const double *x, double *g, int i, int j, int k, double ka, double t0) {
    double v1[3], v2[3];
    sub3(x, i, j, v1);
    sub3(x, k, j, v2);
    double r1 = len3(v1) + MEGA_KEPS, r2 = len3(v2) + MEGA_KEPS;
    double ct = dclamp(dot3(v1, v2) / (r1 * r2), -1.0 + 1e-7, 1.0 - 1e-7);
    double th = acos(ct);
    double st = dmax(sin(th), MEGA_KEPS);
    double dt = (th - t0) * RAD2DEG;
    double e = MEGA_ABC * ka * 0.5 * dt * dt * (1.0 + MEGA_CB * dt);
    if (g) {
        double dE = MEGA_ABC * ka * dt * (1.0 + 1.5 * MEGA_CB * dt) * RAD2DEG;
        for (int c = 0; c < 3; c++) {
            double gi = dE * (ct * v1[c] / r1 - v2[c] / r2) / (r1 * st);
            double gk = dE * (ct * v2[c] / r2 - v1[c] / r1) / (r2 * st);
            g[3 * i + c] += gi;
            g[3 * j + c] += -gi - gk;
            g[3 * k + c] += gk;
        }
    }
    return e;
}

/* :373-398 ; t0 in radians */
static double mg_strbend(const double *x, double *g, int i, int j, int k, double kij, double kji,
                         double r0ij, double r0kj, double t0) {
    double vij[3], vkj[3];
    sub3(x, i, j, vij);
    sub3(x, k, j, vkj);
    double rij = len3(vij) + MEGA_KEPS, rkj = len3(vkj) + MEGA_KEPS;
    double ct = dclamp(dot3(vij, vkj) / (rij * rkj), -1.0 + 1e-7, 1.0 - 1e-7);
    double th = acos(ct);
    double st = dmax(sin(th), MEGA_KEPS);
    double drij = rij - r0ij, drkj = rkj - r0kj;
    double dtd = (th - t0) * RAD2DEG;
    double e = MEGA_SBC * (kij * drij + kji * drkj) * dtd;
    if (g) {
        double dEdth = MEGA_SBC * (kij * drij + kji * drkj) * RAD2DEG;
        double dEij = MEGA_SBC * kij * dtd, dEkj = MEGA_SBC * kji * dtd;
        for (int c = 0; c < 3; c++) {
            double dti = (ct * vij[c] / rij - vkj[c] / rkj) / (rij * st);
            double dtk = (ct * vkj[c] / rkj - vij[c] / rij) / (rkj * st);
            double gi = dEdth * dti + dEij * vij[c] / rij;
            double gk = dEdth * dtk + dEkj * vkj[c] / rkj;
            g[3 * i + c] += gi;
            g[3 * j + c] += -gi - gk;
            g[3 * k + c] += gk;
        }
    }
    return e;
}

/* :403-430 */
static double mg_oop(const double *x, double *g, int a1, int a2, int a3, int a4, double koop) {
    double u[3], v[3], w[3], nv[3];
    sub3(x, a1, a2, u);
    sub3(x, a3, a2, v);
    sub3(x, a4, a2, w);
    cross3(u, v, nv);
    double A = len3(nv) + MEGA_KEPS, B = len3(w) + MEGA_KEPS;
    double AB = A * B;
    double sinc = dclamp(dot3(nv, w) / AB, -1.0, 1.0);
    double chi = asin(sinc);
    double chid = chi * RAD2DEG;
    double cosc = dmax(cos(chi), MEGA_KEPS);
    double e = MEGA_ABC * koop * 0.5 * chid * chid;
    if (g) {
        double dE = MEGA_ABC * koop * chid * RAD2DEG / cosc;
        double nh[3], wh[3], vxw[3], wxu[3], vxn[3], nxu[3];
        for (int c = 0; c < 3; c++) {
            nh[c] = nv[c] / A;
            wh[c] = w[c] / B;
        }
        cross3(v, w, vxw);
        cross3(w, u, wxu);
        cross3(v, nh, vxn);
        cross3(nh, u, nxu);
        for (int c = 0; c < 3; c++) {
            double ds1 = vxw[c] / AB - sinc * vxn[c] / A;
            double ds3 = wxu[c] / AB - sinc * nxu[c] / A;
            double ds4 = (nh[c] - sinc * wh[c]) / B;
            double ds2 = -(ds1 + ds3 + ds4);
            g[3 * a1 + c] += dE * ds1;
            g[3 * a2 + c] += dE * ds2;
            g[3 * a3 + c] += dE * ds3;
            g[3 * a4 + c] += dE * ds4;
        }
    }
    return e;
}

/* :435-462 */
static double mg_torsion(const double *x, double *g, int a1, int a2, int a3, int a4, double V1,
                         double V2, double V3) {
    double b1[3], b2[3], b3[3], c1[3], c2[3], b2h[3], mv[3];
    sub3(x, a2, a1, b1);
    sub3(x, a3, a2, b2);
    sub3(x, a4, a3, b3);
    cross3(b1, b2, c1);
    cross3(b2, b3, c2);
    double b2n = len3(b2) + MEGA_KEPS;
    for (int c = 0; c < 3; c++) b2h[c] = b2[c] / b2n;
    cross3(c1, b2h, mv);
    double n1n2 = sqrt(dot3(c1, c1) * dot3(c2, c2) + MEGA_KEPS);
    double sw = dot3(mv, c2) / n1n2, cw = dot3(c1, c2) / n1n2;
    double om = atan2(sw, cw);
    double e = 0.5 * (V1 * (1.0 + cos(om)) + V2 * (1.0 - cos(2.0 * om)) + V3 * (1.0 + cos(3.0 * om)));
    if (g) {
        double dE = 0.5 * (-V1 * sin(om) + 2.0 * V2 * sin(2.0 * om) - 3.0 * V3 * sin(3.0 * om));
        double rij[3], rkj[3], rkl[3], m[3], n[3];
        sub3(x, a1, a2, rij);
        sub3(x, a3, a2, rkj);
        sub3(x, a3, a4, rkl);
        cross3(rij, rkj, m);
        cross3(rkj, rkl, n);
        double msq = dot3(m, m) + MEGA_KEPS, nsq = dot3(n, n) + MEGA_KEPS;
        double rkjn = len3(rkj) + MEGA_KEPS, rkjn2 = rkjn * rkjn;
        double pij = dot3(rij, rkj) / rkjn2, pkl = dot3(rkl, rkj) / rkjn2;
        for (int c = 0; c < 3; c++) {
            double fi = (-rkjn / msq) * m[c] * dE;
            double fl = (rkjn / nsq) * n[c] * dE;
            double fj = -fi + pij * fi - pkl * fl;
            double fk = -fl - pij * fi + pkl * fl;
            g[3 * a1 + c] += fi;
            g[3 * a2 + c] += fj;
            g[3 * a3 + c] += fk;
            g[3 * a4 + c] += fl;
        }
    }
    return e;
}

/* :467-485 */
static double mg_vdw(const double *x, double *g, int i, int j, double R, double eps) {
    double d[3];
    sub3(x, i, j, d);
    double r = len3(d) + MEGA_KEPS;
    double rho = r / R;
    double rho7 = rho * rho * rho * rho * rho * rho * rho;
    double t1 = 1.07 / (rho + 0.07);
    double t2 = 1.12 / (rho7 + 0.12) - 2.0;
    double t17 = t1 * t1 * t1 * t1 * t1 * t1 * t1;
    double e = eps * t17 * t2;
    if (g) {
        double dt1 = -1.07 / ((rho + 0.07) * (rho + 0.07));
        double dt2 = -1.12 * 7.0 * rho * rho * rho * rho * rho * rho / ((rho7 + 0.12) * (rho7 + 0.12));
        double dE = eps * (7.0 * t17 / t1 * dt1 * t2 + t17 * dt2) / R;
        for (int c = 0; c < 3; c++) {
            g[3 * i + c] += dE * d[c] / r;
            g[3 * j + c] -= dE * d[c] / r;
        }
    }
    return e;
}

/* :490-502 ; sc = 0.75 (1-4) or 1 (mmff_metal_kernel.py:651-652) */
static double mg_ele(const double *x, double *g, int i, int j, double qq, double sc) {
    double d[3];
    sub3(x, i, j, d);
    double r = len3(d) + MEGA_KEPS;
    double den = r + MEGA_ELD;
    double e = sc * MEGA_ELC * qq / den;
    if (g) {
        double dE = -sc * MEGA_ELC * qq / (den * den);
        for (int c = 0; c < 3; c++) {
            g[3 * i + c] += dE * d[c] / r;
            g[3 * j + c] -= dE * d[c] / r;
        }
    }
    return e;
}

/* FP32-storage emulation, an analysis switch that is off by default and used only to study
 * how the HIP kernels' number formats move trajectories: when on, each TG-family term energy
 * is rounded to float before the FP64 sum, the summed gradient is rounded to float, and the nk
 * minimizers round coordinates to float at the start and after every step (the kernels keep
 * coordinates and term results in FP32 and sum energies in FP64). Set it before a run; every
 * thread of a batch reads the same setting. The dense inverse Hessian, its products and the
 * L-BFGS vector updates are then rounded to float too (their dot products stay FP64). */
static int g_fp32_emulation = 0;
void oracle_set_fp32_emulation(int on) { g_fp32_emulation = on != 0; }
static inline double rf(double v) { return g_fp32_emulation ? (double)(float)v : v; }

double oracle_energy_grad(const oracle_mol *m, int family, int diel, const double *x, double *g,
                          uint32_t mask, double *et) {
    double e[7] = {0, 0, 0, 0, 0, 0, 0};
    int n3 = 3 * m->n_atoms;
    if (g) memset(g, 0, sizeof(double) * (size_t)n3);
    if (family == ORACLE_FAMILY_TG) {
        if (mask & T_BOND)
            for (int t = 0; t < m->nb; t++) e[0] += rf(tg_bond(x, g, m->b_i[t], m->b_j[t], m->b_kb[t], m->b_r0[t]));
        if (mask & T_ANGLE)
            for (int t = 0; t < m->na; t++)
                e[1] += rf(tg_angle(x, g, m->a_i[t], m->a_j[t], m->a_k[t], m->a_ka[t], m->a_t0[t], m->a_lin[t] != 0));
        if (mask & T_STRBEND)
            for (int t = 0; t < m->nsb; t++)
                e[2] += rf(tg_strbend(x, g, m->s_i[t], m->s_j[t], m->s_k[t], m->s_r0ij[t], m->s_r0kj[t], m->s_t0[t],
                                      m->s_kij[t], m->s_kkj[t]));
        if (mask & T_OOP)
            for (int t = 0; t < m->noop; t++)
                e[3] += rf(tg_oop(x, g, m->o_i[t], m->o_j[t], m->o_k[t], m->o_l[t], m->o_koop[t]));
        if (mask & T_TORSION)
            for (int t = 0; t < m->nt; t++)
                e[4] += rf(tg_torsion(x, g, m->t_i[t], m->t_j[t], m->t_k[t], m->t_l[t], m->t_v1[t], m->t_v2[t],
                                      m->t_v3[t]));
        if (mask & T_VDW)
            for (int t = 0; t < m->nv; t++) e[5] += rf(tg_vdw(x, g, m->v_i[t], m->v_j[t], m->v_r[t], m->v_eps[t]));
        if (mask & T_ELE)
            for (int t = 0; t < m->ne; t++)
                e[6] += rf(tg_ele(x, g, m->e_i[t], m->e_j[t], m->e_qq[t], diel, m->e_14[t] != 0));
        if (g && g_fp32_emulation)
            for (int i = 0; i < n3; i++) g[i] = rf(g[i]);
    } else {
        /* theta0 is stored in degrees; the mega packer converts to radians
         * (mmff_metal_kernel.py:283-286). */
        if (mask & T_BOND)
            for (int t = 0; t < m->nb; t++) e[0] += mg_bond(x, g, m->b_i[t], m->b_j[t], m->b_kb[t], m->b_r0[t]);
        if (mask & T_ANGLE)
            for (int t = 0; t < m->na; t++)
                e[1] += mg_angle(x, g, m->a_i[t], m->a_j[t], m->a_k[t], m->a_ka[t], (double)m->a_t0[t] * DEG2RAD);
        if (mask & T_STRBEND)
            for (int t = 0; t < m->nsb; t++)
                e[2] += mg_strbend(x, g, m->s_i[t], m->s_j[t], m->s_k[t], m->s_kij[t], m->s_kkj[t], m->s_r0ij[t],
                                   m->s_r0kj[t], (double)m->s_t0[t] * DEG2RAD);
        if (mask & T_OOP)
            for (int t = 0; t < m->noop; t++) e[3] += mg_oop(x, g, m->o_i[t], m->o_j[t], m->o_k[t], m->o_l[t], m->o_koop[t]);
        if (mask & T_TORSION)
            for (int t = 0; t < m->nt; t++)
                e[4] += mg_torsion(x, g, m->t_i[t], m->t_j[t], m->t_k[t], m->t_l[t], m->t_v1[t], m->t_v2[t], m->t_v3[t]);
        if (mask & T_VDW)
            for (int t = 0; t < m->nv; t++) e[5] += mg_vdw(x, g, m->v_i[t], m->v_j[t], m->v_r[t], m->v_eps[t]);
        if (mask & T_ELE)
            for (int t = 0; t < m->ne; t++) e[6] += mg_ele(x, g, m->e_i[t], m->e_j[t], m->e_qq[t], m->e_14[t] ? 0.75 : 1.0);
    }
    double tot = 0.0;
    for (int k = 0; k < 7; k++) {
        tot += e[k];
        if (et) et[k] = e[k];
    }
    return tot;
}

/* Sum over terms of |term energy|, per term type: the natural scale of FP32 rounding
 * error in a term-type total (used to set relative tolerances). */
void oracle_energy_terms_abs(const oracle_mol *m, int family, int diel, const double *x, uint32_t mask,
                             double *et) {
    double *g = NULL;
    double e[7] = {0, 0, 0, 0, 0, 0, 0};
    if (family == ORACLE_FAMILY_TG) {
        if (mask & T_BOND)
            for (int t = 0; t < m->nb; t++) e[0] += fabs(tg_bond(x, g, m->b_i[t], m->b_j[t], m->b_kb[t], m->b_r0[t]));
        if (mask & T_ANGLE)
            for (int t = 0; t < m->na; t++)
                e[1] += fabs(tg_angle(x, g, m->a_i[t], m->a_j[t], m->a_k[t], m->a_ka[t], m->a_t0[t], m->a_lin[t] != 0));
        if (mask & T_STRBEND)
            for (int t = 0; t < m->nsb; t++)
                e[2] += fabs(tg_strbend(x, g, m->s_i[t], m->s_j[t], m->s_k[t], m->s_r0ij[t], m->s_r0kj[t], m->s_t0[t],
                                   m->s_kij[t], m->s_kkj[t]));
        if (mask & T_OOP)
            for (int t = 0; t < m->noop; t++) e[3] += fabs(tg_oop(x, g, m->o_i[t], m->o_j[t], m->o_k[t], m->o_l[t], m->o_koop[t]));
        if (mask & T_TORSION)
            for (int t = 0; t < m->nt; t++)
                e[4] += fabs(tg_torsion(x, g, m->t_i[t], m->t_j[t], m->t_k[t], m->t_l[t], m->t_v1[t], m->t_v2[t], m->t_v3[t]));
        if (mask & T_VDW)
            for (int t = 0; t < m->nv; t++) e[5] += fabs(tg_vdw(x, g, m->v_i[t], m->v_j[t], m->v_r[t], m->v_eps[t]));
        if (mask & T_ELE)
            for (int t = 0; t < m->ne; t++) e[6] += fabs(tg_ele(x, g, m->e_i[t], m->e_j[t], m->e_qq[t], diel, m->e_14[t] != 0));
    } else {
        /* theta0 is stored in degrees; the mega packer converts to radians
         * (mmff_metal_kernel.py:283-286). */
        if (mask & T_BOND)
            for (int t = 0; t < m->nb; t++) e[0] += fabs(mg_bond(x, g, m->b_i[t], m->b_j[t], m->b_kb[t], m->b_r0[t]));
        if (mask & T_ANGLE)
            for (int t = 0; t < m->na; t++)
                e[1] += fabs(mg_angle(x, g, m->a_i[t], m->a_j[t], m->a_k[t], m->a_ka[t], (double)m->a_t0[t] * DEG2RAD));
        if (mask & T_STRBEND)
            for (int t = 0; t < m->nsb; t++)
                e[2] += fabs(mg_strbend(x, g, m->s_i[t], m->s_j[t], m->s_k[t], m->s_kij[t], m->s_kkj[t], m->s_r0ij[t],
                                   m->s_r0kj[t], (double)m->s_t0[t] * DEG2RAD));
        if (mask & T_OOP)
            for (int t = 0; t < m->noop; t++) e[3] += fabs(mg_oop(x, g, m->o_i[t], m->o_j[t], m->o_k[t], m->o_l[t], m->o_koop[t]));
        if (mask & T_TORSION)
            for (int t = 0; t < m->nt; t++)
                e[4] += fabs(mg_torsion(x, g, m->t_i[t], m->t_j[t], m->t_k[t], m->t_l[t], m->t_v1[t], m->t_v2[t], m->t_v3[t]));
        if (mask & T_VDW)
            for (int t = 0; t < m->nv; t++) e[5] += fabs(mg_vdw(x, g, m->v_i[t], m->v_j[t], m->v_r[t], m->v_eps[t]));
        if (mask & T_ELE)
            for (int t = 0; t < m->ne; t++) e[6] += fabs(mg_ele(x, g, m->e_i[t], m->e_j[t], m->e_qq[t], m->e_14[t] ? 0.75 : 1.0));
    }
    for (int k = 0; k < 7; k++) et[k] = e[k];
}


/* ======================= nk optimizers (TG family) ======================= */

/* scale_grad_serial, mmff_bfgs_header.metal:435-454: g *= scale, halve while max|g| > 10.
 * grad_scale is monotone non-increasing across calls. */
static void scale_grad(double *g, int n, double *gs) {
    double s = *gs, mx = 0.0;
    for (int i = 0; i < n; i++) {
        g[i] *= s;
        mx = dmax(mx, fabs(g[i]));
    }
    while (mx > GRAD_CAP) {
        s *= 0.5;
        mx = 0.0;
        for (int i = 0; i < n; i++) {
            g[i] *= 0.5;
            mx = dmax(mx, fabs(g[i]));
        }
    }
    *gs = s;
}

/* Backtracking line search shared by BFGS and L-BFGS:
 * mmff_bfgs_source_tg.metal:153-256. x_old holds the start point on entry.
 * Returns the accepted energy (or old_e when the search gives up, in which case x = x_old). */
static double line_search(const oracle_mol *m, int diel, const double *x_old, double *x, const double *d,
                          double slope, double lam_min, double old_e, int *n_evals, int *accepted) {
    int n = 3 * m->n_atoms;
    double lam = 1.0, prev_lam = 1.0, prev_e = old_e;
    *accepted = 0;
    for (int ls = 0; ls < MAX_LS_ITERS; ls++) {
        if (lam < lam_min) {
            memcpy(x, x_old, sizeof(double) * (size_t)n);
            return old_e;
        }
        for (int i = 0; i < n; i++) x[i] = rf(x_old[i] + lam * d[i]);
        double te = oracle_energy_grad(m, ORACLE_FAMILY_TG, diel, x, NULL, T_ALL, NULL);
        (*n_evals)++;
        if (te - old_e <= FUNCTOL * lam * slope) {
            *accepted = 1;
            return te;
        }
        double tmp;
        if (ls == 0) {
            tmp = -slope / (2.0 * (te - old_e - slope));
        } else {
            double rhs1 = te - old_e - lam * slope;
            double rhs2 = prev_e - old_e - prev_lam * slope;
            double l2 = lam * lam, pl2 = prev_lam * prev_lam, dv = lam - prev_lam;
            if (fabs(dv) < 1e-30) {
                tmp = 0.5 * lam;
            } else {
                double a = (rhs1 / l2 - rhs2 / pl2) / dv;
                double b = (-prev_lam * rhs1 / l2 + lam * rhs2 / pl2) / dv;
                if (fabs(a) < 1e-30) {
                    tmp = fabs(b) > 1e-30 ? -slope / (2.0 * b) : 0.5 * lam;
                } else {
                    double disc = b * b - 3.0 * a * slope;
                    if (disc < 0.0) tmp = 0.5 * lam;
                    else if (b <= 0.0) tmp = (-b + sqrt(disc)) / (3.0 * a);
                    else tmp = -slope / (b + sqrt(disc));
                }
            }
        }
        tmp = dmin(tmp, 0.5 * lam);
        tmp = dmax(tmp, 0.1 * lam);
        prev_lam = lam;
        prev_e = te;
        lam = tmp;
    }
    memcpy(x, x_old, sizeof(double) * (size_t)n);
    return old_e;
}

/* Common prologue/iteration body of mmff_bfgs_source_tg.metal and
 * mmff_lbfgs_source_tg.metal. hist == NULL selects dense BFGS. */
typedef struct {
    int m, count, idx;
    double *S, *Y, *rho, *alpha;
} lbfgs_hist;

/* Dense inverse Hessian start: the identity, or (h0 != NULL) the block-diagonal matrix of the
 * symmetric parts (B + B^T) / 2 of the per-atom 3x3 blocks h0, scaled by GRAD_SCALE_INIT / gs --
 * the learned preconditioner of the HIP minimizer, which has no counterpart in the reference. */
static void init_hessian(double *H, int n, const double *h0, double psc) {
    memset(H, 0, sizeof(double) * (size_t)n * n);
    if (!h0) {
        for (int i = 0; i < n; i++) H[(size_t)i * n + i] = 1.0;
        return;
    }
    for (int a = 0; a < n / 3; a++)
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++)
                H[(size_t)(3 * a + r) * n + 3 * a + c] = psc * 0.5 * (h0[9 * a + 3 * r + c] + h0[9 * a + 3 * c + r]);
}

static int nk_minimize(const oracle_mol *mol, double *x, int max_iters, double grad_tol, int diel,
                       lbfgs_hist *hist, const double *h0, double *e_out, int32_t *counters) {
    int n = 3 * mol->n_atoms;
    size_t nb = sizeof(double) * (size_t)n;
    double *g = malloc(nb), *d = malloc(nb), *xo = malloc(nb), *dg = malloc(nb), *hdg = malloc(nb);
    double *H = NULL;
    if (hist) h0 = NULL;
    int n_evals = 0, n_grads = 1, iters = 0;
    for (int i = 0; i < n; i++) x[i] = rf(x[i]);
    /* initial energy + gradient, :115-135 */
    double energy = oracle_energy_grad(mol, ORACLE_FAMILY_TG, diel, x, g, T_ALL, NULL);
    double gs = GRAD_SCALE_INIT;
    scale_grad(g, n, &gs);
    if (!hist) {
        H = malloc(sizeof(double) * (size_t)n * n);
        init_hessian(H, n, h0, GRAD_SCALE_INIT / gs);
    }
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        if (h0)
            for (int j = 0; j < n; j++) s += H[(size_t)i * n + j] * g[j];
        else
            s = g[i];
        d[i] = -s;
    }
    double ss = 0.0;
    for (int i = 0; i < n; i++) ss += x[i] * x[i];
    double max_step = MAX_STEP_FACTOR * dmax(sqrt(ss), (double)n); /* :141-145 */
    int status = 1;
    for (int it = 0; it < max_iters; it++) {
        iters = it + 1;
        memcpy(xo, x, nb);
        double old_e = energy;
        double dn = 0.0;
        for (int i = 0; i < n; i++) dn += d[i] * d[i];
        dn = sqrt(dn);
        if (dn > max_step) {
            double sc = max_step / dn;
            for (int i = 0; i < n; i++) d[i] *= sc;
        }
        double slope = 0.0;
        for (int i = 0; i < n; i++) slope += d[i] * g[i];
        if (slope >= 0.0) { /* :170-174 ascent bail-out */
            status = 0;
            break;
        }
        double test = 0.0;
        for (int i = 0; i < n; i++) test = dmax(test, fabs(d[i]) / dmax(fabs(x[i]), 1.0));
        double lam_min = MOVETOL / dmax(test, 1e-30);
        int accepted;
        energy = line_search(mol, diel, xo, x, d, slope, lam_min, old_e, &n_evals, &accepted);
        if (!accepted && h0) {
            /* learned start failed: restart from the identity (x is back at x_old) */
            h0 = NULL;
            init_hessian(H, n, NULL, 1.0);
            for (int i = 0; i < n; i++) d[i] = -g[i];
            continue;
        }
        /* xi = x - x_old (stored in xo), TOLX :258-279 */
        double tx = 0.0;
        for (int i = 0; i < n; i++) {
            xo[i] = x[i] - xo[i];
            tx = dmax(tx, fabs(xo[i]) / dmax(fabs(x[i]), 1.0));
        }
        if (tx < TOLX) {
            status = 0;
            break;
        }
        /* new gradient, :285-298 */
        memcpy(dg, g, nb);
        oracle_energy_grad(mol, ORACLE_FAMILY_TG, diel, x, g, T_ALL, NULL);
        n_grads++;
        scale_grad(g, n, &gs);
        for (int i = 0; i < n; i++) dg[i] = g[i] - dg[i];
        double gt = 0.0;
        for (int i = 0; i < n; i++) gt = dmax(gt, fabs(g[i]) * dmax(fabs(x[i]), 1.0));
        if (gt / dmax(energy * gs, 1.0) < grad_tol) { /* :300-316 */
            status = 0;
            break;
        }
        if (!hist) {
            /* BFGS inverse-Hessian update, :318-366 */
            for (int i = 0; i < n; i++) {
                double s = 0.0;
                const double *row = H + (size_t)i * n;
                for (int j = 0; j < n; j++) s += row[j] * dg[j];
                hdg[i] = rf(s);
            }
            double fac = 0, fae = 0, sdg = 0, sxi = 0;
            for (int i = 0; i < n; i++) {
                fac += dg[i] * xo[i];
                fae += dg[i] * hdg[i];
                sdg += dg[i] * dg[i];
                sxi += xo[i] * xo[i];
            }
            if (fac * fac > EPS_GUARD * sdg * sxi && fac > 0) {
                double fi = 1.0 / fac, fei = 1.0 / fae;
                for (int i = 0; i < n; i++) {
                    double ai = fi * xo[i] - fei * hdg[i];
                    double *row = H + (size_t)i * n;
                    for (int j = 0; j < n; j++) {
                        double aj = fi * xo[j] - fei * hdg[j];
                        row[j] = rf(row[j] + (fi * xo[i] * xo[j] - fei * hdg[i] * hdg[j] + fae * ai * aj));
                    }
                }
            }
            for (int i = 0; i < n; i++) {
                double s = 0.0;
                const double *row = H + (size_t)i * n;
                for (int j = 0; j < n; j++) s += row[j] * g[j];
                d[i] = -rf(s);
            }
        } else {
            /* L-BFGS history + two-loop, mmff_lbfgs_source_tg.metal:249-292 */
            int M = hist->m;
            double ys = 0.0;
            for (int i = 0; i < n; i++) ys += dg[i] * xo[i];
            if (ys > 1e-10) {
                int slot = hist->idx % M;
                memcpy(hist->S + (size_t)slot * n, xo, nb);
                memcpy(hist->Y + (size_t)slot * n, dg, nb);
                hist->rho[slot] = 1.0 / ys;
                hist->idx++;
                if (hist->count < M) hist->count++;
            }
            double *q = hdg;
            memcpy(q, g, nb);
            for (int j = hist->count - 1; j >= 0; j--) {
                int slot = (hist->idx - 1 - (hist->count - 1 - j)) % M;
                if (slot < 0) slot += M;
                const double *S = hist->S + (size_t)slot * n, *Y = hist->Y + (size_t)slot * n;
                double sq = 0.0;
                for (int i = 0; i < n; i++) sq += S[i] * q[i];
                double aj = hist->rho[slot] * sq;
                hist->alpha[j] = aj;
                const double ajv = rf(aj);
                for (int i = 0; i < n; i++) q[i] = rf(q[i] - ajv * Y[i]);
            }
            if (hist->count > 0) {
                int nw = (hist->idx - 1) % M;
                if (nw < 0) nw += M;
                const double *S = hist->S + (size_t)nw * n, *Y = hist->Y + (size_t)nw * n;
                double sy = 0.0, yy = 0.0;
                for (int i = 0; i < n; i++) {
                    sy += S[i] * Y[i];
                    yy += Y[i] * Y[i];
                }
                double gam = rf(sy / dmax(yy, 1e-30));
                for (int i = 0; i < n; i++) q[i] = rf(q[i] * gam);
            }
            for (int j = 0; j < hist->count; j++) {
                int slot = (hist->idx - 1 - (hist->count - 1 - j)) % M;
                if (slot < 0) slot += M;
                const double *S = hist->S + (size_t)slot * n, *Y = hist->Y + (size_t)slot * n;
                double yq = 0.0;
                for (int i = 0; i < n; i++) yq += Y[i] * q[i];
                double bj = hist->rho[slot] * yq;
                double aj = hist->alpha[j];
                const double cf = rf(aj - bj);
                for (int i = 0; i < n; i++) q[i] = rf(q[i] + cf * S[i]);
            }
            for (int i = 0; i < n; i++) d[i] = -q[i];
        }
    }
    if (e_out) *e_out = energy;
    if (counters) {
        counters[0] = iters;
        counters[1] = n_grads;
        counters[2] = n_evals;
    }
    free(g);
    free(d);
    free(xo);
    free(dg);
    free(hdg);
    free(H);
    return status;
}

int oracle_bfgs_nk(const oracle_mol *m, double *x, int max_iters, double grad_tol, int diel, double *e_out,
                   int32_t *counters) {
    return nk_minimize(m, x, max_iters, grad_tol, diel, NULL, NULL, e_out, counters);
}

int oracle_bfgs_nk_h0(const oracle_mol *m, double *x, int max_iters, double grad_tol, int diel, const double *h0,
                      double *e_out, int32_t *counters) {
    return nk_minimize(m, x, max_iters, grad_tol, diel, NULL, h0, e_out, counters);
}

int oracle_lbfgs_nk(const oracle_mol *m, double *x, int max_iters, double grad_tol, int diel, int m_hist,
                    double *e_out, int32_t *counters) {
    int n = 3 * m->n_atoms;
    lbfgs_hist h;
    h.m = m_hist;
    h.count = 0;
    h.idx = 0;
    h.S = calloc((size_t)m_hist * n, sizeof(double));
    h.Y = calloc((size_t)m_hist * n, sizeof(double));
    h.rho = calloc((size_t)m_hist, sizeof(double));
    h.alpha = calloc((size_t)m_hist, sizeof(double));
    int st = nk_minimize(m, x, max_iters, grad_tol, diel, &h, NULL, e_out, counters);
    free(h.S);
    free(h.Y);
    free(h.rho);
    free(h.alpha);
    return st;
}

/* ======================= BB fused (MEGA family) ======================= */

/* mmff_metal_kernel.py:845-1075 */
void oracle_bb_fused(const oracle_mol *m, double *x, int max_iters, double max_step, double grad_lr, double *e_out) {
    int N = m->n_atoms, n = 3 * N;
    size_t nb = sizeof(double) * (size_t)n;
    double *g = malloc(nb), *pp = malloc(nb), *pg = calloc((size_t)n, sizeof(double));
    memcpy(pp, x, nb);
    double alpha = grad_lr, E_prev = 1e30, E = 0.0;
    for (int it = 0; it < max_iters; it++) {
        E = oracle_energy_grad(m, ORACLE_FAMILY_MEGA, 1, x, g, T_ALL, NULL);
        if (it > 0 && E > E_prev + 1.0) { /* energy guard :1031-1038 */
            memcpy(x, pp, nb);
            alpha *= 0.5;
            alpha = dmax(alpha, 1e-7);
            E = E_prev;
            continue;
        }
        if (it > 0) { /* BB step :1041-1052 */
            double ss = 0.0, sy = 0.0;
            for (int i = 0; i < n; i++) {
                double si = x[i] - pp[i], yi = g[i] - pg[i];
                ss += si * si;
                sy += si * yi;
            }
            if (sy > 1e-20) alpha = dclamp(ss / sy, 1e-6, 0.3);
        }
        E_prev = E;
        memcpy(pp, x, nb);
        memcpy(pg, g, nb);
        for (int a = 0; a < N; a++) { /* per-atom capped step :1063-1071 */
            double gx = g[3 * a], gy = g[3 * a + 1], gz = g[3 * a + 2];
            double gn = sqrt(gx * gx + gy * gy + gz * gz + 1e-30);
            double s = dmin(max_step / gn, alpha);
            x[3 * a] -= s * gx;
            x[3 * a + 1] -= s * gy;
            x[3 * a + 2] -= s * gz;
        }
    }
    if (e_out) *e_out = E;
    free(g);
    free(pp);
    free(pg);
}

/* ======================= threaded batch drivers ======================= */

typedef struct {
    int kind; /* 0 nk, 1 eg, 2 bb */
    const oracle_mol *mols;
    const int32_t *conf_mol;
    const int64_t *conf_off;
    int n_confs;
    double *x;
    const double *xin;
    int max_iters;
    double grad_tol;
    int diel, use_lbfgs, m_hist, family;
    uint32_t mask;
    double max_step, grad_lr;
    double *e_out, *g_out, *et_out;
    int32_t *status_out, *counters_out;
    const double *h0; /* per-atom blocks at (conf_off / 3 + atom) * 9, or NULL */
    int next;
    pthread_mutex_t lock;
} batch_job;

static void *batch_worker(void *arg) {
    batch_job *J = (batch_job *)arg;
    for (;;) {
        pthread_mutex_lock(&J->lock);
        int c = J->next++;
        pthread_mutex_unlock(&J->lock);
        if (c >= J->n_confs) break;
        const oracle_mol *m = &J->mols[J->conf_mol[c]];
        if (J->kind == 0) {
            double *xc = J->x + J->conf_off[c];
            int32_t *cnt = J->counters_out ? J->counters_out + 3 * (size_t)c : NULL;
            double e;
            int st = J->use_lbfgs ? oracle_lbfgs_nk(m, xc, J->max_iters, J->grad_tol, J->diel, J->m_hist, &e, cnt)
                                  : oracle_bfgs_nk_h0(m, xc, J->max_iters, J->grad_tol, J->diel,
                                                      J->h0 ? J->h0 + (J->conf_off[c] / 3) * 9 : NULL, &e, cnt);
            J->e_out[c] = e;
            J->status_out[c] = st;
        } else if (J->kind == 1) {
            const double *xc = J->xin + J->conf_off[c];
            double *gc = J->g_out ? J->g_out + J->conf_off[c] : NULL;
            J->e_out[c] = oracle_energy_grad(m, J->family, J->diel, xc, gc, J->mask,
                                             J->et_out ? J->et_out + 7 * (size_t)c : NULL);
        } else {
            double *xc = J->x + J->conf_off[c];
            oracle_bb_fused(m, xc, J->max_iters, J->max_step, J->grad_lr, &J->e_out[c]);
        }
    }
    return NULL;
}

static void run_batch(batch_job *J, int n_threads) {
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;
    J->next = 0;
    pthread_mutex_init(&J->lock, NULL);
    pthread_t th[256];
    for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, batch_worker, J);
    for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&J->lock);
}

void oracle_bfgs_nk_batch(const oracle_mol *mols, const int32_t *conf_mol, const int64_t *conf_off, int n_confs,
                          double *x, int max_iters, double grad_tol, int diel, int use_lbfgs, int m_hist,
                          double *e_out, int32_t *status_out, int32_t *counters_out, int n_threads) {
    oracle_bfgs_nk_batch_h0(mols, conf_mol, conf_off, n_confs, x, max_iters, grad_tol, diel, use_lbfgs, m_hist, NULL,
                            e_out, status_out, counters_out, n_threads);
}

void oracle_bfgs_nk_batch_h0(const oracle_mol *mols, const int32_t *conf_mol, const int64_t *conf_off, int n_confs,
                             double *x, int max_iters, double grad_tol, int diel, int use_lbfgs, int m_hist,
                             const double *h0, double *e_out, int32_t *status_out, int32_t *counters_out,
                             int n_threads) {
    batch_job J;
    memset(&J, 0, sizeof J);
    J.h0 = h0;
    J.kind = 0;
    J.mols = mols;
    J.conf_mol = conf_mol;
    J.conf_off = conf_off;
    J.n_confs = n_confs;
    J.x = x;
    J.max_iters = max_iters;
    J.grad_tol = grad_tol;
    J.diel = diel;
    J.use_lbfgs = use_lbfgs;
    J.m_hist = m_hist;
    J.e_out = e_out;
    J.status_out = status_out;
    J.counters_out = counters_out;
    run_batch(&J, n_threads);
}

void oracle_energy_grad_batch(const oracle_mol *mols, const int32_t *conf_mol, const int64_t *conf_off, int n_confs,
                              int family, int diel, const double *x, double *e_out, double *g_out, uint32_t mask,
                              double *et_out, int n_threads) {
    batch_job J;
    memset(&J, 0, sizeof J);
    J.kind = 1;
    J.mols = mols;
    J.conf_mol = conf_mol;
    J.conf_off = conf_off;
    J.n_confs = n_confs;
    J.xin = x;
    J.family = family;
    J.diel = diel;
    J.mask = mask;
    J.e_out = e_out;
    J.g_out = g_out;
    J.et_out = et_out;
    run_batch(&J, n_threads);
}

void oracle_bb_fused_batch(const oracle_mol *mols, const int32_t *conf_mol, const int64_t *conf_off, int n_confs,
                           double *x, int max_iters, double max_step, double grad_lr, double *e_out, int n_threads) {
    batch_job J;
    memset(&J, 0, sizeof J);
    J.kind = 2;
    J.mols = mols;
    J.conf_mol = conf_mol;
    J.conf_off = conf_off;
    J.n_confs = n_confs;
    J.x = x;
    J.max_iters = max_iters;
    J.max_step = max_step;
    J.grad_lr = grad_lr;
    J.e_out = e_out;
    run_batch(&J, n_threads);
}

/* ======================= multi_mol lock-step host L-BFGS (MEGA family) ======================= */

/* One conformer's state for mmff_optimize_metal_multi_mol (mmff_metal_optimizer.py:217-535). */
typedef struct {
    int n, N;
    double *x, *g, *S, *Y, *rho, *best_x, *d, *q, *alpha, *xn, *gn;
    int len, head; /* history ring: len pairs, oldest at head */
    double best_e, e_new;
} mm_state;

/* _lbfgs_direction_batch, mmff_metal_optimizer.py:47-74 (history ordered oldest->newest) */
static void mm_direction(mm_state *s, int M) {
    int n = s->n, m = s->len;
    memcpy(s->q, s->g, sizeof(double) * (size_t)n);
    for (int i = m - 1; i >= 0; i--) {
        int sl = (s->head + i) % M;
        const double *S = s->S + (size_t)sl * n, *Y = s->Y + (size_t)sl * n;
        double sq = 0.0;
        for (int k = 0; k < n; k++) sq += S[k] * s->q[k];
        double a = s->rho[sl] * sq;
        s->alpha[i] = a;
        for (int k = 0; k < n; k++) s->q[k] -= a * Y[k];
    }
    double gam = 1.0;
    if (m > 0) {
        int sl = (s->head + m - 1) % M;
        const double *S = s->S + (size_t)sl * n, *Y = s->Y + (size_t)sl * n;
        double sy = 0.0, yy = 0.0;
        for (int k = 0; k < n; k++) {
            sy += S[k] * Y[k];
            yy += Y[k] * Y[k];
        }
        gam = sy / (yy + 1e-30);
    }
    for (int k = 0; k < n; k++) s->q[k] *= gam;
    for (int i = 0; i < m; i++) {
        int sl = (s->head + i) % M;
        const double *S = s->S + (size_t)sl * n, *Y = s->Y + (size_t)sl * n;
        double yr = 0.0;
        for (int k = 0; k < n; k++) yr += Y[k] * s->q[k];
        double beta = s->rho[sl] * yr;
        for (int k = 0; k < n; k++) s->q[k] += S[k] * (s->alpha[i] - beta);
    }
    for (int k = 0; k < n; k++) s->d[k] = -s->q[k];
}

/* one iteration of the main/polish loop body, :363-402 / :436-468 */
static void mm_step(const oracle_mol *mol, mm_state *s, int M, double max_step, double gsc) {
    int n = s->n, N = s->N;
    mm_direction(s, M);
    for (int a = 0; a < N; a++) {
        double *da = s->d + 3 * a;
        double disp = sqrt(da[0] * da[0] + da[1] * da[1] + da[2] * da[2] + 1e-30);
        double f = dmin(max_step / disp, 1.0);
        da[0] *= f;
        da[1] *= f;
        da[2] *= f;
    }
    for (int k = 0; k < n; k++) s->xn[k] = s->x[k] + s->d[k];
    s->e_new = oracle_energy_grad(mol, ORACLE_FAMILY_MEGA, 1, s->xn, s->gn, T_ALL, NULL);
    for (int k = 0; k < n; k++) s->gn[k] *= gsc;
    int sl;
    if (s->len >= M) {
        sl = s->head;
        s->head = (s->head + 1) % M;
    } else {
        sl = (s->head + s->len) % M;
        s->len++;
    }
    double *S = s->S + (size_t)sl * n, *Y = s->Y + (size_t)sl * n;
    double sy = 0.0;
    for (int k = 0; k < n; k++) {
        S[k] = s->xn[k] - s->x[k];
        Y[k] = s->gn[k] - s->g[k];
        sy += S[k] * Y[k];
    }
    s->rho[sl] = sy > 1e-10 ? 1.0 / (sy + 1e-30) : 0.0;
    memcpy(s->x, s->xn, sizeof(double) * (size_t)n);
    memcpy(s->g, s->gn, sizeof(double) * (size_t)n);
}

static double mm_max_atom_gnorm(const mm_state *s) {
    double mx = 0.0;
    for (int a = 0; a < s->N; a++) {
        const double *g = s->g + 3 * a;
        mx = dmax(mx, sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]));
    }
    return mx;
}

typedef struct {
    const oracle_mol *mols;
    const int32_t *conf_mol;
    mm_state *st;
    int n_confs, n_steps, M, next;
    double max_step, gsc;
    pthread_mutex_t lock;
} mm_job;

static void *mm_worker(void *arg) {
    mm_job *J = (mm_job *)arg;
    for (;;) {
        pthread_mutex_lock(&J->lock);
        int c = J->next++;
        pthread_mutex_unlock(&J->lock);
        if (c >= J->n_confs) break;
        for (int k = 0; k < J->n_steps; k++) mm_step(&J->mols[J->conf_mol[c]], &J->st[c], J->M, J->max_step, J->gsc);
    }
    return NULL;
}

static void mm_run(mm_job *J, int n_steps, int n_threads) {
    J->n_steps = n_steps;
    J->next = 0;
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, mm_worker, J);
    for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
}

static double mm_l2_gnorm(const mm_state *s) {
    double a = 0.0;
    for (int k = 0; k < s->n; k++) a += s->g[k] * s->g[k];
    return sqrt(a);
}

/* flags: ORACLE_MM_BATCH selects mmff_optimize_metal_batch (mmff_metal_optimizer.py:77-202):
 * max_iters steps with the best-x guard at every sync point and no convergence exit, no
 * polish pass, and gnorm_out = ||g_sc|| of the last iterate (:186). */
int oracle_multi_mol_ex(const oracle_mol *mols, const int32_t *conf_mol, const int64_t *conf_off, int n_confs,
                        double *x, int max_iters, int eval_freq, double max_step, int M, double gsc, double tol,
                        double *e_out, double *gnorm_out, int32_t *per_conf_iters, int32_t *conv, int n_threads,
                        int flags) {
    const int batch = (flags & ORACLE_MM_BATCH) != 0;
    mm_state *st = calloc((size_t)n_confs, sizeof(mm_state));
    for (int c = 0; c < n_confs; c++) {
        mm_state *s = &st[c];
        const oracle_mol *m = &mols[conf_mol[c]];
        s->N = m->n_atoms;
        s->n = 3 * s->N;
        size_t n = (size_t)s->n;
        s->x = malloc(n * sizeof(double));
        s->g = malloc(n * sizeof(double));
        s->best_x = malloc(n * sizeof(double));
        s->d = malloc(n * sizeof(double));
        s->q = malloc(n * sizeof(double));
        s->xn = malloc(n * sizeof(double));
        s->gn = malloc(n * sizeof(double));
        s->S = calloc((size_t)M * n, sizeof(double));
        s->Y = calloc((size_t)M * n, sizeof(double));
        s->rho = calloc((size_t)M, sizeof(double));
        s->alpha = calloc((size_t)M, sizeof(double));
        memcpy(s->x, x + conf_off[c], n * sizeof(double));
        /* initial E+G, :309-319 */
        double e0 = oracle_energy_grad(m, ORACLE_FAMILY_MEGA, 1, s->x, s->g, T_ALL, NULL);
        for (size_t k = 0; k < n; k++) s->g[k] *= gsc;
        s->best_e = e0;
        memcpy(s->best_x, s->x, n * sizeof(double));
        per_conf_iters[c] = max_iters;
        conv[c] = 0;
    }
    mm_job J;
    J.mols = mols;
    J.conf_mol = conf_mol;
    J.st = st;
    J.n_confs = n_confs;
    J.M = M;
    J.max_step = max_step;
    J.gsc = gsc;
    pthread_mutex_init(&J.lock, NULL);
    int iters_used = max_iters;
    /* main loop, :362-416: steps run in lock-step chunks between sync points */
    int it = 0;
    while (it < max_iters) {
        int to_sync = eval_freq - (it % eval_freq);
        int steps = to_sync;
        if (it + steps > max_iters) steps = max_iters - it;
        mm_run(&J, steps, n_threads);
        it += steps;
        if (it % eval_freq == 0) {
            int all = 1;
            for (int c = 0; c < n_confs; c++) {
                mm_state *s = &st[c];
                if (s->e_new < s->best_e) memcpy(s->best_x, s->x, sizeof(double) * (size_t)s->n);
                s->best_e = dmin(s->e_new, s->best_e);
                if (batch) continue;
                if (!conv[c] && mm_max_atom_gnorm(s) <= tol) {
                    per_conf_iters[c] = it;
                    conv[c] = 1;
                }
                all &= conv[c];
            }
            if (all && !batch) {
                iters_used = it;
                break;
            }
        }
    }
    /* polish, :418-481 (not in the batch form) */
    for (int c = 0; c < n_confs && !batch; c++) {
        mm_state *s = &st[c];
        const oracle_mol *m = &mols[conf_mol[c]];
        memcpy(s->x, s->best_x, sizeof(double) * (size_t)s->n);
        oracle_energy_grad(m, ORACLE_FAMILY_MEGA, 1, s->x, s->g, T_ALL, NULL);
        for (int k = 0; k < s->n; k++) s->g[k] *= gsc;
        s->len = 0;
        s->head = 0;
    }
    int polish = batch ? 0 : max_iters / 4;
    it = 0;
    while (it < polish) {
        int steps = eval_freq - (it % eval_freq);
        if (it + steps > polish) steps = polish - it;
        mm_run(&J, steps, n_threads);
        it += steps;
        if (it % eval_freq == 0) {
            int all = 1;
            for (int c = 0; c < n_confs; c++) {
                mm_state *s = &st[c];
                if (s->e_new < s->best_e) memcpy(s->best_x, s->x, sizeof(double) * (size_t)s->n);
                s->best_e = dmin(s->e_new, s->best_e);
                all &= (mm_max_atom_gnorm(s) <= tol);
            }
            if (all) break;
        }
    }
    pthread_mutex_destroy(&J.lock);
    for (int c = 0; c < n_confs; c++) {
        mm_state *s = &st[c];
        const oracle_mol *m = &mols[conf_mol[c]];
        memcpy(x + conf_off[c], s->best_x, sizeof(double) * (size_t)s->n);
        e_out[c] = oracle_energy_grad(m, ORACLE_FAMILY_MEGA, 1, s->best_x, NULL, T_ALL, NULL);
        gnorm_out[c] = batch ? mm_l2_gnorm(s) : mm_max_atom_gnorm(s);
        free(s->x);
        free(s->g);
        free(s->best_x);
        free(s->d);
        free(s->q);
        free(s->xn);
        free(s->gn);
        free(s->S);
        free(s->Y);
        free(s->rho);
        free(s->alpha);
    }
    free(st);
    return iters_used;
}

int oracle_multi_mol(const oracle_mol *mols, const int32_t *conf_mol, const int64_t *conf_off, int n_confs,
                     double *x, int max_iters, int eval_freq, double max_step, int M, double gsc, double tol,
                     double *e_out, double *gnorm_out, int32_t *per_conf_iters, int32_t *conv, int n_threads) {
    return oracle_multi_mol_ex(mols, conf_mol, conf_off, n_confs, x, max_iters, eval_freq, max_step, M, gsc, tol,
                               e_out, gnorm_out, per_conf_iters, conv, n_threads, 0);
}
