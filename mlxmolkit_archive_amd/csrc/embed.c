/*
 * Start-geometry embedder for the minimizer workloads (host code, C, threaded).
 *
 * Stands in for the RDKit ETKDG embedding that produces the reference's start
 * geometries (mlxmolkit/conformer_pipeline_v2.py:176-290 runs DG -> 4D->3D -> ETK on the
 * GPU; the benches use RDKit EmbedMultipleConfs). Each conformer starts from seeded
 * random 4D coordinates and minimises a distance-restraint energy:
 *   equality springs   k_eq (d - d0)^2          for 1-2, 1-3 and ring-closure targets
 *   lower-bound walls  k_lb (d_min - d)^2 if d < d_min for every other pair
 * with the fourth coordinate penalised in a second phase and dropped in a third,
 * the same 4D-then-collapse idea the reference pipeline uses (:216-229).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t s;
} rng_t;

static inline uint64_t rng_next(rng_t *r) { /* splitmix64 */
    uint64_t z = (r->s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline double rng_unif(rng_t *r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }
static inline double rng_norm(rng_t *r) {
    double u1 = rng_unif(r), u2 = rng_unif(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

typedef struct {
    int n_atoms;
    int n_restr;
    const int32_t *ri, *rj; /* restraint pairs */
    const float *rd;        /* target distance */
    const int8_t *rkind;    /* 0 equality, 1 lower bound */
} embed_mol;

/* energy and gradient over the restraints listed in act[0..nact) */
static double restraint_energy_grad(const embed_mol *m, const int32_t *act, int nact, const double *x, double *g,
                                    int dim, double w4) {
    int n = m->n_atoms;
    memset(g, 0, sizeof(double) * (size_t)n * dim);
    double e = 0.0;
    for (int a = 0; a < nact; a++) {
        const int t = act[a];
        int i = m->ri[t], j = m->rj[t];
        double d2 = 0.0, dv[4];
        for (int c = 0; c < dim; c++) {
            dv[c] = x[i * dim + c] - x[j * dim + c];
            d2 += dv[c] * dv[c];
        }
        double d0 = m->rd[t];
        /* an inactive lower-bound wall costs no square root */
        if (m->rkind[t] != 0 && d2 + 1e-12 >= d0 * d0) continue;
        double d = sqrt(d2 + 1e-12);
        double diff, k;
        if (m->rkind[t] == 0) {
            diff = d - d0;
            k = 100.0;
        } else {
            if (d >= d0) continue;
            diff = d - d0;
            k = 10.0;
        }
        e += k * diff * diff;
        double f = 2.0 * k * diff / d;
        for (int c = 0; c < dim; c++) {
            g[i * dim + c] += f * dv[c];
            g[j * dim + c] -= f * dv[c];
        }
    }
    if (dim == 4 && w4 > 0.0) {
        for (int i = 0; i < n; i++) {
            double w = x[i * 4 + 3];
            e += w4 * w * w;
            g[i * 4 + 3] += 2.0 * w4 * w;
        }
    }
    return e;
}

/* Restraints that can act within the next REBUILD steps: every equality spring, and the
 * lower-bound walls of pairs closer than d_min + 3 A. An atom moves at most 0.3 A per
 * step (trial steps included), so over REBUILD + 1 = 5 steps a pair closes by at most 3 A
 * and no wall outside the list can switch on before the next rebuild. */
#define REBUILD 4
static int active_restraints(const embed_mol *m, const double *x, int dim, int32_t *act) {
    int nact = 0;
    for (int t = 0; t < m->n_restr; t++) {
        if (m->rkind[t] == 0) {
            act[nact++] = t;
            continue;
        }
        const int i = m->ri[t], j = m->rj[t];
        double d2 = 0.0;
        for (int c = 0; c < dim; c++) {
            const double v = x[i * dim + c] - x[j * dim + c];
            d2 += v * v;
        }
        const double r = m->rd[t] + 3.0;
        if (d2 < r * r) act[nact++] = t;
    }
    return nact;
}

/* gradient descent with a per-atom step cap and simple step adaptation */
static void relax(const embed_mol *m, double *x, double *g, double *xt, double *gt, int32_t *act, int dim, double w4,
                  int steps) {
    int n = m->n_atoms;
    double step = 0.002;
    int nact = active_restraints(m, x, dim, act);
    double e = restraint_energy_grad(m, act, nact, x, g, dim, w4);
    for (int s = 0; s < steps; s++) {
        if (s > 0 && s % REBUILD == 0) nact = active_restraints(m, x, dim, act);
        for (int i = 0; i < n; i++) {
            double gn = 0.0;
            for (int c = 0; c < dim; c++) gn += g[i * dim + c] * g[i * dim + c];
            gn = sqrt(gn);
            double sc = step;
            if (gn * sc > 0.3) sc = 0.3 / gn;
            for (int c = 0; c < dim; c++) xt[i * dim + c] = x[i * dim + c] - sc * g[i * dim + c];
        }
        double et = restraint_energy_grad(m, act, nact, xt, gt, dim, w4);
        if (et <= e) {
            memcpy(x, xt, sizeof(double) * (size_t)n * dim);
            memcpy(g, gt, sizeof(double) * (size_t)n * dim);
            e = et;
            step *= 1.2;
        } else {
            step *= 0.5;
            if (step < 1e-7) step = 1e-7;
        }
    }
}

typedef struct {
    const embed_mol *mols;
    const int32_t *conf_mol;
    const int64_t *conf_off;
    const uint64_t *seeds;
    int n_confs;
    float noise;
    float *out;
    int next;
    pthread_mutex_t lock;
} embed_job;

static void *embed_worker(void *arg) {
    embed_job *J = (embed_job *)arg;
    double *buf = NULL;
    size_t cap = 0;
    int32_t *act = NULL;
    int act_cap = 0;
    for (;;) {
        pthread_mutex_lock(&J->lock);
        int c = J->next++;
        pthread_mutex_unlock(&J->lock);
        if (c >= J->n_confs) break;
        const embed_mol *m = &J->mols[J->conf_mol[c]];
        int n = m->n_atoms;
        size_t need = (size_t)n * 4 * 4;
        if (need > cap) {
            free(buf);
            buf = malloc(sizeof(double) * need);
            cap = need;
        }
        if (m->n_restr > act_cap) {
            free(act);
            act = malloc(sizeof(int32_t) * (size_t)m->n_restr);
            act_cap = m->n_restr;
        }
        double *x = buf, *g = buf + n * 4, *xt = buf + n * 8, *gt = buf + n * 12;
        rng_t r = {J->seeds[c]};
        double box = 1.5 * cbrt((double)n) + 1.0;
        for (int i = 0; i < n * 4; i++) x[i] = (rng_unif(&r) * 2.0 - 1.0) * box;
        relax(m, x, g, xt, gt, act, 4, 0.0, 250);
        relax(m, x, g, xt, gt, act, 4, 2.0, 150);
        /* collapse to 3D */
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) x[i * 3 + k] = x[i * 4 + k];
        relax(m, x, g, xt, gt, act, 3, 0.0, 200);
        float *o = J->out + J->conf_off[c];
        for (int i = 0; i < n * 3; i++) o[i] = (float)(x[i] + (double)J->noise * rng_norm(&r));
    }
    free(buf);
    free(act);
    return NULL;
}

/* Embed n_confs conformers; conformer c uses mols[conf_mol[c]], writes 3N floats at
 * out + conf_off[c], seeded by seeds[c], then adds N(0, noise) to every coordinate. */
void mmff_embed_batch(const embed_mol *mols, const int32_t *conf_mol, const int64_t *conf_off,
                      const uint64_t *seeds, int n_confs, float noise, float *out, int n_threads) {
    embed_job J;
    J.mols = mols;
    J.conf_mol = conf_mol;
    J.conf_off = conf_off;
    J.seeds = seeds;
    J.n_confs = n_confs;
    J.noise = noise;
    J.out = out;
    J.next = 0;
    pthread_mutex_init(&J.lock, NULL);
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, embed_worker, &J);
    for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&J.lock);
}
