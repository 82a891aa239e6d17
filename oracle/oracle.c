/*
 * CPU oracle of the exmarg segment analysis: the table gather and the
 * arrival-time selection, restated in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests, the smoke
 * check and the CPU-baseline leg of the benchmark call it; the analysis
 * path never does.
 *
 * Built with -O2 -fopenmp -ffp-contract=off: every operation rounds as the
 * written expression says (no fused multiply-add), which is how numba
 * evaluates the reference kernel (no fastmath flags).
 *
 *  oracle_gather       src/exmarg/search/gather.py:32-100 (_gather_cpu), step
 *                      by step; float64 arithmetic on float32 inputs, with the
 *                      phase difference formed in float32 first exactly as
 *                      numba types phH - phL (gather.py:54).
 *  oracle_gather_synth the same with table values produced by the synthetic
 *                      store recipe instead of read from memory (stores too
 *                      large for host memory).
 *  oracle_select       src/exmarg/search/gather.py:374-420 (select_times) with
 *                      the score e^{2 lqL} + e^{2 max_tau lqH} evaluated in
 *                      float64 by oracle_pexp and ties ranked by smaller t.
 *  oracle_pexp         e^x by Cody-Waite reduction and a degree-13 Taylor
 *                      polynomial; the HIP library evaluates the identical
 *                      operation sequence, so both rank times identically.
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NPHI 96

/* ------------------------------------------------------------ pexp */
static const double PEXP_LOG2E = 1.4426950408889634;
static const double PEXP_LN2_HI = 6.93147180369123816490e-01;
static const double PEXP_LN2_LO = 1.90821492927058770002e-10;
static const double PEXP_C[14] = {1.0, 1.0, 0.5, 0.16666666666666666, 0.041666666666666664, 0.008333333333333333, 0.001388888888888889, 0.0001984126984126984, 2.48015873015873e-05, 2.7557319223985893e-06, 2.755731922398589e-07, 2.505210838544172e-08, 2.08767569878681e-09, 1.6059043836821613e-10};

// threads of the OpenMP loops below (the CPU baseline is timed on all cores and on one)
void oracle_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

double oracle_pexp(double x) {
    double k = nearbyint(x * PEXP_LOG2E);
    double r = (x - k * PEXP_LN2_HI) - k * PEXP_LN2_LO;
    double p = PEXP_C[13];
    for (int i = 12; i >= 0; --i) p = p * r + PEXP_C[i];
    return ldexp(p, (int)k);
}

/* ------------------------------------------------------------ synthetic tables */
static inline uint64_t splitmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* value of synthetic table `node` at flat index idx of the (ntau, nphif, nb, nd) layout */
float oracle_synth_value(uint64_t seed, int64_t node, int kind, int64_t idx, int nphif, int nb, int nd,
                         const int64_t* tau_idx) {
    uint64_t key = (seed * 0x100000001B3ull) ^ ((uint64_t)node << 40) ^ (uint64_t)idx;
    uint64_t h = splitmix(key);
    if (kind == 1) {
        int64_t hv = (int64_t)(h >> 41);
        return (float)((double)(hv - (1 << 22)) / (double)(1 << 22));
    }
    int64_t hv = (int64_t)(h >> 55);
    int64_t d = idx % nd, b = (idx / nd) % nb, c = (idx / ((int64_t)nd * nb)) % nphif;
    int64_t r = idx / ((int64_t)nd * nb * nphif);
    int64_t tau = tau_idx[r] < 0 ? -tau_idx[r] : tau_idx[r];
    int64_t bb = b - 56 > 0 ? b - 56 : 0;
    int64_t dd = d - nd / 2;
    if (dd < 0) dd = -dd;
    int64_t num = 8 * bb * (64 - c) - 64 * dd - 32 * tau + hv;
    return (float)((double)num / 4096.0);
}

void oracle_synth_fill(uint64_t seed, int64_t node, int kind, int ntau, int nphif, int nb, int nd,
                       const int64_t* tau_idx, float* out) {
    int64_t n = (int64_t)ntau * nphif * nb * nd;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = oracle_synth_value(seed, node, kind, i, nphif, nb, nd, tau_idx);
}

/* ------------------------------------------------------------ gather */
typedef struct {
    const float* const* tables;  /* table pointers (memory) */
    const int64_t* synth_node;   /* node id per table (synthetic) or NULL */
    uint64_t seed;
    int kind;
} table_src;

static inline float tval(const table_src* ts, int64_t tab, int64_t idx, int nphif, int nb, int nd,
                         const int64_t* tau_idx) {
    if (ts->synth_node) return oracle_synth_value(ts->seed, ts->synth_node[tab], ts->kind, idx, nphif, nb, nd, tau_idx);
    return ts->tables[tab][idx];
}

/* one (template j, lag r, slot n) of the gather: the online (max, sum) over the walked times */
static void gather_one(const float* lqH, const float* phH, const float* lqL, const float* phL, int64_t n_time,
                       const int64_t* mass_ids, const int64_t* tab_ids, int64_t nsel, int ns, int64_t j, int r, int n,
                       const int64_t* tau_idx, int ntau, const table_src* ts, int nphif, int nb, int nd,
                       const double* offbs, double rdb, double offd, double rdd,
                       const int32_t* sel, const float* lw, int64_t n_walk, int use_sel,
                       double* out_m, double* out_s) {
    const double rph = NPHI / 6.283185307179586;
    int64_t mid = mass_ids[j];
    const float* qH = lqH + mid * n_time;
    const float* pH = phH + mid * n_time;
    const float* qL = lqL + mid * n_time;
    const float* pL = phL + mid * n_time;
    int64_t tau0 = ((tau_idx[r] % n_time) + n_time) % n_time;
    int64_t tab = tab_ids[j * ns + n];
    double m = -1e300, s = 0.0;
    for (int64_t k = 0; k < n_walk; ++k) {
        int64_t t = use_sel ? (int64_t)sel[mid * n_walk + k] : k;
        int64_t i1 = t + tau0;
        if (i1 >= n_time) i1 -= n_time;
        float dph = pH[i1] - pL[t];               /* float32 - float32 stays float32 */
        double x3 = (double)dph * rph;
        if (x3 < 0.0) x3 += NPHI;
        if (x3 >= NPHI) x3 -= NPHI;
        int c0 = (int)x3;
        if (c0 >= NPHI) c0 = NPHI - 1;
        else if (c0 < 0) c0 = 0;
        double fc = x3 - c0;
        int c1 = (c0 + 1) % NPHI;
        double aH = qH[i1], aL = qL[t];
        double xbr = aL * rdb + offbs[j * ns + n];
        double xb = xbr;
        if (xb < 0.0) xb = 0.0;
        else if (xb > nb - 1 - 1e-6) xb = nb - 1 - 1e-6;
        double aLc = aL + (xb - xbr) / rdb;
        double xd = (aH - aLc) * rdd + offd;
        if (xd < 0.0) xd = 0.0;
        else if (xd > nd - 1 - 1e-6) xd = nd - 1 - 1e-6;
        int d0 = (int)xd;
        double fdf = xd - d0, fd0 = 1.0 - fdf;
        int d1 = d0 + 1 < nd ? d0 + 1 : nd - 1;
        int b0 = (int)xb;
        double fbf = xb - b0, fb0 = 1.0 - fbf;
        int b1 = b0 + 1 < nb ? b0 + 1 : nb - 1;
        double acc = 0.0;
        for (int cc = 0; cc < 2; ++cc) {
            int ci = cc ? c1 : c0;
            if (ci > NPHI / 2) ci = NPHI - ci;
            double fci = cc ? fc : 1.0 - fc;
            int64_t rc = ((int64_t)r * nphif + ci) * nb;
            double t00 = tval(ts, tab, (rc + b0) * nd + d0, nphif, nb, nd, tau_idx);
            double t01 = tval(ts, tab, (rc + b0) * nd + d1, nphif, nb, nd, tau_idx);
            double t10 = tval(ts, tab, (rc + b1) * nd + d0, nphif, nb, nd, tau_idx);
            double t11 = tval(ts, tab, (rc + b1) * nd + d1, nphif, nb, nd, tau_idx);
            double v = fb0 * (fd0 * t00 + fdf * t01) + fbf * (fd0 * t10 + fdf * t11);
            acc += fci * v;
        }
        if (use_sel) acc += (double)lw[mid * n_walk + k];
        if (acc > m) {
            s = s * exp(m - acc) + 1.0;
            m = acc;
        } else {
            s += exp(acc - m);
        }
    }
    out_m[((int64_t)n * ntau + r) * nsel + j] = m;
    out_s[((int64_t)n * ntau + r) * nsel + j] = s;
}

/* every (template, lag, slot).  slab_order 0: the reference's job split, one job per (template, lag)
 * pair handed to the threads (gather.py:41, prange over nsel * ntau), the slots inside a job;
 * slab_order 1: slot by slot and lag by lag, the threads sharing the templates, so that they read one
 * lag slab of one table at a time (the same values and operations, a cache-friendly order) */
static void gather_core(const float* lqH, const float* phH, const float* lqL, const float* phL, int64_t n_time,
                        const int64_t* mass_ids, const int64_t* tab_ids, int64_t nsel, int ns,
                        const int64_t* tau_idx, int ntau, const table_src* ts, int nphif, int nb, int nd,
                        const double* offbs, double rdb, double offd, double rdd,
                        const int32_t* sel, const float* lw, int64_t n_walk_sel, int use_sel,
                        double* out_m, double* out_s, int slab_order) {
    const int64_t n_walk = use_sel ? n_walk_sel : n_time;
    if (!slab_order) {
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t job = 0; job < nsel * ntau; ++job)
            for (int n = 0; n < ns; ++n)
                gather_one(lqH, phH, lqL, phL, n_time, mass_ids, tab_ids, nsel, ns, job / ntau, (int)(job % ntau), n,
                           tau_idx, ntau, ts, nphif, nb, nd, offbs, rdb, offd, rdd, sel, lw, n_walk, use_sel, out_m,
                           out_s);
        return;
    }
    for (int n = 0; n < ns; ++n)
        for (int r = 0; r < ntau; ++r) {
#pragma omp parallel for schedule(dynamic, 8)
            for (int64_t j = 0; j < nsel; ++j)
                gather_one(lqH, phH, lqL, phL, n_time, mass_ids, tab_ids, nsel, ns, j, r, n, tau_idx, ntau, ts, nphif,
                           nb, nd, offbs, rdb, offd, rdd, sel, lw, n_walk, use_sel, out_m, out_s);
        }
}

void oracle_gather(const float* lqH, const float* phH, const float* lqL, const float* phL, int64_t n_time,
                   const int64_t* mass_ids, const int64_t* tab_ids, int64_t nsel, int ns,
                   const int64_t* tau_idx, int ntau, const float* const* tables, int nphif, int nb, int nd,
                   const double* offbs, double rdb, double offd, double rdd,
                   const int32_t* sel, const float* lw, int64_t n_walk, int use_sel,
                   double* out_m, double* out_s, int slab_order) {
    table_src ts = {tables, NULL, 0, 0};
    gather_core(lqH, phH, lqL, phL, n_time, mass_ids, tab_ids, nsel, ns, tau_idx, ntau, &ts, nphif, nb, nd,
                offbs, rdb, offd, rdd, sel, lw, n_walk, use_sel, out_m, out_s, slab_order);
}

void oracle_gather_synth(const float* lqH, const float* phH, const float* lqL, const float* phL, int64_t n_time,
                         const int64_t* mass_ids, const int64_t* tab_ids, int64_t nsel, int ns,
                         const int64_t* tau_idx, int ntau, const int64_t* node_of_table, uint64_t seed, int kind,
                         int nphif, int nb, int nd, const double* offbs, double rdb, double offd, double rdd,
                         const int32_t* sel, const float* lw, int64_t n_walk, int use_sel,
                         double* out_m, double* out_s, int slab_order) {
    table_src ts = {NULL, node_of_table, seed, kind};
    gather_core(lqH, phH, lqL, phL, n_time, mass_ids, tab_ids, nsel, ns, tau_idx, ntau, &ts, nphif, nb, nd,
                offbs, rdb, offd, rdd, sel, lw, n_walk, use_sel, out_m, out_s, slab_order);
}

/* ------------------------------------------------------------ select_times */
typedef struct {
    double score;
    int32_t t;
} scored;

static int cmp_rank(const void* a, const void* b) {   /* larger score first, then smaller t */
    const scored* x = (const scored*)a;
    const scored* y = (const scored*)b;
    if (x->score > y->score) return -1;
    if (x->score < y->score) return 1;
    return (x->t > y->t) - (x->t < y->t);
}

static int cmp_int(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* scores of one template: e^{2 lqL(t)} + e^{2 max_{lo<=tau<=hi} lqH(t+tau)}, cyclic in t */
void oracle_scores(const float* lqH, const float* lqL, int64_t nt, int lo, int hi, double* score) {
    for (int64_t t = 0; t < nt; ++t) {
        float mx = -INFINITY;
        for (int tau = lo; tau <= hi; ++tau) {
            int64_t i = ((t + tau) % nt + nt) % nt;
            if (lqH[i] > mx) mx = lqH[i];
        }
        score[t] = oracle_pexp(2.0 * (double)lqL[t]) + oracle_pexp(2.0 * (double)mx);
    }
}

/* sel (nm, K+S) int32 sorted times and lw (nm, K+S) float32; pick (nm, S) int64 in [0, nt-K) */
void oracle_select(const float* lqH, const float* lqL, int64_t nm, int64_t nt, int lo, int hi, int K, int S,
                   const int64_t* pick, float lw_rand, int32_t* sel, float* lw) {
#pragma omp parallel
    {
        scored* buf = (scored*)malloc(sizeof(scored) * nt);
        double* score = (double*)malloc(sizeof(double) * nt);
        unsigned char* istop = (unsigned char*)malloc(nt);
        int32_t* rest = (int32_t*)malloc(sizeof(int32_t) * nt);
        int32_t* picked = (int32_t*)malloc(sizeof(int32_t) * (S > 0 ? S : 1));
        int32_t* top = (int32_t*)malloc(sizeof(int32_t) * (K > 0 ? K : 1));
#pragma omp for schedule(dynamic, 1)
        for (int64_t m = 0; m < nm; ++m) {
            oracle_scores(lqH + m * nt, lqL + m * nt, nt, lo, hi, score);
            for (int64_t t = 0; t < nt; ++t) {
                buf[t].score = score[t];
                buf[t].t = (int32_t)t;
            }
            qsort(buf, nt, sizeof(scored), cmp_rank);
            memset(istop, 0, nt);
            for (int i = 0; i < K; ++i) {
                top[i] = buf[i].t;
                istop[buf[i].t] = 1;
            }
            int64_t nr = 0;
            for (int64_t t = 0; t < nt; ++t)
                if (!istop[t]) rest[nr++] = (int32_t)t;
            for (int s = 0; s < S; ++s) picked[s] = rest[pick[m * S + s]];
            qsort(top, K, sizeof(int32_t), cmp_int);
            qsort(picked, S, sizeof(int32_t), cmp_int);
            /* merge the two sorted lists (disjoint values) */
            int i = 0, j = 0, o = 0;
            int32_t* so = sel + m * (int64_t)(K + S);
            float* lo_ = lw + m * (int64_t)(K + S);
            while (i < K || j < S) {
                if (j >= S || (i < K && top[i] < picked[j])) {
                    so[o] = top[i++];
                    lo_[o++] = 0.0f;
                } else {
                    so[o] = picked[j++];
                    lo_[o++] = lw_rand;
                }
            }
        }
        free(buf);
        free(score);
        free(istop);
        free(rest);
        free(picked);
        free(top);
    }
}
