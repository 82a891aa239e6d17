/*
 * The C oracle run under AddressSanitizer and UndefinedBehaviorSanitizer (test infrastructure).
 *
 * make -C oracle sanitize builds this driver together with oracle.c, both instrumented, and runs it:
 * every public function of the oracle on small inputs that reach its edges (lags that wrap, beta and
 * Delta clamps at both ends, phase bins on both sides of the mirror, every slot, both job orders,
 * walks over every time and over a selection, selections with K or S equal to one), so that an
 * out-of-bounds access, an overflow or a use of uninitialised memory fails the run.  The values are
 * checked too where a closed form fixes them:
 *   a constant table c gives m + ln s = c + ln(n_walk) (every time) and c + ln sum_k e^{lw_k}
 *   (a selection);
 *   the selection is sorted, holds K + S distinct times and weighs drawn times by lw_rand;
 *   oracle_pexp agrees with the C library's exp to 2 ulp over [-700, 700].
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void oracle_set_threads(int n);
double oracle_pexp(double x);
float oracle_synth_value(uint64_t seed, int64_t node, int kind, int64_t idx, int nphif, int nb, int nd,
                         const int64_t* tau_idx);
void oracle_synth_fill(uint64_t seed, int64_t node, int kind, int ntau, int nphif, int nb, int nd,
                       const int64_t* tau_idx, float* out);
void oracle_gather(const float* lqH, const float* phH, const float* lqL, const float* phL, int64_t n_time,
                   const int64_t* mass_ids, const int64_t* tab_ids, int64_t nsel, int ns, const int64_t* tau_idx,
                   int ntau, const float* const* tables, int nphif, int nb, int nd, const double* offbs, double rdb,
                   double offd, double rdd, const int32_t* sel, const float* lw, int64_t n_walk, int use_sel,
                   double* out_m, double* out_s, int slab_order);
void oracle_gather_synth(const float* lqH, const float* phH, const float* lqL, const float* phL, int64_t n_time,
                         const int64_t* mass_ids, const int64_t* tab_ids, int64_t nsel, int ns,
                         const int64_t* tau_idx, int ntau, const int64_t* node_of_table, uint64_t seed, int kind,
                         int nphif, int nb, int nd, const double* offbs, double rdb, double offd, double rdd,
                         const int32_t* sel, const float* lw, int64_t n_walk, int use_sel, double* out_m,
                         double* out_s, int slab_order);
void oracle_scores(const float* lqH, const float* lqL, int64_t nt, int lo, int hi, double* score);
void oracle_select(const float* lqH, const float* lqL, int64_t nm, int64_t nt, int lo, int hi, int K, int S,
                   const int64_t* pick, float lw_rand, int32_t* sel, float* lw);

static int failures = 0;

static void expect(int ok, const char* what, double got, double want) {
    if (!ok) {
        fprintf(stderr, "FAIL %s: got %.17g, want %.17g\n", what, got, want);
        ++failures;
    }
}

static uint64_t rng = 0x2545F4914F6CDD1Dull;
static double uniform(void) {   /* xorshift64*, in [0, 1) */
    rng ^= rng >> 12;
    rng ^= rng << 25;
    rng ^= rng >> 27;
    return (double)((rng * 0x2545F4914F6CDD1Dull) >> 11) / 9007199254740992.0;
}

enum { NT = 64, NM = 3, NTAU = 3, NPHIF = 49, NB = 4, ND = 5, NS = 2, NTAB = 3 };

static const int64_t TAU[NTAU] = {0, 1, -1};

int main(void) {
    oracle_set_threads(2);
    const int64_t tsize = (int64_t)NTAU * NPHIF * NB * ND;
    float* tab[NTAB];
    for (int k = 0; k < NTAB; ++k) {
        tab[k] = (float*)malloc(sizeof(float) * tsize);
        for (int64_t i = 0; i < tsize; ++i) tab[k][i] = k == 0 ? 1.25f : (float)(4.0 * uniform() - 2.0);
    }
    /* trajectories reaching both clamps of beta and Delta and every phase bin */
    float *lqH = malloc(sizeof(float) * NM * NT), *phH = malloc(sizeof(float) * NM * NT);
    float *lqL = malloc(sizeof(float) * NM * NT), *phL = malloc(sizeof(float) * NM * NT);
    for (int i = 0; i < NM * NT; ++i) {
        lqH[i] = (float)(12.0 * uniform() - 6.0);
        lqL[i] = (float)(12.0 * uniform() - 6.0);
        phH[i] = (float)(6.283185307179586 * uniform() - 3.141592653589793);
        phL[i] = (float)(6.283185307179586 * uniform() - 3.141592653589793);
    }
    const double rdb = 1.0, offd = 2.0, rdd = 1.0;
    int64_t mass_ids[NM] = {2, 0, 1};
    int64_t tab_all[NM * NS] = {1, 2, 0, 0, 2, 1};
    int64_t tab_const[NM * NS] = {0, 0, 0, 0, 0, 0};
    double offbs[NM * NS];
    for (int i = 0; i < NM * NS; ++i) offbs[i] = 1.5 * uniform();
    double *om = malloc(sizeof(double) * NS * NTAU * NM), *os = malloc(sizeof(double) * NS * NTAU * NM);
    double *om2 = malloc(sizeof(double) * NS * NTAU * NM), *os2 = malloc(sizeof(double) * NS * NTAU * NM);

    /* every time: a constant table gives c + ln n_time; both job orders agree bit for bit */
    oracle_gather(lqH, phH, lqL, phL, NT, mass_ids, tab_const, NM, NS, TAU, NTAU, (const float* const*)tab, NPHIF, NB,
                  ND, offbs, rdb, offd, rdd, NULL, NULL, NT, 0, om, os, 0);
    for (int i = 0; i < NS * NTAU * NM; ++i)
        expect(fabs(om[i] + log(os[i]) - (1.25 + log((double)NT))) < 1e-12, "constant table, every time",
               om[i] + log(os[i]), 1.25 + log((double)NT));
    oracle_gather(lqH, phH, lqL, phL, NT, mass_ids, tab_all, NM, NS, TAU, NTAU, (const float* const*)tab, NPHIF, NB, ND,
                  offbs, rdb, offd, rdd, NULL, NULL, NT, 0, om, os, 0);
    oracle_gather(lqH, phH, lqL, phL, NT, mass_ids, tab_all, NM, NS, TAU, NTAU, (const float* const*)tab, NPHIF, NB, ND,
                  offbs, rdb, offd, rdd, NULL, NULL, NT, 0, om2, os2, 1);
    expect(memcmp(om, om2, sizeof(double) * NS * NTAU * NM) == 0 && memcmp(os, os2, sizeof(double) * NS * NTAU * NM) == 0,
           "job orders agree", 0, 0);

    /* a selection of K + S times per template */
    for (int K = 1; K <= 8; K += 7)
        for (int S = 1; S <= 8; S += 7) {
            const int nw = K + S;
            int64_t* pick = malloc(sizeof(int64_t) * NM * S);
            for (int i = 0; i < NM * S; ++i) pick[i] = (int64_t)(uniform() * (NT - K));
            int32_t* sel = malloc(sizeof(int32_t) * NM * nw);
            float* lw = malloc(sizeof(float) * NM * nw);
            const float lw_rand = (float)log((double)(NT - K) / S);
            oracle_select(lqH, lqL, NM, NT, -1, 1, K, S, pick, lw_rand, sel, lw);
            for (int m = 0; m < NM; ++m) {
                int nrand = 0;
                for (int k = 0; k < nw; ++k) {
                    const int32_t t = sel[m * nw + k];
                    expect(t >= 0 && t < NT, "selected time in range", t, 0);
                    if (k > 0) expect(sel[m * nw + k - 1] <= t, "selection sorted", t, sel[m * nw + k - 1]);
                    nrand += lw[m * nw + k] == lw_rand;
                    expect(lw[m * nw + k] == 0.0f || lw[m * nw + k] == lw_rand, "selection weight", lw[m * nw + k], 0);
                }
                expect(nrand == S || (lw_rand == 0.0f && nrand == nw), "drawn times weighed by lw_rand", nrand, S);
            }
            oracle_gather(lqH, phH, lqL, phL, NT, mass_ids, tab_const, NM, NS, TAU, NTAU, (const float* const*)tab,
                          NPHIF, NB, ND, offbs, rdb, offd, rdd, sel, lw, nw, 1, om, os, 1);
            for (int n = 0; n < NS; ++n)
                for (int r = 0; r < NTAU; ++r)
                    for (int j = 0; j < NM; ++j) {
                        const int64_t m = mass_ids[j];
                        double want = 0.0;
                        for (int k = 0; k < nw; ++k) want += exp((double)lw[m * nw + k]);
                        want = 1.25 + log(want);
                        const int64_t o = ((int64_t)n * NTAU + r) * NM + j;
                        expect(fabs(om[o] + log(os[o]) - want) < 1e-12, "constant table, selection",
                               om[o] + log(os[o]), want);
                    }
            oracle_gather(lqH, phH, lqL, phL, NT, mass_ids, tab_all, NM, NS, TAU, NTAU, (const float* const*)tab, NPHIF,
                          NB, ND, offbs, rdb, offd, rdd, sel, lw, nw, 1, om, os, 0);
            free(pick);
            free(sel);
            free(lw);
        }

    /* synthetic tables: the fill agrees with the per-value recipe, and the gather reads them */
    const int64_t node_of_table[NTAB] = {0, 7, 52};
    float* fill = malloc(sizeof(float) * tsize);
    for (int kind = 0; kind < 2; ++kind) {
        oracle_synth_fill(3, 7, kind, NTAU, NPHIF, NB, ND, TAU, fill);
        for (int64_t i = 0; i < tsize; i += 97)
            expect(fill[i] == oracle_synth_value(3, 7, kind, i, NPHIF, NB, ND, TAU), "synthetic fill", fill[i], 0);
        oracle_gather_synth(lqH, phH, lqL, phL, NT, mass_ids, tab_all, NM, NS, TAU, NTAU, node_of_table, 3, kind,
                            NPHIF, NB, ND, offbs, rdb, offd, rdd, NULL, NULL, NT, 0, om, os, kind);
        for (int i = 0; i < NS * NTAU * NM; ++i) expect(isfinite(om[i]) && os[i] >= 1.0, "synthetic gather", os[i], 1);
    }

    /* the scores and the exponential */
    double* score = malloc(sizeof(double) * NT);
    oracle_scores(lqH, lqL, NT, -1, 1, score);
    for (int t = 0; t < NT; ++t) expect(score[t] > 0.0 && isfinite(score[t]), "score", score[t], 0);
    for (double x = -700.0; x <= 700.0; x += 0.37) {
        const double a = oracle_pexp(x), b = exp(x);
        expect(fabs(a - b) <= 2.0 * (nextafter(b, INFINITY) - b), "pexp", a, b);
    }

    for (int k = 0; k < NTAB; ++k) free(tab[k]);
    free(lqH);
    free(phH);
    free(lqL);
    free(phL);
    free(om);
    free(os);
    free(om2);
    free(os2);
    free(fill);
    free(score);
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("oracle under sanitizers: all checks passed\n");
    return 0;
}
