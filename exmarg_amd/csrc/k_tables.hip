// Stage 3 of the tables: the sum over the orders (scripts/build_tables/build_bd.py:52-79).
//
// For one node, stage 2 gives lnK_n(tau, q1, q2, phi) for the orders n = 0 .. nord - 1.  Stage 3 needs
// only its last column (q2 = q_max) and last row (q1 = q_max) of every (q1, q2) block:
//   g_n(Delta) = col_n(q1) - 2 n ln q_max     for Delta = ln q1 - ln q_max <= 0,
//   g_n(Delta) = row_n(q2) - 2 n ln q2        for Delta = ln q_max - ln q2 >= 0,
// and, on the beta grid, E_n(beta, Delta) = n beta + g_n(Delta) and for every truncation order n_max
//   T(tau, phi, beta, Delta) = max_{n <= n_max} E_n + ln sum_{n <= n_max} exp(E_n - max)
// in float64, rounded to float32.  The products and sums are formed as the reference forms them (no
// fused multiply-add), the sum runs over n in increasing order.
//
// One block of 512 threads per (tau, phi, 64 Delta columns): g_n of the block's columns is staged in LDS
// (float64), lane l takes column l and the 8 waves take every 8th beta row.
#include <cmath>

#include "exm_common.h"

#pragma clang fp contract(off)

namespace exm {

namespace {

constexpr int BD_THREADS = 512;
constexpr int BD_GROUPS = BD_THREADS / 64;
constexpr int BD_DT = 64;            // Delta columns per block
constexpr int BD_MAXORD = 128;       // orders one block stages (64 KB of LDS)
constexpr int BD_MAXLAD = 32;        // truncation orders per call

__global__ __launch_bounds__(BD_THREADS) void k_bd_sum_orders(const float* __restrict__ col,
                                                             const float* __restrict__ row,
                                                             const double* __restrict__ lnq,
                                                             const double* __restrict__ beta, int nord, int ntau,
                                                             int nq, int nphif, int nb, const int* __restrict__ ladder,
                                                             int nlad, float* __restrict__ out) {
    extern __shared__ double gs[];   // [nord][BD_DT]
    const int nd = 2 * nq - 1;
    const int tp = blockIdx.y;       // tau * nphif + phi
    const int ti = tp / nphif, pi = tp - ti * nphif;
    const int d0 = blockIdx.x * BD_DT;
    const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const double lq_max = lnq[nq - 1];
    for (int i = threadIdx.x; i < nord * BD_DT; i += BD_THREADS) {
        const int n = i / BD_DT, dl = i - n * BD_DT, d = d0 + dl;
        double g = 0.0;
        if (d < nd) {
            const double twon = 2.0 * (double)n;
            if (d < nq) {
                const double c = (double)col[(((int64_t)n * ntau + ti) * nq + d) * nphif + pi];
                g = c - twon * lq_max;
            } else {
                const int q2 = 2 * nq - 2 - d;
                const double rw = (double)row[(((int64_t)n * ntau + ti) * nq + q2) * nphif + pi];
                g = rw - twon * lnq[q2];
            }
        }
        gs[i] = g;
    }
    __syncthreads();
    const int d = d0 + lane;
    if (d >= nd) return;             // no barrier follows
    const int64_t plane = (int64_t)nb * nd;
    for (int b = grp; b < nb; b += BD_GROUPS) {
        const double bt = beta[b];
        for (int l = 0; l < nlad; ++l) {
            const int nm = ladder[l];
            double mx = -INFINITY;
            for (int n = 0; n <= nm; ++n) {
                const double e = (double)n * bt + gs[n * BD_DT + lane];
                mx = e > mx ? e : mx;
            }
            double tot = 0.0;
            for (int n = 0; n <= nm; ++n) tot += exp(((double)n * bt + gs[n * BD_DT + lane]) - mx);
            out[((int64_t)l * ntau * nphif + tp) * plane + (int64_t)b * nd + d] = (float)(mx + log(tot));
        }
    }
}

}  // namespace

// the stage-3 tables of one node for every truncation order of `ladder` (checked by the caller)
int tables_sum_orders(const float* col, const float* row, int nord, int ntau, int nq, int nphif, const double* lnq,
                      const double* beta, int nb, const int* ladder, int nlad, float* const* out) {
    EXM_CHECK(col && row && lnq && beta && ladder && out, EXM_ERR_ARG, "NULL argument");
    EXM_CHECK(nord >= 1 && nord <= BD_MAXORD, EXM_ERR_ARG, "%d orders outside [1, %d]", nord, BD_MAXORD);
    EXM_CHECK(ntau >= 1 && nq >= 2 && nphif >= 1 && nb >= 1, EXM_ERR_ARG, "bad stage-3 shape");
    EXM_CHECK(nq <= 32768, EXM_ERR_ARG, "q grid of %d points too long", nq);
    EXM_CHECK(nlad >= 1 && nlad <= BD_MAXLAD, EXM_ERR_ARG, "%d truncation orders outside [1, %d]", nlad, BD_MAXLAD);
    for (int l = 0; l < nlad; ++l) {
        EXM_CHECK(ladder[l] >= 0 && ladder[l] < nord, EXM_ERR_ARG, "truncation order %d outside [0, %d)", ladder[l],
                  nord);
        EXM_CHECK(l == 0 || ladder[l] > ladder[l - 1], EXM_ERR_ARG, "truncation orders must ascend");
        EXM_CHECK(out[l], EXM_ERR_ARG, "NULL output table %d", l);
    }
    Context& c = ctx();
    const int nd = 2 * nq - 1;
    const size_t in_bytes = sizeof(float) * (size_t)nord * ntau * nq * nphif;
    const size_t tab_bytes = sizeof(float) * (size_t)ntau * nphif * nb * nd;
    const size_t small = sizeof(double) * (size_t)(nq + nb) + sizeof(int) * (size_t)nlad;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t total = 2 * al(in_bytes) + al(small) + al(tab_bytes * (size_t)nlad);
    void* buf = nullptr;
    EXM_TRY(dev_alloc(&buf, total));
    char* base = (char*)buf;
    float* d_col = (float*)base;
    float* d_row = (float*)(base + al(in_bytes));
    double* d_lnq = (double*)(base + 2 * al(in_bytes));
    double* d_beta = d_lnq + nq;
    int* d_lad = (int*)(d_beta + nb);
    float* d_out = (float*)(base + 2 * al(in_bytes) + al(small));
    int rc = EXM_OK;
    auto hip_ok = [&](hipError_t e, const char* what) {
        if (rc == EXM_OK && e != hipSuccess) {
            set_error("table sum over the orders: %s failed (%s)", what, hipGetErrorString(e));
            rc = EXM_ERR_HIP;
        }
    };
    hip_ok(hipMemcpyAsync(d_col, col, in_bytes, hipMemcpyHostToDevice, c.stream), "upload");
    hip_ok(hipMemcpyAsync(d_row, row, in_bytes, hipMemcpyHostToDevice, c.stream), "upload");
    hip_ok(hipMemcpyAsync(d_lnq, lnq, sizeof(double) * nq, hipMemcpyHostToDevice, c.stream), "upload");
    hip_ok(hipMemcpyAsync(d_beta, beta, sizeof(double) * nb, hipMemcpyHostToDevice, c.stream), "upload");
    hip_ok(hipMemcpyAsync(d_lad, ladder, sizeof(int) * nlad, hipMemcpyHostToDevice, c.stream), "upload");
    const size_t lds = sizeof(double) * (size_t)nord * BD_DT;
    if (rc == EXM_OK) {
        hip_ok(hipFuncSetAttribute((const void*)k_bd_sum_orders, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
               "set LDS");
        ProfileScope prof("tables");
        hipLaunchKernelGGL(k_bd_sum_orders, dim3((unsigned)((nd + BD_DT - 1) / BD_DT), (unsigned)(ntau * nphif)),
                           dim3(BD_THREADS), lds, c.stream, d_col, d_row, d_lnq, d_beta, nord, ntau, nq, nphif, nb,
                           d_lad, nlad, d_out);
        hip_ok(hipGetLastError(), "launch");
    }
    for (int l = 0; l < nlad && rc == EXM_OK; ++l)
        hip_ok(hipMemcpyAsync(out[l], d_out + (size_t)l * (tab_bytes / sizeof(float)), tab_bytes, hipMemcpyDeviceToHost,
                              c.stream),
               "download");
    hip_ok(hipStreamSynchronize(c.stream), "kernel");
    dev_free(buf, total);
    return rc;
}

}  // namespace exm
