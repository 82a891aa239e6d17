// The table read of Eq. (C7) (src/exmarg/search/gather.py:1-24): for every job (template m,
// table k, ln a) and lag tau, the running (max, sum) of exp(lnS) over the walked arrival times,
// then the log-sum-exp over the lags, minus ln n_time (gather.py:103-107).
//
// k_gather_ref evaluates one lookup exactly as the reference CPU kernel does
// (src/exmarg/search/gather.py:45-98): float64 arithmetic on float32 inputs, phase difference
// formed in float32, every operation separately rounded.  One thread per (job, lag), the jobs of a
// lag adjacent so a wavefront shares one lag slab of the tables.
#include <cmath>

#include "exm_common.h"

#pragma clang fp contract(off)

namespace exm {

__global__ __launch_bounds__(256) void k_gather_ref(
    const float* __restrict__ lqH, const float* __restrict__ phH, const float* __restrict__ lqL,
    const float* __restrict__ phL, int64_t n_time, const float* __restrict__ T, int64_t tstride, int nphif, int nb,
    int nd, const int32_t* __restrict__ tau, int ntau, const int32_t* __restrict__ jm,
    const int32_t* __restrict__ jtab, const double* __restrict__ joffb, int64_t njobs, double rdb, double offd,
    double rdd, const int32_t* __restrict__ sel, const float* __restrict__ lw, int n_walk_sel,
    double* __restrict__ om, double* __restrict__ os) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (e >= njobs) return;
    const int m = jm[e];
    const double offb = joffb[e];
    const float* Tn = T + (int64_t)jtab[e] * tstride + (int64_t)r * nb * nphif * nd;
    const int64_t base = (int64_t)m * n_time;
    int64_t tau0 = ((int64_t)tau[r] % n_time + n_time) % n_time;
    const double rph = NPHI / 6.283185307179586;
    const int64_t n_walk = sel ? n_walk_sel : n_time;
    const int64_t row = (int64_t)nphif * nd;   // floats per beta row
    double mx = -1e300, s = 0.0;
    for (int64_t k = 0; k < n_walk; ++k) {
        const int64_t t = sel ? (int64_t)sel[(int64_t)m * n_walk_sel + k] : k;
        int64_t i1 = t + tau0;
        if (i1 >= n_time) i1 -= n_time;
        const float dph = phH[base + i1] - phL[base + t];
        double x3 = (double)dph * rph;
        if (x3 < 0.0) x3 += NPHI;
        if (x3 >= NPHI) x3 -= NPHI;
        int c0 = (int)x3;
        if (c0 >= NPHI) c0 = NPHI - 1;
        else if (c0 < 0) c0 = 0;
        const double fc = x3 - c0;
        const int c1 = (c0 + 1) % NPHI;
        const double aH = lqH[base + i1], aL = lqL[base + t];
        const double xbr = aL * rdb + offb;
        double xb = xbr;
        if (xb < 0.0) xb = 0.0;
        else if (xb > nb - 1 - 1e-6) xb = nb - 1 - 1e-6;
        const double aLc = aL + (xb - xbr) / rdb;
        double xd = (aH - aLc) * rdd + offd;
        if (xd < 0.0) xd = 0.0;
        else if (xd > nd - 1 - 1e-6) xd = nd - 1 - 1e-6;
        int d0 = (int)xd;
        d0 = d0 < 0 ? 0 : (d0 > nd - 1 ? nd - 1 : d0);   // NaN inputs stay inside the table
        const double fdf = xd - d0, fd0 = 1.0 - fdf;
        const int d1 = d0 + 1 < nd ? d0 + 1 : nd - 1;
        int b0 = (int)xb;
        b0 = b0 < 0 ? 0 : (b0 > nb - 1 ? nb - 1 : b0);
        const double fbf = xb - b0, fb0 = 1.0 - fbf;
        const int b1 = b0 + 1 < nb ? b0 + 1 : nb - 1;
        double acc = 0.0;
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
            int ci = cc ? c1 : c0;
            if (ci > NPHI / 2) ci = NPHI - ci;
            const double fci = cc ? fc : 1.0 - fc;
            // device rows are Delta-major: [beta][Delta][phase bin]
            const float* r0 = Tn + b0 * row + ci;
            const float* r1 = Tn + b1 * row + ci;
            const double v = fb0 * (fd0 * (double)r0[d0 * nphif] + fdf * (double)r0[d1 * nphif]) +
                             fbf * (fd0 * (double)r1[d0 * nphif] + fdf * (double)r1[d1 * nphif]);
            acc += fci * v;
        }
        if (sel) acc += (double)lw[(int64_t)m * n_walk_sel + k];
        if (acc > mx) {
            s = s * exp(mx - acc) + 1.0;
            mx = acc;
        } else {
            s += exp(acc - mx);
        }
    }
    om[e * ntau + r] = mx;
    os[e * ntau + r] = s;
}

// ln Lambda[e] = max_r om + ln sum_r os e^{om - max} - ln n_time.  One wave per read: the lanes take the
// lags (a read's om and os rows are contiguous, so the loads coalesce), the max and the sum are reduced
// over the wave by lane shuffles (float64; the sum's order is the reduction tree's)
__global__ __launch_bounds__(256) void k_lse_lags(const double* __restrict__ om, const double* __restrict__ os,
                                                  int64_t njobs, int ntau, const int32_t* __restrict__ jm,
                                                  const double* __restrict__ ln_nt_g, double ln_nt,
                                                  double* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (e >= njobs) return;                             // the whole wave leaves together
    const double* a = om + e * ntau;
    const double* b = os + e * ntau;
    double mx = a[0];
    for (int r = lane; r < ntau; r += 64) mx = a[r] > mx ? a[r] : mx;
#pragma unroll
    for (int k = 32; k > 0; k >>= 1) {
        const double o = __shfl_xor(mx, k, 64);
        mx = o > mx ? o : mx;
    }
    double tot = 0.0;
    for (int r = lane; r < ntau; r += 64) tot += b[r] * exp(a[r] - mx);
#pragma unroll
    for (int k = 32; k > 0; k >>= 1) tot += __shfl_xor(tot, k, 64);
    if (lane == 0) {
        const double ln_n_time = ln_nt_g ? ln_nt_g[jm[e]] : ln_nt;
        out[e] = mx + log(tot) - ln_n_time;
    }
}

int launch_gather_ref(Traj* traj, Store* st, const JobSet& js, Times* times, double* om, double* os) {
    Context& c = ctx();
    const double rdb = 2.0 / st->dbeta;
    const double offd = -st->delta0 / st->ddelta;
    const double rdd = 1.0 / st->ddelta;
    dim3 grid((unsigned)((js.njobs + 255) / 256), (unsigned)st->ntau);
    hipLaunchKernelGGL(k_gather_ref, grid, dim3(256), 0, cur_stream(), traj->lqH(), traj->phH(), traj->lqL(),
                       traj->phL(), traj->n_time, st->d_T, st->tstride, st->nphif, st->nb, st->nd, st->d_tau,
                       st->ntau, js.d_m, js.d_tab, js.d_offb, js.njobs, rdb, offd, rdd,
                       times ? times->d_sel : nullptr, times ? times->d_lw : nullptr, times ? times->n_walk : 0, om,
                       os);
    EXM_HIP(hipGetLastError());
    return EXM_OK;
}

int launch_gather(Traj* traj, Store* st, const JobSet& js, Times* times, double* d_out) {
    double* pairs = nullptr;
    EXM_TRY(scratch(0, sizeof(double) * 2 * (size_t)js.njobs * st->ntau, (void**)&pairs));
    double* om = pairs;
    double* os = pairs + (size_t)js.njobs * st->ntau;
    const int mode = gather_mode();
    if (mode != 1) {
        EXM_TRY(launch_gather_band(traj, st, js, times, om, os));
    } else {
        ProfileScope prof("gather");
        EXM_TRY(launch_gather_ref(traj, st, js, times, om, os));
    }
    return launch_lse(js, st->ntau, om, os, nullptr, std::log((double)traj->n_time), d_out);
}

size_t gather_scratch_bytes(const Traj* traj, const Store* st, const JobSet& js, const Times* times) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t need = al(sizeof(double) * 2 * (size_t)js.njobs * st->ntau);          // slot 0: (max, sum) pairs
    if (gather_mode() != 1) need += band_scratch_bytes(traj, st, js, times);
    return need;
}

int launch_lse(const JobSet& js, int ntau, const double* om, const double* os, const double* ln_nt_g, double ln_nt,
               double* d_out) {
    hipLaunchKernelGGL(k_lse_lags, dim3((unsigned)((js.njobs + 3) / 4)), dim3(256), 0, cur_stream(), om, os,
                       js.njobs, ntau, js.d_m, ln_nt_g, ln_nt, d_out);
    EXM_HIP(hipGetLastError());
    return EXM_OK;
}

}  // namespace exm
