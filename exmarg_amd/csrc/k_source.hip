// Node values of one segment -> ln Lambda on the source-frame grid (m1, z) and its prior average ln B.
//
// Restates src/exmarg/search/pe.py:87-142 (_to_source_frame, _log_mean) for the device, float64
// throughout:
//   surf[iz, m]  linear interpolation in y of template m's finite node values (y > 0, finite lamN)
//                at yq[iz] = D0 / D_L(z[iz]); -inf outside [first, last] point or with < 2 points
//   out[i, iz]   linear interpolation in M_det across the templates finite at iz (M ascending) at
//                Md = 2 m1[i] (1 + z[iz]); -inf outside [first, last] finite mass or with < 2
//   ln_b         ln sum w e^out over the finite cells with w > 0 (-inf when there is none)
// Between neighbouring points the value is v0 + t (v1 - v0), t = (x - x0) / (x1 - x0) (t = 1 in y and
// t = 0 in M_det for coincident points), as the reference's vectorized form computes it.

#include "exm_common.h"

namespace exm {

namespace {

constexpr int SRC_ROWS = 64;        // node rows per template (53 nodes + the boundary row fit)
constexpr int SURF_THREADS = 64;
constexpr int OUT_THREADS = 256;
constexpr int LM_THREADS = 256;

// one wave per template: its finite points sorted by y (rank sort), then every yq
__global__ __launch_bounds__(SURF_THREADS) void k_surface(const double* __restrict__ lamN, const double* __restrict__ yN,
                                                          int nrow, int nm, const double* __restrict__ yq, int nz,
                                                          double* __restrict__ surf) {
    __shared__ double ys[SRC_ROWS], ls[SRC_ROWS];
    __shared__ int cnt;
    const int m = blockIdx.x, t = threadIdx.x;
    double y = INFINITY, l = 0.0;
    bool ok = false;
    if (t < nrow) {
        y = yN[(int64_t)t * nm + m];
        l = lamN[(int64_t)t * nm + m];
        ok = isfinite(l) && y > 0.0;
    }
    // rank among the finite points: smaller y first, ties by row (a stable sort)
    const unsigned long long okmask = __ballot(ok);
    int rank = 0;
    for (int j = 0; j < nrow; ++j) {
        const double yj = __shfl(y, j, 64);
        if (((okmask >> j) & 1ull) && (yj < y || (yj == y && j < t))) ++rank;
    }
    if (ok) {
        ys[rank] = y;
        ls[rank] = l;
    }
    if (t == 0) cnt = __popcll(okmask);
    __syncthreads();
    const int n = cnt;
    for (int iz = t; iz < nz; iz += SURF_THREADS) {
        const double x = yq[iz];
        double v = -INFINITY;
        if (n >= 2 && x >= ys[0] && x <= ys[n - 1]) {
            // j = last point <= x
            int lo = 0, hi = n - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (ys[mid] <= x) lo = mid; else hi = mid - 1;
            }
            const int a = min(lo, n - 2), b = a + 1;
            const double y0 = ys[a], y1 = ys[b];
            const double tt = y1 > y0 ? (x - y0) / (y1 - y0) : 1.0;
            v = ls[a] + tt * (ls[b] - ls[a]);
        }
        surf[(int64_t)iz * nm + m] = v;
    }
}

// one workgroup per z: the finite templates of surf[iz] in mass order, then every m1
__global__ __launch_bounds__(OUT_THREADS) void k_mass_interp(const double* __restrict__ surf, int nm,
                                                             const double* __restrict__ M, const double* __restrict__ m1,
                                                             int nm1, const double* __restrict__ z, int nz,
                                                             double* __restrict__ out) {
    extern __shared__ int good[];     // indices of the finite templates, ascending
    __shared__ int part[OUT_THREADS];
    const int iz = blockIdx.x, t = threadIdx.x;
    const double* row = surf + (int64_t)iz * nm;
    const int per = (nm + OUT_THREADS - 1) / OUT_THREADS;
    const int a = min(t * per, nm), b = min(a + per, nm);
    int c = 0;
    for (int m = a; m < b; ++m) c += isfinite(row[m]) ? 1 : 0;
    part[t] = c;
    __syncthreads();
    // exclusive scan of the per-thread counts (one thread: 256 entries)
    if (t == 0) {
        int s = 0;
        for (int k = 0; k < OUT_THREADS; ++k) {
            const int v = part[k];
            part[k] = s;
            s += v;
        }
    }
    __syncthreads();
    int o = part[t];
    for (int m = a; m < b; ++m)
        if (isfinite(row[m])) good[o++] = m;
    __shared__ int ngood_s;
    if (t == OUT_THREADS - 1) ngood_s = o;
    __syncthreads();
    const int ng = ngood_s;
    const double zf = 1.0 + z[iz];
    for (int i = t; i < nm1; i += OUT_THREADS) {
        const double Md = 2.0 * m1[i] * zf;
        double v = -INFINITY;
        if (ng >= 2 && Md >= M[good[0]] && Md <= M[good[ng - 1]]) {
            // below = finite masses with M < Md
            int lo = 0, hi = ng;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (M[good[mid]] < Md) lo = mid + 1; else hi = mid;
            }
            const int below = lo;
            const int ih = good[min(below, ng - 1)];
            const int il = below >= 1 ? good[below - 1] : ih;
            const double m0 = M[il], mm1 = M[ih];
            const double tt = mm1 > m0 ? (Md - m0) / (mm1 - m0) : 0.0;
            v = row[il] + tt * (row[ih] - row[il]);
        }
        out[(int64_t)i * nz + iz] = v;
    }
}

// ln sum w e^L over the finite cells with w > 0, in two stages.  Stage 1: workgroup b takes the cells
// [b per, (b + 1) per) and writes its (max, sum w e^(L - max)) to part[2 b], part[2 b + 1] ((-inf, 0) when
// none of its cells counts).  Stage 2 (one workgroup): M = the largest partial max, ln_b = M + ln sum_b
// part_sum_b e^(part_max_b - M), or -inf when no cell counts.
__device__ __forceinline__ double block_reduce(double v, bool is_max, double* red) {
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    for (int d = 32; d > 0; d >>= 1) {
        const double o = __shfl_xor(v, d, 64);
        v = is_max ? fmax(v, o) : v + o;
    }
    __syncthreads();                    // red may still be read by the previous reduction
    if (lane == 0) red[wid] = v;
    __syncthreads();
    v = red[0];
    for (int k = 1; k < LM_THREADS / 64; ++k) v = is_max ? fmax(v, red[k]) : v + red[k];
    return v;
}

__global__ __launch_bounds__(LM_THREADS) void k_log_mean_part(const double* __restrict__ L,
                                                              const double* __restrict__ w, int64_t n, int64_t per,
                                                              double* __restrict__ part) {
    __shared__ double red[LM_THREADS / 64];
    const int64_t a = (int64_t)blockIdx.x * per, b = min(n, a + per);
    double mx = -INFINITY;
    for (int64_t i = a + threadIdx.x; i < b; i += LM_THREADS)
        if (w[i] > 0.0 && isfinite(L[i])) mx = fmax(mx, L[i]);
    mx = block_reduce(mx, true, red);
    double s = 0.0;
    if (isfinite(mx))
        for (int64_t i = a + threadIdx.x; i < b; i += LM_THREADS)
            if (w[i] > 0.0 && isfinite(L[i])) s += w[i] * exp(L[i] - mx);
    s = block_reduce(s, false, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = mx;
        part[2 * blockIdx.x + 1] = s;
    }
}

__global__ __launch_bounds__(LM_THREADS) void k_log_mean(const double* __restrict__ part, int nblk,
                                                         double* __restrict__ ln_b) {
    __shared__ double red[LM_THREADS / 64];
    double mx = -INFINITY;
    for (int k = threadIdx.x; k < nblk; k += LM_THREADS) mx = fmax(mx, part[2 * k]);
    mx = block_reduce(mx, true, red);
    if (!isfinite(mx)) {
        if (threadIdx.x == 0) *ln_b = -INFINITY;
        return;
    }
    double s = 0.0;
    for (int k = threadIdx.x; k < nblk; k += LM_THREADS)
        if (isfinite(part[2 * k])) s += part[2 * k + 1] * exp(part[2 * k] - mx);
    s = block_reduce(s, false, red);
    if (threadIdx.x == 0) *ln_b = mx + log(s);
}

}  // namespace

int source_frame_limits(int nrow, int nm) {
    EXM_CHECK(nrow >= 1 && nrow <= SRC_ROWS, EXM_ERR_ARG, "%d node rows exceed %d", nrow, SRC_ROWS);
    EXM_CHECK(nm >= 1 && (size_t)nm * sizeof(int) <= 128 * 1024, EXM_ERR_ARG, "%d templates exceed the LDS index", nm);
    return EXM_OK;
}

int launch_source_frame(const double* lamN, const double* yN, int nrow, int nm, const double* M, const double* yq,
                        const double* m1, int nm1, const double* z, int nz, const double* w, double* surf,
                        double* out, double* ln_b, double* part) {
    EXM_TRY(source_frame_limits(nrow, nm));
    const hipStream_t s = cur_stream();
    hipLaunchKernelGGL(k_surface, dim3(nm), dim3(SURF_THREADS), 0, s, lamN, yN, nrow, nm, yq, nz, surf);
    EXM_HIP(hipGetLastError());
    const size_t lds = sizeof(int) * (size_t)nm;
    EXM_HIP(hipFuncSetAttribute((const void*)k_mass_interp, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_mass_interp, dim3(nz), dim3(OUT_THREADS), lds, s, surf, nm, M, m1, nm1, z, nz, out);
    EXM_HIP(hipGetLastError());
    if (w) {
        const int64_t n = (int64_t)nm1 * nz;
        const int64_t per = (n + LOG_MEAN_BLOCKS - 1) / LOG_MEAN_BLOCKS;       // n >= 1, so per >= 1
        const int nblk = (int)((n + per - 1) / per);
        hipLaunchKernelGGL(k_log_mean_part, dim3(nblk), dim3(LM_THREADS), 0, s, out, w, n, per, part);
        EXM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_log_mean, dim3(1), dim3(LM_THREADS), 0, s, part, nblk, ln_b);
        EXM_HIP(hipGetLastError());
    }
    return EXM_OK;
}

}  // namespace exm
