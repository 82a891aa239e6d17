// Arrival-time selection of every template (src/exmarg/search/gather.py:374-420).
//
// One workgroup of 1024 threads per template:
//   A  score(t) = e^{2 lqL(t)} + e^{2 max_{lag_lo <= tau <= lag_hi} lqH(t + tau)} (cyclic in t), in
//      float64 by pexp, stored as its IEEE bits (order-preserving for positive values);
//      the window maximum reads an LDS tile of lqH with its lag halo.
//   B  radix select (11-bit digits, LDS histogram) of the K-th largest key;
//   C  ordered compaction: the K winners (keys above the K-th, then keys equal to it by
//      increasing t) in increasing t;
//   D  the drawn times rest[pick] of the remaining times in time order,
//      rest[p] = p + #{i : top[i] - i <= p}, sorted by a bitonic sort and merged with the winners.
#include "exm_common.h"

namespace exm {

constexpr int SEL_THREADS = 1024;
constexpr int SEL_WAVES = SEL_THREADS / 64;
constexpr int SEL_TILE = 4096;
constexpr int RADIX = 2048;

// exclusive scan over the block; *total receives the block sum
__device__ int block_excl_scan(int v, int* wsum, int* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int s = lane < SEL_WAVES ? wsum[lane] : 0;
#pragma unroll
        for (int d = 1; d < SEL_WAVES; d <<= 1) {
            const int y = __shfl_up(s, d, 64);
            if (lane >= d) s += y;
        }
        if (lane < SEL_WAVES) wsum[lane] = s;
    }
    __syncthreads();
    const int base = wid ? wsum[wid - 1] : 0;
    *total = wsum[SEL_WAVES - 1];
    __syncthreads();
    return base + x - v;
}

__device__ __forceinline__ int upper_bound_g(const int32_t* top, int K, int64_t p) {
    // number of i with top[i] - i <= p (top[i] - i is non-decreasing)
    int lo = 0, hi = K;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int64_t)top[mid] - mid <= p) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int lower_bound_i(const int32_t* a, int n, int32_t v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(SEL_THREADS) void k_select(const float* __restrict__ lqH, const float* __restrict__ lqL,
                                                        int64_t nt, int lag_lo, int lag_hi, int K, int S, int Sp,
                                                        const int64_t* __restrict__ pick, float lw_rand, int m0,
                                                        uint64_t* __restrict__ keys, int32_t* __restrict__ sel,
                                                        float* __restrict__ lw) {
    extern __shared__ unsigned char smem[];
    const int span = lag_hi - lag_lo;
    float* tile = (float*)smem;                                   // SEL_TILE + span
    uint32_t* hist = (uint32_t*)(tile + SEL_TILE + span + 1);     // RADIX
    int32_t* top = (int32_t*)(hist + RADIX);                      // K
    int32_t* picked = top + K;                                    // Sp
    int* wsum = picked + Sp;                                      // SEL_WAVES
    int* misc = wsum + SEL_WAVES;                                 // 4

    const int tid = threadIdx.x;
    const int m = m0 + blockIdx.x;
    const float* qH = lqH + (int64_t)m * nt;
    const float* qL = lqL + (int64_t)m * nt;
    uint64_t* key = keys + (int64_t)blockIdx.x * nt;

    // ---------------- A: keys
    for (int64_t t0 = 0; t0 < nt; t0 += SEL_TILE) {
        const int64_t nthis = (nt - t0) < SEL_TILE ? (nt - t0) : SEL_TILE;
        for (int i = tid; i < nthis + span; i += SEL_THREADS) {
            int64_t src = (t0 + lag_lo + i) % nt;
            if (src < 0) src += nt;
            tile[i] = qH[src];
        }
        __syncthreads();
        for (int i = tid; i < nthis; i += SEL_THREADS) {
            float mx = tile[i];
            for (int j = 1; j <= span; ++j) mx = fmaxf(mx, tile[i + j]);
            const double sc = __dadd_rn(pexp(2.0 * (double)qL[t0 + i]), pexp(2.0 * (double)mx));
            key[t0 + i] = (uint64_t)__double_as_longlong(sc);
        }
        __syncthreads();
    }

    // ---------------- B: radix select of the K-th largest key
    uint64_t prefix = 0, pmask = 0;
    int krem = K;
    const int shifts[6] = {53, 42, 31, 20, 9, 0};
    const int widths[6] = {11, 11, 11, 11, 11, 9};
    for (int pass = 0; pass < 6; ++pass) {
        const int shift = shifts[pass];
        const uint64_t dmask = (1ull << widths[pass]) - 1;
        for (int i = tid; i < RADIX; i += SEL_THREADS) hist[i] = 0;
        __syncthreads();
        for (int64_t t = tid; t < nt; t += SEL_THREADS) {
            const uint64_t k = key[t];
            if ((k & pmask) == prefix) atomicAdd(&hist[(k >> shift) & dmask], 1u);
        }
        __syncthreads();
        // bins in descending order: thread i holds bins RADIX-1-2i and RADIX-2-2i
        const int b0 = RADIX - 1 - 2 * tid, b1 = RADIX - 2 - 2 * tid;
        const int h0 = (int)hist[b0], h1 = (int)hist[b1];
        int total;
        const int before = block_excl_scan(h0 + h1, wsum, &total);
        if (before < krem && krem <= before + h0) {
            misc[0] = b0;
            misc[1] = krem - before;
        } else if (before + h0 < krem && krem <= before + h0 + h1) {
            misc[0] = b1;
            misc[1] = krem - before - h0;
        }
        __syncthreads();
        prefix |= (uint64_t)misc[0] << shift;
        pmask |= dmask << shift;
        krem = misc[1];
        __syncthreads();
    }
    const uint64_t thr = prefix;
    const int need = krem;

    // ---------------- C: ordered compaction of the winners
    int eq_base = 0, take_base = 0;
    for (int64_t t0 = 0; t0 < nt; t0 += SEL_THREADS) {
        const int64_t t = t0 + tid;
        const bool valid = t < nt;
        const uint64_t k = valid ? key[t] : 0;
        const int gt = valid && k > thr;
        const int eq = valid && k == thr;
        int tot_eq, tot_take;
        const int ex_eq = block_excl_scan(eq, wsum, &tot_eq);
        const int take = gt || (eq && eq_base + ex_eq < need);
        const int ex_take = block_excl_scan(take, wsum, &tot_take);
        if (take) top[take_base + ex_take] = (int32_t)t;
        eq_base += tot_eq;
        take_base += tot_take;
    }
    __syncthreads();

    // ---------------- D: drawn times, sorted, merged with the winners
    for (int s = tid; s < Sp; s += SEL_THREADS) {
        if (s < S) {
            const int64_t p = pick[(int64_t)m * S + s];
            picked[s] = (int32_t)(p + upper_bound_g(top, K, p));
        } else {
            picked[s] = INT32_MAX;
        }
    }
    __syncthreads();
    for (int size = 2; size <= Sp; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = tid; i < Sp; i += SEL_THREADS) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const int32_t a = picked[i], b = picked[j];
                    if ((a > b) == up) {
                        picked[i] = b;
                        picked[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    const int nw = K + S;
    int32_t* so = sel + (int64_t)m * nw;
    float* lo_ = lw + (int64_t)m * nw;
    for (int i = tid; i < K; i += SEL_THREADS) {
        const int pos = i + lower_bound_i(picked, S, top[i]);
        so[pos] = top[i];
        lo_[pos] = 0.0f;
    }
    for (int j = tid; j < S; j += SEL_THREADS) {
        const int pos = j + lower_bound_i(top, K, picked[j]);
        so[pos] = picked[j];
        lo_[pos] = lw_rand;
    }
}

int launch_select(Traj* traj, int lag_lo, int lag_hi, int K, int S, const int64_t* pick, const int64_t* d_pick,
                  float lw_rand, Times* times) {
    Context& c = ctx();
    const int64_t nt = traj->n_time;
    int Sp = 1;
    while (Sp < S) Sp <<= 1;
    const int span = lag_hi - lag_lo;
    const size_t lds = sizeof(float) * (SEL_TILE + span + 1) + sizeof(uint32_t) * RADIX +
                       sizeof(int32_t) * (K + Sp) + sizeof(int) * (SEL_WAVES + 4);
    EXM_CHECK(lds <= 160 * 1024, EXM_ERR_ARG, "selection needs %zu bytes of LDS", lds);
    const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(traj->n_tmpl, (1ll << 30) / (8 * nt)));
    uint64_t* keys = nullptr;
    EXM_TRY(scratch(0, sizeof(uint64_t) * (size_t)chunk * nt, (void**)&keys));
    // the draw: uploaded from the host (pick) or already on the device (d_pick)
    const int64_t* dpick = d_pick;
    if (!dpick) {
        int64_t* up = nullptr;
        EXM_TRY(scratch(2, sizeof(int64_t) * (size_t)traj->n_tmpl * S, (void**)&up));
        EXM_HIP(hipMemcpyAsync(up, pick, sizeof(int64_t) * (size_t)traj->n_tmpl * S, hipMemcpyHostToDevice,
                               c.stream));
        dpick = up;
    }
    EXM_HIP(hipFuncSetAttribute((const void*)k_select, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    ProfileScope prof("select");
    for (int m0 = 0; m0 < traj->n_tmpl; m0 += chunk) {
        const int nb = std::min(chunk, traj->n_tmpl - m0);
        hipLaunchKernelGGL(k_select, dim3(nb), dim3(SEL_THREADS), lds, c.stream, traj->lqH(), traj->lqL(), nt,
                           lag_lo, lag_hi, K, S, Sp, dpick, lw_rand, m0, keys, times->d_sel, times->d_lw);
        EXM_HIP(hipGetLastError());
    }
    return EXM_OK;
}

}  // namespace exm
