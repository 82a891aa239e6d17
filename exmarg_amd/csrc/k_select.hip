// Arrival-time selection of every template (src/exmarg/search/gather.py:374-420).
//
// One workgroup of 512 threads per template (about 52 KB of LDS at K = 1024, so that three workgroups
// share a CU and one's barriers overlap the others' work):
//   A  score(t) = e^{2 lqL(t)} + e^{2 max_{lag_lo <= tau <= lag_hi} lqH(t + tau)} (cyclic in t), in
//      float64 by pexp, stored as its IEEE bits (order-preserving for positive values).  The window
//      maximum comes from a sparse table over an LDS tile of lqH with its lag halo: maxima of 2^k
//      consecutive values by k doubling steps, then two overlapping blocks cover the window.  Every
//      stride-th key is kept as a sample.
//   B  the K largest keys: the samples are sorted in LDS and a threshold taken below the expected
//      K-th key; one pass collects every key at or above it (at most CAND of them) and an LDS bitonic
//      sort by (key descending, time ascending) gives the K winners exactly (ties to the earlier
//      time).  From 262144 times on (one sample in 128 keys or fewer) the same pass counts the keys at
//      or above four alternate thresholds, and when the threshold catches fewer than K or more than
//      CAND keys a second pass collects at the highest alternate that holds K to CAND keys.  Failing
//      that, an exact radix select (11-bit digits, LDS histogram) and an ordered compaction pass give
//      the same winners;
//   C  the winners in increasing t;
//   D  the drawn times rest[pick] of the remaining times in time order,
//      rest[p] = p + #{i : top[i] - i <= p}, sorted by a bitonic sort and merged with the winners.
#include "exm_common.h"

namespace exm {

constexpr int SEL_THREADS = 512;
constexpr int SEL_WAVES = SEL_THREADS / 64;
constexpr int RADIX = 2048;
constexpr int RADIX_PER_THREAD = RADIX / SEL_THREADS;
constexpr int NSAMP = 2048;          // at most this many samples of the keys
constexpr int CAND = 4096;           // candidates sorted in LDS
// with sparse samples, the candidate threshold's expected count in units of K, and its alternates'
// (increasing)
constexpr double THR_F = 2.2;
constexpr int NALT = 4;
__device__ constexpr double ALT_F[NALT] = {1.3, 1.7, 2.8, 3.5};
// The candidate arrays (CAND keys, then CAND times) also hold, while they are not in use, the sparse-
// table tile of phase A (after the samples), the radix histogram and the drawn times of phase D.
constexpr int TILE_FLOATS = ((CAND - NSAMP) * 8 + CAND * 4) / 4;   // room for the two tile halves

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

__device__ __forceinline__ void bitonic_desc_pairs(uint64_t* k, int32_t* t, int n) {
    // sort (key, time) pairs: key descending, then time ascending; n a power of two
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < n; i += SEL_THREADS) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool first = (i & size) == 0;           // this half in the requested order
                    const uint64_t ka = k[i], kb = k[j];
                    const int32_t ta = t[i], tb = t[j];
                    const bool a_before_b = ka > kb || (ka == kb && ta < tb);
                    if (a_before_b != first) {
                        k[i] = kb;
                        k[j] = ka;
                        t[i] = tb;
                        t[j] = ta;
                    }
                }
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ void bitonic_asc_int(int32_t* a, int n) {
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < n; i += SEL_THREADS) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const int32_t x = a[i], y = a[j];
                    if ((x > y) == up) {
                        a[i] = y;
                        a[j] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(SEL_THREADS) void k_select(const float* __restrict__ lqH, const float* __restrict__ lqL,
                                                        int64_t nt, int lag_lo, int lag_hi, int K, int Kp, int S,
                                                        int Sp, const int64_t* __restrict__ pick, float lw_rand,
                                                        int m0, uint64_t* __restrict__ keys,
                                                        int32_t* __restrict__ sel, float* __restrict__ lw) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int span = lag_hi - lag_lo;
    const int tlen = TILE_FLOATS / 2;                             // one tile half: times plus the lag halo
    const int sel_tile = tlen - span - 1;                         // times per tile
    uint64_t* ckey = (uint64_t*)smem;                             // CAND (samples first use NSAMP of it)
    int32_t* ctime = (int32_t*)(ckey + CAND);                     // CAND
    float* tile = (float*)(ckey + NSAMP);                         // 2 tlen, phase A only
    uint32_t* hist = (uint32_t*)ckey;                             // RADIX, exact radix select only
    int32_t* picked = (int32_t*)ckey;                             // Sp, phase D only
    int32_t* top = ctime + CAND;                                  // Kp
    int* wsum = top + Kp;                                         // SEL_WAVES
    int* misc = wsum + SEL_WAVES;                                 // 4 + NALT

    const int tid = threadIdx.x;
    const int m = m0 + blockIdx.x;
    const float* qH = lqH + (int64_t)m * nt;
    const float* qL = lqL + (int64_t)m * nt;
    uint64_t* key = keys + (int64_t)blockIdx.x * nt;
    // window of W = span + 1 lags: blocks of 2^kw and the offset of the second block
    const int W = span + 1;
    int kw = 0;
    while ((2 << kw) <= W) ++kw;
    const int off2 = W - (1 << kw);
    // one sample every stride = 2^ssh times, the least power of two with NSAMP stride >= nt (a shift and
    // a mask per key instead of a 64-bit division)
    int ssh = 0;
    while (((int64_t)NSAMP << ssh) < nt) ++ssh;
    const int64_t stride = (int64_t)1 << ssh;
    const int nsamp = (int)((nt + stride - 1) >> ssh);            // <= NSAMP
    // lqH is read cyclically at t + tau: with |lags| < nt one conditional add or subtract wraps every
    // index the tiles read (t0 + lag_lo + i lies in [lag_lo, nt + lag_hi)); a 64-bit remainder otherwise
    const bool wrap_once = lag_lo > -nt && lag_hi < nt;

    // ---------------- A: keys and samples.  The lqH tile (with its lag halo) of the next tile is
    // requested into registers while this one is processed, and this tile's lqL before its doubling
    // steps, so that the loads' latency overlaps work.
    constexpr int QH_PT = (TILE_FLOATS / 2 + SEL_THREADS - 1) / SEL_THREADS;   // halo tile values per thread
    constexpr int QL_PT = QH_PT;                                               // tile times per thread
    float qh_next[QH_PT];
    auto load_qh = [&](int64_t t0, int nload) {
#pragma unroll
        for (int k = 0; k < QH_PT; ++k) {
            const int i = tid + k * SEL_THREADS;
            int64_t src = t0 + lag_lo + i;
            if (wrap_once) {
                src = src < 0 ? src + nt : (src >= nt ? src - nt : src);
            } else {
                src %= nt;
                if (src < 0) src += nt;
            }
            qh_next[k] = i < nload ? qH[src] : 0.0f;
        }
    };
    load_qh(0, (int)((nt < sel_tile ? nt : sel_tile) + span));
    for (int64_t t0 = 0; t0 < nt; t0 += sel_tile) {
        const int nthis = (int)((nt - t0) < sel_tile ? (nt - t0) : sel_tile);
        float* src_buf = tile;
#pragma unroll
        for (int k = 0; k < QH_PT; ++k) {
            const int i = tid + k * SEL_THREADS;
            if (i < nthis + span) src_buf[i] = qh_next[k];
        }
        float ql[QL_PT];
#pragma unroll
        for (int k = 0; k < QL_PT; ++k) {
            const int i = tid + k * SEL_THREADS;
            ql[k] = i < nthis ? qL[t0 + i] : 0.0f;
        }
        __syncthreads();
        const int64_t t1 = t0 + sel_tile;
        if (t1 < nt) load_qh(t1, (int)((nt - t1) < sel_tile ? (nt - t1) : sel_tile) + span);
        // doubling: after step j, src_buf[i] = max of 2^j values starting at i
        int len = nthis + span;
        for (int j = 0; j < kw; ++j) {
            float* dst = src_buf == tile ? tile + tlen : tile;
            const int h = 1 << j;
            const int nlen = len - h;
            for (int i = tid; i < nlen; i += SEL_THREADS) dst[i] = fmaxf(src_buf[i], src_buf[i + h]);
            __syncthreads();
            src_buf = dst;
            len = nlen;
        }
#pragma unroll
        for (int k = 0; k < QL_PT; ++k) {
            const int i = tid + k * SEL_THREADS;
            if (i < nthis) {
                const float mx = fmaxf(src_buf[i], src_buf[i + off2]);
                const int64_t t = t0 + i;
                const double sc = __dadd_rn(pexp(2.0 * (double)ql[k]), pexp(2.0 * (double)mx));
                const uint64_t kk = (uint64_t)__double_as_longlong(sc);
                key[t] = kk;
                if ((t & (stride - 1)) == 0) ckey[t >> ssh] = kk;
            }
        }
        __syncthreads();
    }

    // ---------------- B: the threshold from the samples, the candidates, the exact K largest
    int np2 = 1;
    while (np2 < nsamp) np2 <<= 1;
    for (int i = nsamp + tid; i < np2; i += SEL_THREADS) ckey[i] = 0;
    for (int i = tid; i < np2; i += SEL_THREADS) ctime[i] = i;
    __syncthreads();
    bitonic_desc_pairs(ckey, ctime, np2);
    // expected count at or above sample r is (r + 1) stride.  With a sample in at most 64 keys, aim for
    // 1.5 K plus eight samples (the count stays within K .. 2048, one bitonic sort of 2048 at short
    // lengths); with sparser samples (one in 128 or more keys from 262144 times on, where the window
    // maximum also repeats over neighbouring times) the count strays further, so aim for 2.2 K, midway
    // between K and CAND, and count the keys at or above alternate thresholds at ALT_F[j] K (decreasing
    // thresholds) in the same pass: one more pass at the highest alternate holding K to CAND keys costs
    // far less than the exact radix select
    const bool sparse = stride >= 128;
    const int64_t rr = sparse ? (int64_t)(THR_F * K / (double)stride + 0.5) - 1
                              : (int64_t)(1.5 * K / (double)stride) + 8;
    const uint64_t thr0 = rr < nsamp ? ckey[rr < 0 ? 0 : rr] : 0;
    uint64_t thr_alt[NALT];
#pragma unroll
    for (int j = 0; j < NALT; ++j) {
        const int64_t r = (int64_t)(ALT_F[j] * K / (double)stride + 0.5) - 1;
        thr_alt[j] = r < nsamp ? ckey[r < 0 ? 0 : r] : 0;
    }
    if (tid == 0) {
        misc[2] = 0;
        for (int j = 0; j < NALT; ++j) misc[4 + j] = 0;
    }
    __syncthreads();
    // the keys at or above thr into the candidate arrays (their count into misc[2]); with count_alt, the
    // number of keys at or above each alternate into misc[4 ..]; eight keys per thread are requested
    // before any is tested
    auto collect = [&](uint64_t thr, bool count_alt) {
        constexpr int KB = 8;
        int nalt[NALT] = {};
        for (int64_t t0 = 0; t0 < nt; t0 += KB * SEL_THREADS) {
            uint64_t kb[KB];
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                const int64_t t = t0 + j * SEL_THREADS + tid;
                kb[j] = t < nt ? key[t] : 0;
            }
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                const int64_t t = t0 + j * SEL_THREADS + tid;
                if (count_alt) {
#pragma unroll
                    for (int a = 0; a < NALT; ++a) nalt[a] += __popcll(__ballot(t < nt && kb[j] >= thr_alt[a]));
                }
                if (t < nt && kb[j] >= thr) {
                    const int pos = atomicAdd(&misc[2], 1);
                    if (pos < CAND) {
                        ckey[pos] = kb[j];
                        ctime[pos] = (int32_t)t;
                    }
                }
            }
        }
        if (count_alt && (tid & 63) == 0)
            for (int a = 0; a < NALT; ++a) atomicAdd(&misc[4 + a], nalt[a]);
    };
    collect(thr0, sparse);
    __syncthreads();
    int ncand = misc[2];
    if (sparse && (ncand < K || ncand > CAND)) {
        int best = -1;
        for (int a = 0; a < NALT && best < 0; ++a)
            if (misc[4 + a] >= K && misc[4 + a] <= CAND) best = a;
        if (best >= 0) {
            __syncthreads();
            if (tid == 0) misc[2] = 0;
            __syncthreads();
            collect(thr_alt[best], false);
            __syncthreads();
            ncand = misc[2];
        }
    }
    __syncthreads();
    if (ncand >= K && ncand <= CAND) {
        int nc2 = 1;
        while (nc2 < ncand) nc2 <<= 1;
        for (int i = ncand + tid; i < nc2; i += SEL_THREADS) {
            ckey[i] = 0;
            ctime[i] = INT32_MAX;
        }
        __syncthreads();
        bitonic_desc_pairs(ckey, ctime, nc2);
        for (int i = tid; i < Kp; i += SEL_THREADS) top[i] = i < K ? ctime[i] : INT32_MAX;
        __syncthreads();
        bitonic_asc_int(top, Kp);
    } else {
        // exact radix select of the K-th largest key
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
            // bins in descending order: thread i holds bins RADIX-1-RPT i down to RADIX-RPT-RPT i
            int h[RADIX_PER_THREAD], hsum = 0;
#pragma unroll
            for (int k = 0; k < RADIX_PER_THREAD; ++k) {
                h[k] = (int)hist[RADIX - 1 - RADIX_PER_THREAD * tid - k];
                hsum += h[k];
            }
            int total;
            int run = block_excl_scan(hsum, wsum, &total);
#pragma unroll
            for (int k = 0; k < RADIX_PER_THREAD; ++k) {
                if (run < krem && krem <= run + h[k]) {
                    misc[0] = RADIX - 1 - RADIX_PER_THREAD * tid - k;
                    misc[1] = krem - run;
                }
                run += h[k];
            }
            __syncthreads();
            prefix |= (uint64_t)misc[0] << shift;
            pmask |= dmask << shift;
            krem = misc[1];
            __syncthreads();
        }
        const uint64_t thr = prefix;
        const int need = krem;
        // ordered compaction of the winners (keys above the K-th, then keys equal to it by time)
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
    }

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
    int Kp = 1;
    while (Kp < K) Kp <<= 1;
    const int span = lag_hi - lag_lo;
    EXM_CHECK(span + 1 <= TILE_FLOATS / 4, EXM_ERR_ARG, "selection window of %d lags too wide", span + 1);
    EXM_CHECK(Sp <= 2 * CAND && RADIX <= 2 * CAND, EXM_ERR_ARG, "selection of %d drawn times too large", S);
    const size_t lds = (sizeof(uint64_t) + sizeof(int32_t)) * CAND + sizeof(int32_t) * Kp + sizeof(int) * (SEL_WAVES + 4 + NALT);
    EXM_CHECK(lds <= 160 * 1024, EXM_ERR_ARG, "selection needs %zu bytes of LDS", lds);
    // templates per launch: the keys of up to SEL_KEY_BYTES (4096 templates of 524288 times), so that a
    // bank of a few thousand templates is one launch of several workgroups per CU slot and no launch ends
    // on a mostly idle last wave (at 131072 times, 1 GiB of keys in two launches of 1024 and 976 templates
    // took 2.79 ms against 2.39 in one)
    const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(traj->n_tmpl, SEL_KEY_BYTES / (8 * nt)));
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
                           lag_lo, lag_hi, K, Kp, S, Sp, dpick, lw_rand, m0, keys, times->d_sel, times->d_lw);
        EXM_HIP(hipGetLastError());
    }
    return EXM_OK;
}

}  // namespace exm
