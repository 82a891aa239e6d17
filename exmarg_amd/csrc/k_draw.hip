// The random draw of select_times on the device: numpy's default_rng(seed).integers(0, high, n) for
// int64 output, reproduced exactly from the generator's state.
//
// The reference draws on the host, pick = default_rng(seed + row).integers(0, n_time - K,
// (n_tmpl, S)) (src/exmarg/search/gather.py:397), and copies it to the device.  numpy's Generator
// uses the PCG64 bit generator (128-bit LCG, multiplier 0x2360ED051FC65DA44385DF649FCCF645, XSL-RR
// output: rotr64(hi ^ lo, state >> 122) of the advanced state), hands out 32-bit words as the low
// then the high half of each 64-bit output, and bounds them with Lemire's method: m = x * high,
// accept when (m mod 2^32) >= (2^32 - high) mod high, value m >> 32.  Acceptance depends on each word
// alone, so the stream parallelizes: every thread jumps its LCG ahead (Brown's O(log n) advance) to its
// block of words, counts the accepted ones, an exclusive scan gives each thread its first output, and
// a second pass writes the values.  The caller passes the generator state (state, inc) that numpy
// derives from the seed; nothing else of numpy is needed.

#include "exm_common.h"

namespace exm {

namespace {

typedef unsigned __int128 u128;

constexpr int DRAW_THREADS = 256;
constexpr int WORDS64 = 32;                 // 64-bit outputs per thread (64 candidate 32-bit words)
constexpr u128 PCG_MULT = ((u128)0x2360ED051FC65DA4ull << 64) | (u128)0x4385DF649FCCF645ull;

__device__ __forceinline__ u128 lcg_advance(u128 state, u128 inc, uint64_t delta) {
    u128 acc_mult = 1, acc_plus = 0, cur_mult = PCG_MULT, cur_plus = inc;
    while (delta > 0) {
        if (delta & 1) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1;
    }
    return acc_mult * state + acc_plus;
}

__device__ __forceinline__ uint64_t pcg_output(u128 s) {
    const uint64_t x = (uint64_t)(s >> 64) ^ (uint64_t)s;
    const unsigned rot = (unsigned)(s >> 122);
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

// pass 1 (write == false): accepted words per thread; pass 2: the values at their output positions
template <bool WRITE>
__global__ __launch_bounds__(DRAW_THREADS) void k_draw(uint64_t s_hi, uint64_t s_lo, uint64_t i_hi, uint64_t i_lo,
                                                       int64_t nthreads, uint32_t high, uint32_t threshold,
                                                       int32_t* __restrict__ counts, const int32_t* __restrict__ first,
                                                       int64_t n_out, int64_t* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * DRAW_THREADS + threadIdx.x;
    if (t >= nthreads) return;
    const u128 inc = ((u128)i_hi << 64) | i_lo;
    u128 s = lcg_advance(((u128)s_hi << 64) | s_lo, inc, (uint64_t)t * WORDS64);
    int64_t pos = WRITE ? first[t] : 0;
    int c = 0;
    for (int j = 0; j < WORDS64; ++j) {
        s = s * PCG_MULT + inc;
        const uint64_t v = pcg_output(s);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint64_t m = (uint64_t)(uint32_t)(h ? v >> 32 : v) * high;
            if ((uint32_t)m >= threshold) {
                if (WRITE) {
                    if (pos < n_out) out[pos] = (int64_t)(m >> 32);
                    ++pos;
                } else {
                    ++c;
                }
            }
        }
    }
    if (!WRITE) counts[t] = c;
}

// exclusive scan of the per-thread counts (one workgroup; total to *total)
__global__ __launch_bounds__(1024) void k_draw_scan(const int32_t* __restrict__ counts, int64_t n,
                                                    int32_t* __restrict__ first, int64_t* __restrict__ total) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (n + 1023) / 1024;
    const int64_t a = min((int64_t)t * per, n), b = min(a + per, n);
    int64_t s = 0;
    for (int64_t i = a; i < b; ++i) s += counts[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        int64_t run = 0;
        for (int k = 0; k < 1024; ++k) {
            const int64_t v = part[k];
            part[k] = run;
            run += v;
        }
        *total = run;
    }
    __syncthreads();
    int64_t run = part[t];
    for (int64_t i = a; i < b; ++i) {
        first[i] = (int32_t)run;
        run += counts[i];
    }
}

}  // namespace

int launch_draw(const uint64_t rng_state[4], int64_t n_out, int64_t high, int64_t* d_out) {
    EXM_CHECK(high >= 1 && high <= 0xFFFFFFFFll, EXM_ERR_ARG, "draw range %lld outside [1, 2^32)", (long long)high);
    EXM_CHECK(n_out >= 0 && n_out < (1ll << 31) - (1ll << 24), EXM_ERR_ARG, "draw of %lld values too large",
              (long long)n_out);
    if (n_out == 0) return EXM_OK;
    const uint32_t h = (uint32_t)high;
    const uint32_t threshold = (uint32_t)((0x100000000ull - (uint64_t)h) % (uint64_t)h);
    // words to generate: the expected need at this acceptance rate with a margin; grown if short
    const double acc = (4294967296.0 - (double)threshold) / 4294967296.0;
    int64_t words = (int64_t)((double)n_out / acc * 1.02) + 8192;
    const hipStream_t s = cur_stream();
    for (int attempt = 0; attempt < 8; ++attempt) {
        const int64_t nthreads = (words + 2 * WORDS64 - 1) / (2 * WORDS64);
        void* buf = nullptr;
        const size_t cbytes = (sizeof(int32_t) * (size_t)nthreads + 255) & ~(size_t)255;
        EXM_TRY(scratch(3, 2 * cbytes + 256, &buf));
        int32_t* counts = (int32_t*)buf;
        int32_t* first = (int32_t*)((char*)buf + cbytes);
        int64_t* d_total = (int64_t*)((char*)buf + 2 * cbytes);
        const unsigned grid = (unsigned)((nthreads + DRAW_THREADS - 1) / DRAW_THREADS);
        hipLaunchKernelGGL(k_draw<false>, dim3(grid), dim3(DRAW_THREADS), 0, s, rng_state[0], rng_state[1],
                           rng_state[2], rng_state[3], nthreads, h, threshold, counts, nullptr, n_out, nullptr);
        EXM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_draw_scan, dim3(1), dim3(1024), 0, s, counts, nthreads, first, d_total);
        EXM_HIP(hipGetLastError());
        int64_t total = 0;
        EXM_HIP(hipMemcpyAsync(&total, d_total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        EXM_HIP(hipStreamSynchronize(s));
        if (total >= n_out) {
            hipLaunchKernelGGL(k_draw<true>, dim3(grid), dim3(DRAW_THREADS), 0, s, rng_state[0], rng_state[1],
                               rng_state[2], rng_state[3], nthreads, h, threshold, nullptr, first, n_out, d_out);
            EXM_HIP(hipGetLastError());
            return EXM_OK;
        }
        words *= 2;
    }
    EXM_CHECK(false, EXM_ERR_ARG, "draw of %lld values below 2^32 / %lld did not converge", (long long)n_out,
              (long long)high);
    return EXM_OK;
}

}  // namespace exm
