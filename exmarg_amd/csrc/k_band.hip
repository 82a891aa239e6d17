// The table read for a selection of walked times (src/exmarg/search/gather.py:32-100 for
// times = select_times(...)), organised around the beta rows of the tables.
//
// For every template, beta = ln a + 2 ln|z_L(t)| of a walked time t is increasing in ln|z_L(t)| at
// every node and lag.  So, once per segment:
//   k_walk_sort    sorts each template's walked times by ln|z_L|;
//   k_walk_stream  writes, for every (template, lag) and walked time in that order, one 8-byte record:
//                  the Delta column and fraction (24-bit fixed point) of the lookup with beta unclamped,
//                  which do not depend on the node; the pair of adjacent stored phase bins that
//                  phi_H(t + tau) - phi_L(t) falls between (bins mirrored to [0, 48]; the record holds
//                  the LDS offset of the lower bin in the Delta column) and the weight of the upper bin
//                  (25-bit fixed point); a drawn-time flag.  ln|z_H(t + tau)| goes to a side array that
//                  only the rows where beta may be clamped read;
//   k_read_bounds  finds, for every read (template, table, ln a), where each beta row starts in the
//                  sorted walk.
// The gather (k_gather_wave, or the piece form k_gather_piece) then runs one workgroup per (table, lag,
// tile of at most 1024 reads of that table) over the interior rows 1 .. nb - 3.  It sweeps the beta rows
// in increasing order with a ring of three rows (every phase bin and Delta column, 3 x 43 KB at the
// reference geometry) in LDS; at row b the lookups of the tile whose beta row is b are cut into pieces
// of consecutive walked times of one read, dealt to the lanes in rounds, and each read's pieces are
// merged by the lane that owns the read.  Every table value a lookup reads comes from LDS and every
// table row is read from HBM once per workgroup.  k_gather_edge adds the lookups of rows 0 and nb - 2,
// where beta may be clamped, with the Delta coordinate recomputed from ln|z_L| and ln|z_H| exactly.
//
// Arithmetic (float64 form): the beta fraction, the interpolation and every sum in float64 (the MI355X
// runs float64 FMAs at the float32 rate), exp of non-positive differences in float32; the float32 form
// interpolates in float32.
#include <climits>
#include <cmath>

#include "exm_common.h"

namespace exm {

constexpr int WALK_THREADS = 1024;
constexpr int BAND_THREADS = 1024;        // the piece form's workgroup
constexpr int WAVE_THREADS = 1024;        // the wave form's workgroup (64 reads per wave)
constexpr double FC_SCALE = 16777216.0;   // phase weights in fixed point, units of 2^-24
constexpr uint32_t FC_MASK = 0x1FFFFFFu;  // 25 bits: the weight of the upper bin, 0 .. 2^24
constexpr uint32_t DRAWN = 0x80000000u;   // walked time from the random draw (log weight lw_rand)
constexpr int BIN_SHIFT = 25;             // 6 bits: the lower stored bin of the pair (0..47)
constexpr int PIECE = 8;                  // default walked times per piece

__device__ __forceinline__ uint32_t float_order_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// the float of an order key (the inverse of float_order_key)
__device__ __forceinline__ float order_key_float(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// One block per template: the walk's nw times in increasing ln|z_L| (an LDS bitonic sort of (key, index)
// pairs).  The walk is positions off .. off + nw - 1 of the template's row of the selection (sel, lw with
// row stride sel_stride), or, without a selection, the arrival times off .. off + nw - 1 themselves (log
// weight 0): walks longer than one launch takes are cut into consecutive slices.
__global__ __launch_bounds__(WALK_THREADS) void k_walk_sort(const float* __restrict__ lqL, int64_t nt,
                                                            const int32_t* __restrict__ sel,
                                                            const float* __restrict__ lw, int64_t sel_stride,
                                                            int64_t off, int nw, int nwp,
                                                            float* __restrict__ wa, float* __restrict__ wl,
                                                            int32_t* __restrict__ wt) {
    extern __shared__ uint64_t skeys[];
    const int m = blockIdx.x;
    const int tid = threadIdx.x;
    const int32_t* s = sel ? sel + (int64_t)m * sel_stride + off : nullptr;
    const float* q = lqL + (int64_t)m * nt;
    for (int i = tid; i < nwp; i += WALK_THREADS)
        skeys[i] = i < nw ? (((uint64_t)float_order_key(q[s ? (int64_t)s[i] : off + i]) << 32) | (uint32_t)i) : ~0ull;
    __syncthreads();
    for (int size = 2; size <= nwp; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = tid; i < nwp; i += WALK_THREADS) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const uint64_t a = skeys[i], b = skeys[j];
                    if ((a > b) == up) {
                        skeys[i] = b;
                        skeys[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int p = tid; p < nw; p += WALK_THREADS) {
        const int i = (int)(uint32_t)skeys[p];
        const int32_t t = s ? s[i] : (int32_t)(off + i);
        wa[(int64_t)m * nw + p] = q[t];
        wl[(int64_t)m * nw + p] = s ? lw[(int64_t)m * sel_stride + off + i] : 0.0f;
        wt[(int64_t)m * nw + p] = t;
    }
}

// One block per (template, 32 walked times).  For every lag and walked time it writes the walk record
// (8 bytes; layout below) and ln|z_H(t + tau)| (4 bytes, read only in the rows where beta may be clamped).
// The records are computed with the lanes over the lags, so that the z_H reads of one walked time (at
// t + tau for consecutive lags) are contiguous, kept in LDS (three planes of ntau rows, odd stride so that
// neither access pattern conflicts on the banks) and written with the lanes over the walked times (one
// contiguous run per lag, two lags per wave).  The small tile keeps the LDS per block at 35 KB at 87 lags.
//
// Walk record, bits 0-63 of (x, y): the Delta fraction of the lookup in units of 2^-24 (bits 0-23), the
// LDS offset of its lower (Delta, phase) corner in a table row, d0 * NPHIF + lower stored phase bin
// (bits 24-37), the weight of the upper phase bin in units of 2^-24 (bits 38-62) and the drawn flag (bit
// 63).  Delta = ln|z_H(t + tau)| - ln|z_L(t)| does not depend on the node as long as beta is not clamped,
// which only happens in the first and the last row pair of a table: every other row reads the record as
// it is; those two rows recompute Delta from ln|z_H| (k_gather_edge).
constexpr int WS_TILE = 32, WS_STRIDE = WS_TILE + 1;
constexpr uint32_t FD_MASK = 0xFFFFFFu;           // 24 bits of Delta fraction
constexpr int REC_O_SHIFT = 24;                   // the corner offset starts at bit 24 (of 64)
constexpr uint32_t REC_O_MASK = 0x3FFFu;          // 14 bits: rows of at most 16384 floats
constexpr int REC_W_SHIFT = 6;                    // the phase weight starts at bit 38 = 32 + 6
constexpr uint32_t REC_DRAWN = 0x80000000u;       // bit 63

// LDS of the walk stream block: the block's walked times, their phi_L, ln|z_L| and flags, the lag
// offsets, then the three planes
__host__ __device__ constexpr size_t walk_stream_lds_bytes(int ntau) {
    return 4 * sizeof(uint32_t) * WS_TILE + sizeof(int32_t) * ntau + 3 * sizeof(uint32_t) * ntau * WS_STRIDE;
}

__global__ __launch_bounds__(256) void k_walk_stream(const float* __restrict__ lqH, const float* __restrict__ phH,
                                                     const float* __restrict__ phL, int64_t nt,
                                                     const int32_t* __restrict__ tau, int ntau,
                                                     const int32_t* __restrict__ wt, const float* __restrict__ wl,
                                                     const float* __restrict__ wa, int nw, double rdd, double offd,
                                                     int nd, uint2* __restrict__ stream, float* __restrict__ aHs) {
    extern __shared__ uint32_t wsh[];
    int32_t* wts = (int32_t*)wsh;                       // the block's walked times
    float* pLs = (float*)(wsh + WS_TILE);               // their phi_L
    uint32_t* drs = wsh + 2 * WS_TILE;                  // their drawn flags
    float* aLs = (float*)(wsh + 3 * WS_TILE);           // their ln|z_L|
    int32_t* tau0 = (int32_t*)(wsh + 4 * WS_TILE);      // lag offsets in [0, nt)
    uint32_t* rx = wsh + 4 * WS_TILE + ntau;
    uint32_t* ry = rx + ntau * WS_STRIDE;
    uint32_t* waH = ry + ntau * WS_STRIDE;
    const int m = blockIdx.x, p0 = blockIdx.y * WS_TILE;
    const int np = min(WS_TILE, nw - p0);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t base = (int64_t)m * nt;
    const double rph = NPHI / 6.283185307179586;
    const double xdmax = nd - 1 - 1e-6;
    for (int r = threadIdx.x; r < ntau; r += 256) tau0[r] = (int32_t)(((int64_t)tau[r] % nt + nt) % nt);
    if (threadIdx.x < np) {
        const int p = p0 + threadIdx.x;
        const int32_t t = wt[(int64_t)m * nw + p];
        wts[threadIdx.x] = t;
        pLs[threadIdx.x] = phL[base + t];
        aLs[threadIdx.x] = wa[(int64_t)m * nw + p];
        drs[threadIdx.x] = wl[(int64_t)m * nw + p] != 0.0f ? REC_DRAWN : 0u;
    }
    __syncthreads();
    // per chunk of 64 lags: the wave's walked times' phi_H and ln|z_H| are all requested, then the
    // records computed (the loads' latency overlaps)
    constexpr int PPW = WS_TILE / 4;                    // walked times per wave
    for (int r0 = 0; r0 < ntau; r0 += 64) {
        const int r = r0 + lane;
        const bool rv = r < ntau;
        const int64_t tr = rv ? tau0[r] : 0;
        float phv[PPW], aHv[PPW];
#pragma unroll
        for (int k = 0; k < PPW; ++k) {
            const int pl = wid + 4 * k;
            phv[k] = aHv[k] = 0.0f;
            if (rv && pl < np) {
                int64_t i1 = wts[pl] + tr;
                if (i1 >= nt) i1 -= nt;
                phv[k] = phH[base + i1];
                aHv[k] = lqH[base + i1];
            }
        }
#pragma unroll
        for (int k = 0; k < PPW; ++k) {
            const int pl = wid + 4 * k;
            if (!(rv && pl < np)) continue;
            // phase (gather.py:54-65): the stored bins of c0 and c0 + 1 (mod 96) are adjacent, ci1 = ci0 + 1
            // below the mirror and ci0 - 1 from it on; the record holds the lower bin of the pair and the
            // weight of the upper one (fc, or 1 - fc when c0's stored bin is the upper)
            const float dph = __fsub_rn(phv[k], pLs[pl]);
            double x3 = (double)dph * rph;
            if (x3 < 0.0) x3 += NPHI;
            if (x3 >= NPHI) x3 -= NPHI;
            int c0 = (int)x3;
            c0 = c0 < 0 ? 0 : (c0 > NPHI - 1 ? NPHI - 1 : c0);
            const double fc = x3 - c0;
            uint32_t fq = (uint32_t)__double2uint_rn(fc * FC_SCALE);
            fq = fq > (1u << 24) - 1 ? (1u << 24) - 1 : fq;
            const uint32_t ci0 = (uint32_t)(c0 > NPHI / 2 ? NPHI - c0 : c0);
            const bool down = c0 >= NPHI / 2;
            const uint32_t lo = down ? ci0 - 1 : ci0;
            const uint32_t wq = down ? (1u << 24) - fq : fq;
            // Delta with beta unclamped (gather.py:67-78): the coordinate, its column and its fraction in
            // 24-bit fixed point (at most 2^24 - 1, so that column d0 + 1 stays in the row)
            double xd = ((double)aHv[k] - (double)aLs[pl]) * rdd + offd;
            xd = xd > 0.0 ? (xd < xdmax ? xd : xdmax) : 0.0;               // NaN -> 0
            const int d0 = (int)xd;
            uint32_t fdq = (uint32_t)__double2uint_rn((xd - d0) * FC_SCALE);
            fdq = fdq > FD_MASK ? FD_MASK : fdq;
            const uint32_t o = (uint32_t)d0 * NPHIF + lo;
            rx[r * WS_STRIDE + pl] = fdq | (o << REC_O_SHIFT);
            ry[r * WS_STRIDE + pl] = (o >> (32 - REC_O_SHIFT)) | (wq << REC_W_SHIFT) | drs[pl];
            waH[r * WS_STRIDE + pl] = __float_as_uint(aHv[k]);
        }
    }
    __syncthreads();
    const int pl = lane & (WS_TILE - 1);
    for (int r = 2 * wid + (lane >> 5); r < ntau; r += 8)
        if (pl < np) {
            const int64_t o = ((int64_t)m * ntau + r) * nw + p0 + pl;
            stream[o] = make_uint2(rx[r * WS_STRIDE + pl], ry[r * WS_STRIDE + pl]);
            aHs[o] = __uint_as_float(waH[r * WS_STRIDE + pl]);
        }
}

// DPP lane moves within a wave64 (VALU, no LDS round trip): lane l receives the value of lane l - n of
// its row of 16 lanes (row_shr:n), or of lane 15 / 31 of the rows below it (row_bcast:15 / :31); lanes
// without a source lane, or outside the row and bank masks, receive `old`.  Combining each lane with
// the moved value in the sequence shr 1, shr 2, shr 4 (banks 1-3), shr 8 (banks 2-3), bcast 15 (rows 1,
// 3), bcast 31 (rows 2, 3), with `old` the identity of the combination, leaves at every lane the
// combination over the lanes up to itself (an inclusive scan).
constexpr int DPP_ROW_SHR = 0x110, DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;

template <int CTRL, int ROWM, int BANKM>
__device__ __forceinline__ int dpp_i(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWM, BANKM, false);
}

template <int CTRL, int ROWM, int BANKM>
__device__ __forceinline__ uint32_t dpp_u(uint32_t old, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROWM, BANKM, false);
}

// beta row of one walked time of one read: the clamped coordinate xb, the unclamped xbr and the row
__device__ __forceinline__ int beta_row(float aL, double offb, double rdb, int nb, double& xb, double& xbr) {
    xbr = (double)aL * rdb + offb;
    const double xbmax = nb - 1 - 1e-6;
    xb = xbr < 0.0 ? 0.0 : (xbr > xbmax ? xbmax : xbr);
    const int b0 = (int)xb;
    return b0 < 0 ? 0 : (b0 > nb - 2 ? nb - 2 : b0);   // NaN inputs stay inside the table
}

// per read: first and last beta row of its walk and the start of every row in the sorted walk,
// bounds[j * (nb + 1) + (b - first)] for b in [first, last + 1].  One wave per read, 64 walked times at a
// time: the rows (first..b] not yet started start at the first walked time whose row reaches b, i.e.
// rows in (running maximum before p, running maximum at p] start at p (a maximum scan over the lanes,
// carried from chunk to chunk), the same bounds as walking the times one by one.
__global__ __launch_bounds__(256) void k_read_bounds(const float* __restrict__ wa, int nw, const int32_t* __restrict__ jm,
                                                     const double* __restrict__ joffb, int64_t njobs, double rdb,
                                                     int nb, int32_t* __restrict__ first_last,
                                                     uint16_t* __restrict__ bounds) {
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= njobs) return;                              // a whole wave
    const float* a = wa + (int64_t)jm[j] * nw;
    const double offb = joffb[j];
    double xb, xbr;
    const int first = beta_row(a[0], offb, rdb, nb, xb, xbr);
    int last = beta_row(a[nw - 1], offb, rdb, nb, xb, xbr);
    last = last < first ? first : last;                 // only NaN input breaks the order
    uint16_t* bj = bounds + j * (int64_t)(nb + 1);
    if (lane == 0) {
        first_last[2 * j] = first;
        first_last[2 * j + 1] = last;
        bj[0] = 0;
    }
    int carry = first;                                  // running maximum of the rows before this chunk
    for (int c = 0; c < nw; c += 64) {
        const int p = c + lane;
        int run = p < nw ? beta_row(a[p], offb, rdb, nb, xb, xbr) : INT_MIN;
        // inclusive maximum scan over the lanes by DPP moves (INT_MIN the identity)
        run = max(run, dpp_i<DPP_ROW_SHR | 1, 0xf, 0xf>(INT_MIN, run));
        run = max(run, dpp_i<DPP_ROW_SHR | 2, 0xf, 0xf>(INT_MIN, run));
        run = max(run, dpp_i<DPP_ROW_SHR | 4, 0xf, 0xe>(INT_MIN, run));
        run = max(run, dpp_i<DPP_ROW_SHR | 8, 0xf, 0xc>(INT_MIN, run));
        run = max(run, dpp_i<DPP_ROW_BCAST15, 0xa, 0xf>(INT_MIN, run));
        run = max(run, dpp_i<DPP_ROW_BCAST31, 0xc, 0xf>(INT_MIN, run));
        run = max(run, carry);
        int prev = __shfl_up(run, 1, 64);
        if (lane == 0) prev = carry;
        if (p >= 1 && p < nw)
            for (int rr = prev + 1; rr <= min(run, last); ++rr) bj[rr - first] = (uint16_t)p;
        carry = __builtin_amdgcn_readlane(run, 63);
    }
    if (lane == 0) bj[last + 1 - first] = (uint16_t)nw;
}

// floats of one ring slot: a table row rounded up to whole 64-float blocks (one block per wave copy)
__host__ __device__ constexpr int row_slot(int rowlen) { return (rowlen + 63) & ~63; }

// byte offset of the piece results after the three ring slots, 16-byte aligned
__host__ __device__ constexpr size_t piece_res_offset(int rowlen) {
    return ((size_t)12 * row_slot(rowlen) + 15) & ~(size_t)15;
}

// bytes of LDS the piece kernel uses for a row of rowlen floats
__host__ __device__ constexpr size_t piece_lds_bytes(int rowlen) {
    return piece_res_offset(rowlen) + 2 * sizeof(double) * BAND_THREADS + sizeof(int) * BAND_THREADS +
           3 * sizeof(uint16_t) * BAND_THREADS + sizeof(int) * 20;
}

// copy one table row (slot floats, whole 64-float blocks) from global memory straight into LDS:
// wave w moves blocks w, w + 16, ...; each lane one float of a block, no register staging.  The
// loads complete asynchronously; the caller waits (s_waitcnt) before the barrier that publishes them.
__device__ __forceinline__ void row_to_lds(const float* __restrict__ src, float* dst, int slot, int wid, int lane,
                                           int nthr) {
    for (int k = wid * 64; k < slot; k += nthr) {
        __builtin_amdgcn_global_load_lds((const void*)(src + k + lane),
                                         (__attribute__((address_space(3))) void*)(dst + k), 4, 0, 0);
    }
}

// e^x over the whole float64 range (x clamped to +-1100, beyond which the result is 0 or inf anyway):
// 2^n (ldexp) for the nearest integer n of x log2(e), times e^t, t = (x log2(e) - n) ln 2, |t| <= 0.347.
// Float64 (V double): e^t by its Taylor polynomial to degree 12 (truncation < 0.347^13 / 13! < 2e-16).
// Float32 form: 2^f by the float32 exp2 (1 ulp).
template <typename V>
__device__ __forceinline__ double exp_wide(double x) {
    const double y = fmin(fmax(x, -1100.0), 1100.0) * 1.4426950408889634;
    const double n = __builtin_rint(y);
    if constexpr (sizeof(V) == 4) {
        return ldexp((double)__builtin_amdgcn_exp2f((float)(y - n)), (int)n);
    } else {
        const double t = (y - n) * 0.6931471805599453;
        constexpr double C[13] = {1.0, 1.0, 1.0 / 2, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720, 1.0 / 5040, 1.0 / 40320,
                                  1.0 / 362880, 1.0 / 3628800, 1.0 / 39916800, 1.0 / 479001600};
        double p = C[12];
#pragma unroll
        for (int k = 11; k >= 0; --k) p = fma(p, t, C[k]);
        return ldexp(p, (int)n);
    }
}

// Combine (m2, s2) into the running (m, s) of a log-sum-exp: one exp of the non-positive difference,
// in float32 (the per-round merges of the wave and piece forms).
__device__ __forceinline__ void lse_merge(double& m, double& s, double m2, double s2) {
    const double d = m2 - m;
    const double e = (double)__expf(-(float)fabs(d));
    s = d > 0.0 ? fma(s, e, s2) : fma(s2, e, s);
    m = fmax(m, m2);
}

// The same merge with the exp in float64 (exp_wide): for merging whole results (edge rows into the interior
// rows', a walk slice into the previous ones), whose maxima may lie far apart while both sums matter.
__device__ __forceinline__ void lse_merge_exact(double& m, double& s, double m2, double s2) {
    const double d = m2 - m;
    const double e = exp_wide<double>(-fabs(d));
    s = d > 0.0 ? fma(s, e, s2) : fma(s2, e, s);
    m = fmax(m, m2);
}

__device__ __forceinline__ double vfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float vfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Geometry of one workgroup's table slice, shared by the band kernels.
struct BandGeom {
    double rdb, offd, rdd, qbd, xbmax, xdmax;
    int nd;
};

// The bilinear (beta, Delta) interpolation at both phase bins and the phase interpolation of one lookup
// from rows R0 (row b) and R1 (row b + 1).  A row is stored Delta-major, [Delta][phase bin], so the four
// values of a row that one lookup reads (Delta columns d0 and d0 + 1, stored bins lo and lo + 1) sit at
// o, o + 1, o + NPHIF and o + NPHIF + 1 from one address: two ds_read2 per row.
template <typename V>
__device__ __forceinline__ V interp(const float* R0, const float* R1, int o, V fd, V fb, V wc) {
    const V a00 = R0[o], a01 = R0[o + 1], a10 = R0[o + NPHIF], a11 = R0[o + NPHIF + 1];
    const V b00 = R1[o], b01 = R1[o + 1], b10 = R1[o + NPHIF], b11 = R1[o + NPHIF + 1];
    // Delta, then beta, at the lower (x0) and upper (x1) bin, then the phase
    const V r0l = vfma(fd, a10 - a00, a00), r0u = vfma(fd, a11 - a01, a01);
    const V r1l = vfma(fd, b10 - b00, b00), r1u = vfma(fd, b11 - b01, b01);
    const V vl = vfma(fb, r1l - r0l, r0l), vu = vfma(fb, r1u - r0u, r0u);
    return vfma(wc, vu - vl, vl);
}

// the walk record's fields
__device__ __forceinline__ int rec_offset(uint2 e) {
    return (int)(__builtin_amdgcn_alignbit(e.y, e.x, REC_O_SHIFT) & REC_O_MASK);
}
__device__ __forceinline__ uint32_t rec_fd(uint2 e) { return e.x & FD_MASK; }
__device__ __forceinline__ uint32_t rec_wc(uint2 e) { return (e.y >> REC_W_SHIFT) & FC_MASK; }

// One lookup of an interior row b (beta not clamped there): the record carries Delta's column and fraction
// and the phase pair; only the beta fraction depends on the read, xb - b = aL rdb + (offb - b), formed in
// float64 (cb = offb - b).  Returns the interpolated value (without the drawn-time log weight).
template <typename V>
__device__ __forceinline__ V lean_lookup(const float* R0, const float* R1, double rdb, double cb, float aLf,
                                         uint2 e) {
    const V fb = (V)fma((double)aLf, rdb, cb);
    const V fd = (V)rec_fd(e) * (V)(1.0 / FC_SCALE);
    const V wc = (V)rec_wc(e) * (V)(1.0 / FC_SCALE);
    return interp<V>(R0, R1, rec_offset(e), fd, fb, wc);
}

// One lookup of App. B (the reference's gather.py:45-98) in any row, beta clamp included: the record's
// phase pair with Delta recomputed from ln|z_L| (aLf) and ln|z_H| (aH) in float64.  Used for the first and
// the last row pair of a table, the rows where beta may be clamped.
template <typename V>
__device__ __forceinline__ double exact_lookup(const float* R0, const float* R1, const BandGeom& g, double bd,
                                               double offb, double lw_rand, float aLf, uint2 e, float aH) {
    const double aL = (double)aLf;
    // beta coordinate (clamped) and its fraction in row b
    const double xbr = fma(aL, g.rdb, offb);
    const double xb = fmin(fmax(xbr, 0.0), g.xbmax);
    const double fb = xb - bd;
    // Delta coordinate: ln q_L moves with the beta clamp, ln q_H stays; xd lies in [0, nd - 1 - 1e-6]
    // (NaN included: fmax returns the number), so d0 = floor(xd) is in [0, nd - 2]
    const double aLc = aL + (xb - xbr) / g.rdb;
    const double xd = fmin(fmax(((double)aH - aLc) * g.rdd + g.offd, 0.0), g.xdmax);
    const int d0 = (int)xd;
    const double fd = xd - d0;
    const int lo = rec_offset(e) % NPHIF;
    const V v = interp<V>(R0, R1, d0 * NPHIF + lo, (V)fd, (V)fb, (V)rec_wc(e) * (V)(1.0 / FC_SCALE));
    return (double)v + ((e.y & REC_DRAWN) ? lw_rand : 0.0);
}

// one workgroup per (tile of reads of one table, lag); PN walked times per piece; V the type of the
// table interpolation (double, or float for the float32 form)
template <int PN, typename V>
__global__ __launch_bounds__(BAND_THREADS) void k_gather_piece(
    const uint2* __restrict__ stream, const float* __restrict__ wa, int nw, const float* __restrict__ lw_g,
    const float* __restrict__ T, int64_t tstride, int nphif, int nb, int nd, int ntau,
    const int32_t* __restrict__ jm, const int32_t* __restrict__ jtab, const double* __restrict__ joffb,
    const int32_t* __restrict__ tiles, const int32_t* __restrict__ first_last, const uint16_t* __restrict__ bounds,
    double rdb, double offd, double rdd, double* __restrict__ om, double* __restrict__ os, int accumulate) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int rowlen = nphif * nd, slot = row_slot(rowlen);
    float* rows = (float*)lds;                                              // 3 ring slots
    double* pm = (double*)(lds + piece_res_offset(rowlen));                 // (max, sum) of each dealt piece
    double* ps = pm + BAND_THREADS;
    int* qex = (int*)(ps + BAND_THREADS);                                   // first piece of each read
    uint16_t* plo = (uint16_t*)(qex + BAND_THREADS);                        // walk range of each read in row b
    uint16_t* phi = plo + BAND_THREADS;
    uint16_t* owner = phi + BAND_THREADS;                                   // read of each dealt piece
    int* wsum = (int*)(owner + BAND_THREADS);                               // 16 wave sums + band range
    const int r = blockIdx.y;
    const int j0 = tiles[2 * blockIdx.x], nj = tiles[2 * blockIdx.x + 1];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const float* Tr = T + (int64_t)jtab[j0] * tstride + (int64_t)r * nb * rowlen;

    int my_first = INT_MAX, my_last = INT_MIN;
    if (tid < nj) {
        my_first = first_last[2 * (j0 + tid)];
        my_last = first_last[2 * (j0 + tid) + 1];
    }
    // running (max, sum) of this lane's read, merged piece by piece in piece order
    double amx = -1e300, asum = 0.0;
    if (tid == 0) {
        wsum[16] = INT_MAX;
        wsum[17] = INT_MIN;
    }
    __syncthreads();
    if (tid < nj) {
        atomicMin(&wsum[16], my_first);
        atomicMax(&wsum[17], my_last);
    }
    __syncthreads();
    // the interior rows of the band (rows 0 and nb - 2, where beta may be clamped, are k_gather_edge's)
    const int b_start = max(wsum[16], 1), b_end = min(wsum[17], nb - 3);
    // rows b_start .. b_start + 2 (as far as the band and the table reach)
    const int b_top = min(b_end + 1, nb - 1);
    for (int q = 0; q < 3 && b_start + q <= b_top; ++q)
        row_to_lds(Tr + (int64_t)(b_start + q) * rowlen, rows + ((b_start + q) % 3) * slot, slot, wid, lane,
                   BAND_THREADS);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    for (int b = b_start; b <= b_end; ++b) {
        // row b + 2 into the slot row b - 1 left; it lands while row b is read
        if (b > b_start && b + 2 <= b_top)
            row_to_lds(Tr + (int64_t)(b + 2) * rowlen, rows + ((b + 2) % 3) * slot, slot, wid, lane, BAND_THREADS);
        // walk range and pieces of every read in this row
        int lo_p = 0, hi_p = 0;
        if (tid < nj && b >= my_first && b <= my_last) {
            const uint16_t* bj = bounds + (int64_t)(j0 + tid) * (nb + 1);
            lo_p = bj[b - my_first];
            hi_p = bj[b + 1 - my_first];
        }
        const int q = (hi_p - lo_p + PN - 1) / PN;
        int x = q;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (lane == 63) wsum[wid] = x;
        __syncthreads();
        if (wid == 0) {
            int sw = lane < 16 ? wsum[lane] : 0;
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) {
                const int y = __shfl_up(sw, d, 64);
                if (lane >= d) sw += y;
            }
            if (lane < 16) wsum[lane] = sw;
        }
        __syncthreads();
        const int my_qex = x - q + (wid ? wsum[wid - 1] : 0);
        const int Q = wsum[15];
        qex[tid] = my_qex;
        plo[tid] = (uint16_t)lo_p;
        phi[tid] = (uint16_t)hi_p;

        const int sb = b % 3;
        const float* R0 = rows + sb * slot;
        const float* R1 = rows + (sb + 1 == 3 ? 0 : sb + 1) * slot;
        const double bd = (double)b;
        // rounds of one piece per lane
        for (int base = 0; base < Q; base += BAND_THREADS) {
            const int lim = min(Q - base, BAND_THREADS);
            const int my_lo = max(my_qex, base), my_hi = min(my_qex + q, base + lim);
            for (int i = my_lo; i < my_hi; ++i) owner[i - base] = (uint16_t)tid;
            __syncthreads();
            if (tid < lim) {
                const int jj = owner[tid];
                const int64_t job = (int64_t)j0 + jj;
                const int pa = (int)plo[jj] + (base + tid - qex[jj]) * PN;
                const int pe = min(pa + PN, (int)phi[jj]);
                const int m = jm[job];
                const double cb = joffb[job] - bd;
                const double lw_rand = (double)lw_g[m];
                const float* wam = wa + (int64_t)m * nw;
                const uint2* sm = stream + ((int64_t)m * ntau + r) * nw;
                double mx = -1e300, s = 0.0;
                // records are requested in chunks of up to 8 before the chunk's first lookup; the slots
                // past the piece's end read its last record and add nothing (a = -1e300, the first slot
                // is always real).
                constexpr int CH = PN < 8 ? PN : 8;
#pragma unroll
                for (int c0 = 0; c0 < PN; c0 += CH) {
                    if (c0 > 0 && pa + c0 >= pe) break;
                    float aLv[CH];
                    uint2 ev[CH];
#pragma unroll
                    for (int k = 0; k < CH; ++k) {
                        const int i = min(pa + c0 + k, pe - 1);
                        aLv[k] = wam[i];
                        ev[k] = sm[i];
                    }
                    // value of lookup k of the chunk (-1e300 past the piece's end)
                    auto lookup = [&](int k) -> double {
                        const double v = (double)lean_lookup<V>(R0, R1, rdb, cb, aLv[k], ev[k]) +
                                         ((ev[k].y & REC_DRAWN) ? lw_rand : 0.0);
                        return pa + c0 + k < pe ? v : -1e300;
                    };
                    if constexpr (sizeof(V) == 4) {
                        // independent lookups (their LDS reads and interpolations overlap), then the
                        // chunk's (max, sum): exps of non-positive differences
                        double av[CH];
#pragma unroll
                        for (int k = 0; k < CH; ++k) av[k] = lookup(k);
                        double cm = av[0];
#pragma unroll
                        for (int k = 1; k < CH; ++k) cm = fmax(cm, av[k]);
                        double cs = 0.0;
#pragma unroll
                        for (int k = 0; k < CH; ++k) cs += (double)__expf((float)(av[k] - cm));
                        if (c0 == 0) {
                            mx = cm;
                            s = cs;
                        } else {
                            const double dlt = cm - mx;
                            const double ex = (double)__expf(-(float)fabs(dlt));
                            s = dlt > 0.0 ? fma(s, ex, cs) : fma(cs, ex, s);
                            mx = fmax(mx, cm);
                        }
                    } else {
                        // running (max, sum), one exp of the non-positive difference per lookup (the
                        // float64 interpolation leaves no registers for overlapping whole lookups)
#pragma unroll
                        for (int k = 0; k < CH; ++k) {
                            const double a = lookup(k);
                            const double dlt = a - mx;
                            const double ex = (double)__expf(-(float)fabs(dlt));
                            s = dlt > 0.0 ? fma(s, ex, 1.0) : s + ex;
                            mx = fmax(mx, a);
                        }
                    }
                }
                pm[tid] = mx;
                ps[tid] = s;
            }
            __syncthreads();
            // the read's lane merges its pieces of this round in piece order
            for (int i = my_lo; i < my_hi; ++i) {
                const double m2 = pm[i - base], s2 = ps[i - base];
                if (m2 > amx) {
                    asum = fma(asum, (double)__expf((float)(amx - m2)), s2);
                    amx = m2;
                } else {
                    asum = fma(s2, (double)__expf((float)(m2 - amx)), asum);
                }
            }
            __syncthreads();
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    }
    if (tid < nj) {
        const int64_t o = (int64_t)(j0 + tid) * ntau + r;
        if (accumulate) lse_merge_exact(amx, asum, om[o], os[o]);
        om[o] = amx;
        os[o] = asum;
    }
}

// records of one piece: PN consecutive walked times (positions past the read's range end lie inside
// the buffers' WALK_PAD_BYTES and are not used)
template <int PN>
__device__ __forceinline__ void load_piece(const float* __restrict__ wp, const uint2* __restrict__ sp,
                                           float (&aL)[PN], uint2 (&ev)[PN]) {
#pragma unroll
    for (int k = 0; k < PN; ++k) {
        aL[k] = wp[k];
        ev[k] = sp[k];
    }
}

// lane-shuffle helpers (wave64)
__device__ __forceinline__ double shfl_d(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ int shfl_i(int v, int src) { return __shfl(v, src, 64); }


// lane l receives lane l - n of its row of 16, or 0 where l - n lies outside the row (every lane
// written, so no previous value is needed)
template <int N>
__device__ __forceinline__ double shr_zero_d(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), DPP_ROW_SHR | N, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), DPP_ROW_SHR | N, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

template <int CTRL, int ROWM, int BANKM>
__device__ __forceinline__ double dpp_d(double old, double v) {
    const int lo = dpp_i<CTRL, ROWM, BANKM>(__double2loint(old), __double2loint(v));
    const int hi = dpp_i<CTRL, ROWM, BANKM>(__double2hiint(old), __double2hiint(v));
    return __hiloint2double(hi, lo);
}

// The wave form of the band kernel: one workgroup per (tile of reads of one table, lag) and the same
// three-row LDS ring as k_gather_piece, but each wave deals the pieces of its own 64 reads (lane l of
// wave w owns read w + 16 l) in rounds of 64 without workgroup barriers: a lane finds its piece's read
// by a binary search over the wave's piece offsets, takes the read's parameters from the owning lane
// by shuffles, computes the piece, and a segmented scan over lanes (DPP moves) leaves each read's
// (max, sum) of the round at its last lane, which the owning lane merges.  Only the ring rotation at
// the end of a row synchronizes the workgroup, so the waves of a CU overlap each other's memory
// latency.  Merges are float64; their order follows lane positions, so a read's result depends on its
// neighbours in the workgroup only through float64 rounding.
template <int PN, typename V>
__global__ __launch_bounds__(WAVE_THREADS) void k_gather_wave(
    const uint2* __restrict__ stream, const float* __restrict__ wa, int nw, const float* __restrict__ lw_g,
    const float* __restrict__ T, int64_t tstride, int nphif, int nb, int nd, int ntau,
    const int32_t* __restrict__ jm, const int32_t* __restrict__ jtab, const double* __restrict__ joffb,
    const int32_t* __restrict__ tiles, const int32_t* __restrict__ first_last, const uint16_t* __restrict__ bounds,
    double rdb, double offd, double rdd, double* __restrict__ om, double* __restrict__ os, int accumulate) {
    extern __shared__ __align__(16) unsigned char lds[];
    constexpr int NWAVE = WAVE_THREADS / 64;
    const int rowlen = nphif * nd, slot = row_slot(rowlen);
    float* rows = (float*)lds;                                              // 3 ring slots
    int* band = (int*)(lds + piece_res_offset(rowlen));                     // band range of the tile
    const int r = blockIdx.y;
    const int j0 = tiles[2 * blockIdx.x], nj = tiles[2 * blockIdx.x + 1];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int jl = wid + NWAVE * lane;                                      // this lane's read in the tile
    const bool mine = jl < nj;
    const float* Tr = T + (int64_t)jtab[j0] * tstride + (int64_t)r * nb * rowlen;

    // this lane's read: rows, template, offsets
    int my_first = INT_MAX, my_last = INT_MIN, my_m = 0;
    double my_offb = 0.0;
    float my_lw = 0.0f;
    if (mine) {
        const int64_t job = (int64_t)j0 + jl;
        my_first = first_last[2 * job];
        my_last = first_last[2 * job + 1];
        my_m = jm[job];
        my_offb = joffb[job];
        my_lw = lw_g[my_m];
    }
    double amx = -1e300, asum = 0.0;
    if (tid == 0) {
        band[0] = INT_MAX;
        band[1] = INT_MIN;
    }
    __syncthreads();
    if (mine) {
        atomicMin(&band[0], my_first);
        atomicMax(&band[1], my_last);
    }
    __syncthreads();
    // the interior rows of the band (rows 0 and nb - 2, where beta may be clamped, are k_gather_edge's)
    const int b_start = max(band[0], 1), b_end = min(band[1], nb - 3);
    const int b_top = min(b_end + 1, nb - 1);
    for (int q = 0; q < 3 && b_start + q <= b_top; ++q)
        row_to_lds(Tr + (int64_t)(b_start + q) * rowlen, rows + ((b_start + q) % 3) * slot, slot, wid, lane,
                   WAVE_THREADS);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    // walk range of this lane's read in row b (empty outside its rows)
    const uint16_t* bj = bounds + ((int64_t)j0 + (mine ? jl : 0)) * (nb + 1);
    auto bounds_of = [&](int b, int& lo_, int& hi_) {
        lo_ = hi_ = 0;
        if (mine && b >= my_first && b <= my_last) {
            lo_ = bj[b - my_first];
            hi_ = bj[b + 1 - my_first];
        }
    };
    // the wave's pieces in one row: this lane's read's walk range, its pieces, its first piece (an
    // exclusive sum over the lanes, by DPP) and the wave's total
    struct Layout {
        int lo, hi, q, qex, Qw;
    };
    auto layout = [&](int lo_, int hi_) -> Layout {
        Layout L;
        L.lo = lo_;
        L.hi = hi_;
        L.q = (int)((unsigned)(hi_ - lo_ + PN - 1) / PN);
        int x = L.q;
        x += dpp_i<DPP_ROW_SHR | 1, 0xf, 0xf>(0, x);
        x += dpp_i<DPP_ROW_SHR | 2, 0xf, 0xf>(0, x);
        x += dpp_i<DPP_ROW_SHR | 4, 0xf, 0xe>(0, x);
        x += dpp_i<DPP_ROW_SHR | 8, 0xf, 0xc>(0, x);
        x += dpp_i<DPP_ROW_BCAST15, 0xa, 0xf>(0, x);
        x += dpp_i<DPP_ROW_BCAST31, 0xc, 0xf>(0, x);
        L.qex = x - L.q;
        L.Qw = __builtin_amdgcn_readlane(x, 63);
        return L;
    };
    // a round's piece: its read (owning lane), walk range, parameters and records
    struct Piece {
        int o, pa, pe, m, last;
        bool act;
        double offb;
        float lw;
        float aL[PN];
        uint2 ev[PN];
    };
    // the piece of this lane in the round starting at piece `base` of a row with layout L; lanes past the
    // row's last piece keep an empty range at the start of a walk
    auto fetch = [&](const Layout& L, int base, Piece& P) {
        const int p = base + lane;
        P.act = p < L.Qw;
        // owner: the largest lane whose first piece is <= p (piece offsets ascend with the lane)
        int o = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
            const int c = o + step;
            const int qc = shfl_i(L.qex, c & 63);
            if (c < 64 && qc <= p) o = c;
        }
        const int lo_o = shfl_i(L.lo, o), hi_o = shfl_i(L.hi, o), qex_o = shfl_i(L.qex, o);
        P.o = o;
        P.m = shfl_i(my_m, o);
        P.offb = shfl_d(my_offb, o);
        P.lw = __shfl(my_lw, o, 64);
        P.pa = P.act ? lo_o + (p - qex_o) * PN : 0;
        P.pe = P.act ? min(P.pa + PN, hi_o) : 0;
        // the last lane of the read's run in this round (any lane for lanes past the last piece)
        P.last = (min(qex_o + (int)((unsigned)(hi_o - lo_o + PN - 1) / PN), base + 64) - 1 - base) & 63;
        load_piece<PN>(wa + (int64_t)P.m * nw + P.pa, stream + ((int64_t)P.m * ntau + r) * nw + P.pa, P.aL, P.ev);
    };

    // Rows flow into each other: row b + 1's layout is formed while row b runs (its bounds were requested
    // a row earlier), and the last round of row b fetches row b + 1's first piece.
    int n_lo, n_hi, c_lo, c_hi;
    bounds_of(b_start, c_lo, c_hi);
    bounds_of(b_start + 1, n_lo, n_hi);
    Layout S = layout(c_lo, c_hi);
    Piece pa_, pb_;
    bool pre = false;                           // pa_ holds row b's first piece
    for (int b = b_start; b <= b_end; ++b) {
        if (b > b_start && b + 2 <= b_top)
            row_to_lds(Tr + (int64_t)(b + 2) * rowlen, rows + ((b + 2) % 3) * slot, slot, wid, lane, WAVE_THREADS);
        int f_lo, f_hi;
        bounds_of(b + 2, f_lo, f_hi);           // row b + 2's walk range, used at the end of this row
        const Layout N = layout(n_lo, n_hi);    // row b + 1
        const int sb = b % 3;
        const float* R0 = rows + sb * slot;
        const float* R1 = rows + (sb + 1 == 3 ? 0 : sb + 1) * slot;
        const double bd = (double)b;
        const int q = S.q, qex = S.qex, Qw = S.Qw;
        // One round: the next piece (this row's next round, or row b + 1's first) is fetched into nxt,
        // then cur is computed and merged.  The round has no branches, so the compiler can overlap the
        // fetch's lane shuffles with the lookups: every slot of a piece is looked up, unused slots with
        // the records that follow the piece in memory (the read's later walked times, the next lag's or
        // template's, or the zeroed padding after the last; their offsets lie inside a row) and their
        // values discarded.
        auto round = [&](int base, Piece& cur, Piece& nxt) __attribute__((always_inline)) {
            const bool same_row = base + 64 < Qw;
            const Layout F{same_row ? S.lo : N.lo, same_row ? S.hi : N.hi, same_row ? S.q : N.q,
                           same_row ? S.qex : N.qex, same_row ? Qw : N.Qw};
            fetch(F, same_row ? base + 64 : 0, nxt);
            const bool act = cur.act;
            double mx = -1e300, sm = 0.0;
            {
                const int nuse = cur.pe - cur.pa;   // slots of the piece that hold its read's walked times
                const double cb = cur.offb - bd;
                if constexpr (sizeof(V) == 4) {
                    // float32 form: the piece's values in float, their max (the reference of the sum), then
                    // the exps, a drawn time's weighted by e^{lw} (at most n_time / n_random), summed in
                    // float64; slots past the piece's end are -inf and add nothing
                    float av[PN];
#pragma unroll
                    for (int k = 0; k < PN; ++k) {
                        const float v = lean_lookup<float>(R0, R1, rdb, cb, cur.aL[k], cur.ev[k]);
                        av[k] = k < nuse ? v : -INFINITY;
                    }
                    float mf = av[0];
#pragma unroll
                    for (int k = 1; k < PN; ++k) mf = fmaxf(mf, av[k]);
                    constexpr float LOG2E = 1.4426950408889634f;
                    const float ew = __expf(cur.lw), ml = mf * LOG2E;
#pragma unroll
                    for (int k = 0; k < PN; ++k) {
                        const float ex = __builtin_amdgcn_exp2f(fmaf(av[k], LOG2E, -ml));
                        sm += (double)((cur.ev[k].y & REC_DRAWN) ? ex * ew : ex);
                    }
                    mx = (double)mf;
                } else {
                    const double lw = (double)cur.lw;
                    auto lookup = [&](int k) -> double {
                        const double a = (double)lean_lookup<V>(R0, R1, rdb, cb, cur.aL[k], cur.ev[k]) +
                                         ((cur.ev[k].y & REC_DRAWN) ? lw : 0.0);
                        return k < nuse ? a : -1e300;
                    };
                    if constexpr (PN <= 6) {
                        // the piece's values first, then its (max, sum); the max, only the reference of the
                        // sum, is taken in float; -inf values leave the sum unchanged
                        double av[PN];
#pragma unroll
                        for (int k = 0; k < PN; ++k) av[k] = lookup(k);
                        float mf = (float)av[0];
#pragma unroll
                        for (int k = 1; k < PN; ++k) mf = fmaxf(mf, (float)av[k]);
                        mx = (double)mf;
#pragma unroll
                        for (int k = 0; k < PN; ++k) sm += (double)__expf((float)(av[k] - mx));
                    } else {
                        // a running (max, sum) over the lookups (fewer live registers for long pieces)
#pragma unroll
                        for (int k = 0; k < PN; ++k) {
                            const double a = lookup(k);
                            const double dlt = a - mx;
                            const double ex = (double)__expf(-(float)fabs(dlt));
                            sm = dlt > 0.0 ? fma(sm, ex, 1.0) : sm + ex;
                            mx = dlt > 0.0 ? a : mx;
                        }
                    }
                }
            }
            // Each read's (max, sum) over its run of lanes in this round, from two unsegmented scans.
            // The max: the owning lane ascends with the lane, so the prefix maximum of the key (owner,
            // piece max) at a lane is the maximum over its own run up to itself; the piece max enters
            // the key truncated to 26 bits of its order (the run max is only the reference of the sums,
            // so any value within 2^-17 relative of it serves).  Lanes past the last piece hold key 0 and
            // come after every active lane.  The sum: every lane's sum rescaled to its run's reference,
            // prefix-summed over the wave (float64); a run's sum is the difference of the prefix sums at
            // its last lane and before its first, and is at least ~1 (the lane holding the run's max
            // contributes its own largest term, ~1), so the difference loses ~1e-14 relative at most.
            uint32_t key = act ? ((uint32_t)cur.o << 26) | (float_order_key((float)mx) >> 6) : 0u;
            key = max(key, dpp_u<DPP_ROW_SHR | 1, 0xf, 0xf>(0u, key));
            key = max(key, dpp_u<DPP_ROW_SHR | 2, 0xf, 0xf>(0u, key));
            key = max(key, dpp_u<DPP_ROW_SHR | 4, 0xf, 0xe>(0u, key));
            key = max(key, dpp_u<DPP_ROW_SHR | 8, 0xf, 0xc>(0u, key));
            key = max(key, dpp_u<DPP_ROW_BCAST15, 0xa, 0xf>(0u, key));
            key = max(key, dpp_u<DPP_ROW_BCAST31, 0xc, 0xf>(0u, key));
            const double Mr = (double)order_key_float(__shfl(key, cur.last, 64) << 6);
            double t = act ? sm * (double)__expf((float)(mx - Mr)) : 0.0;
            t += shr_zero_d<1>(t);
            t += shr_zero_d<2>(t);
            t += shr_zero_d<4>(t);
            t += shr_zero_d<8>(t);
            t += dpp_d<DPP_ROW_BCAST15, 0xa, 0xf>(0.0, t);
            t += dpp_d<DPP_ROW_BCAST31, 0xc, 0xf>(0.0, t);
            // the owning lane takes its read's (max, sum) of this round from the run's lanes
            const bool in_round = q > 0 && qex < base + 64 && qex + q > base;
            const int tail = in_round ? min(qex + q, base + 64) - 1 - base : 0;
            const int head = in_round ? max(qex - base, 0) : 0;
            const double hm = (double)order_key_float(__shfl(key, tail, 64) << 6);
            const double before = shfl_d(t, (head + 63) & 63);
            const double hs = shfl_d(t, tail) - (head > 0 ? before : 0.0);
            if (in_round) lse_merge(amx, asum, hm, hs);
        };
        if (Qw > 0) {
            if (!pre) fetch(S, 0, pa_);
            // two rounds per iteration, the pieces alternating between pa_ and pb_
            for (int base = 0; base < Qw; base += 128) {
                round(base, pa_, pb_);
                if (base + 64 < Qw) round(base + 64, pb_, pa_);
            }
            // row b + 1's first piece came with the last round: into pb_ after an odd number of rounds
            if (((Qw + 63) >> 6) & 1) pa_ = pb_;
            pre = N.Qw > 0;
        } else {
            pre = false;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        S = N;
        n_lo = f_lo;
        n_hi = f_hi;
    }
    if (mine) {
        const int64_t o = ((int64_t)j0 + jl) * ntau + r;
        if (accumulate) lse_merge_exact(amx, asum, om[o], os[o]);
        om[o] = amx;
        os[o] = asum;
    }
}

// ------------------------------------------------------------------------------------------------------
// The scheduled form of the band kernel.
//
// Which walked times of which read fall in which beta row does not depend on the lag, so the pieces a
// workgroup deals (PN consecutive walked times of one read in one row) are listed once per launch, for
// every tile and row, before the gather: k_sched_count counts the pieces of every (tile, row),
// k_sched_emit writes their descriptors in row order and, within a row, in read order.  The gather
// (k_gather_sched, one workgroup per (tile, lag) as the other forms) then sweeps the rows with the same
// three-row LDS ring, and its waves take the pieces of the current row in chunks of 64 from a counter
// in LDS, so they stay busy until the row's last chunk.  Rows 0 and nb - 2, where beta may be clamped,
// are swept too, their chunks with the exact lookup (Delta recomputed with the clamp, ln|z_H| read for
// each walked time), so k_gather_edge only recomputes what this kernel flags.  Each read's sum is kept
// in LDS against a fixed
// reference: the value of its loudest walked time (the last of its sorted walk, found before the sweep
// with the exact lookup), so that a piece contributes s e^{m - M} (its own (max, sum) rescaled to the
// read's reference M) by one float64 LDS atomic add, and no lane has to meet the other pieces of its
// read.  M is a value of the read's own sum, so the sum is at least ~1; its terms exceed e^709 only if
// the read's values span more than ~700 (an overflow the final step detects, flagging the (read, lag)
// for k_gather_edge to recompute exactly).
//
// Piece descriptor, 32 bits: the read's index in its tile (bits 0-10), the slots used minus one (bits
// 11-13) and the walk position of its first slot (bits 14-31).
constexpr int SCHED_THREADS = 1024;
constexpr int SD_USE_SHIFT = 11, SD_POS_SHIFT = 14;
constexpr uint32_t SD_READ_MASK = 0x7FFu;
// reads of a tile per schedule thread: thread t takes reads 2t and 2t + 1, so the read order is kept
constexpr int SCHED_READS_PER_THREAD = SCHED_TILE_JOBS / SCHED_THREADS;
static_assert(SCHED_READS_PER_THREAD == 2, "the schedule kernels take two reads per thread");

// pieces a read may hold at most: each row's run ends in at most one partial piece
__host__ __device__ inline int64_t sched_cap_per_read(int nw, int pn, int nb) { return (nw + pn - 1) / pn + nb; }

// walk range of read j in row b (empty outside the read's rows, and for b = nb - 1, which no walked time
// reaches: rows run over 0 .. nb - 2)
__device__ __forceinline__ void row_run(const int32_t* __restrict__ first_last, const uint16_t* __restrict__ bounds,
                                        int64_t j, int nb, int b, int& lo, int& hi) {
    lo = hi = 0;
    const int first = first_last[2 * j], last = first_last[2 * j + 1];
    if (b <= nb - 2 && b >= first && b <= last) {
        const uint16_t* bj = bounds + j * (nb + 1);
        lo = bj[b - first];
        hi = bj[b + 1 - first];
    }
}

// sum over the workgroup of one int per thread (every thread receives it); red holds NWAVE ints
template <int NWAVE>
__device__ __forceinline__ int block_sum(int v, int* red) {
#pragma unroll
    for (int k = 32; k > 0; k >>= 1) v += __shfl_xor(v, k, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    int s = 0;
#pragma unroll
    for (int w = 0; w < NWAVE; ++w) s += red[w];
    __syncthreads();
    return s;
}

// pieces of every (tile, row): cnt[t * (nb + 1) + b]
template <int PN>
__global__ __launch_bounds__(SCHED_THREADS) void k_sched_count(const int32_t* __restrict__ tiles,
                                                               const int32_t* __restrict__ first_last,
                                                               const uint16_t* __restrict__ bounds, int nb,
                                                               int32_t* __restrict__ cnt) {
    __shared__ int red[SCHED_THREADS / 64];
    const int t = blockIdx.x, b = blockIdx.y;
    const int j0 = tiles[2 * t], nj = tiles[2 * t + 1];
    int q = 0;
    for (int u = 0; u < SCHED_READS_PER_THREAD; ++u) {
        const int jl = SCHED_READS_PER_THREAD * (int)threadIdx.x + u;
        if (jl < nj) {
            int lo, hi;
            row_run(first_last, bounds, (int64_t)j0 + jl, nb, b, lo, hi);
            q += (hi - lo + PN - 1) / PN;
        }
    }
    const int total = block_sum<SCHED_THREADS / 64>(q, red);
    if (threadIdx.x == 0) cnt[(int64_t)t * (nb + 1) + b] = total;
}

// descriptors of every (tile, row): the tile's pieces start at desc + j0 * cap (its first read's index
// times the per-read capacity), row b's at rowoff[t * (nb + 1) + b] after that, rowoff[.. + nb] = total
template <int PN>
__global__ __launch_bounds__(SCHED_THREADS) void k_sched_emit(const int32_t* __restrict__ tiles,
                                                              const int32_t* __restrict__ first_last,
                                                              const uint16_t* __restrict__ bounds, int nb,
                                                              const int32_t* __restrict__ cnt, int64_t cap,
                                                              int32_t* __restrict__ rowoff,
                                                              uint32_t* __restrict__ desc) {
    __shared__ int red[SCHED_THREADS / 64];
    __shared__ int wex[SCHED_THREADS / 64];
    const int t = blockIdx.x, b = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int j0 = tiles[2 * t], nj = tiles[2 * t + 1];
    const int32_t* ct = cnt + (int64_t)t * (nb + 1);
    // where row b starts: the pieces of the rows before it (nb < 1024 rows, one per thread)
    const int start = block_sum<SCHED_THREADS / 64>(tid < b ? ct[tid] : 0, red);
    // this thread's reads 2 tid and 2 tid + 1 (in read order) and their pieces in row b
    int lo[SCHED_READS_PER_THREAD], hi[SCHED_READS_PER_THREAD], q = 0;
    for (int u = 0; u < SCHED_READS_PER_THREAD; ++u) {
        const int jl = SCHED_READS_PER_THREAD * tid + u;
        lo[u] = hi[u] = 0;
        if (jl < nj) row_run(first_last, bounds, (int64_t)j0 + jl, nb, b, lo[u], hi[u]);
        q += (hi[u] - lo[u] + PN - 1) / PN;
    }
    // exclusive sum of q over the workgroup: within the wave by shuffles, then over the waves
    int x = q;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wex[wid] = x;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wid; ++w) before += wex[w];
    const int ex = before + x - q;
    uint32_t* out = desc + (int64_t)j0 * cap + start + ex;
    for (int u = 0; u < SCHED_READS_PER_THREAD; ++u) {
        const uint32_t jl = (uint32_t)(SCHED_READS_PER_THREAD * tid + u);
        for (int pa = lo[u]; pa < hi[u]; pa += PN) {
            const int use = min(PN, hi[u] - pa);
            *out++ = jl | ((uint32_t)(use - 1) << SD_USE_SHIFT) | ((uint32_t)pa << SD_POS_SHIFT);
        }
    }
    if (tid == 0) {
        rowoff[(int64_t)t * (nb + 1) + b] = start;
        if (b == nb - 1) rowoff[(int64_t)t * (nb + 1) + nb] = start + ct[b];
    }
}

// bytes of LDS the scheduled kernel uses for a row of rowlen floats: the ring, then per read of the tile
// its sum (float64), reference (float) and template (int), then the chunk counters and the band range.
// A read's ln-a offset and drawn-time weight are read from memory where a piece needs them (at compute
// time, a chunk after the fetch); the template, which the piece's record addresses need, from LDS.
__host__ __device__ constexpr size_t sched_lds_bytes(int rowlen) {
    return piece_res_offset(rowlen) + (size_t)SCHED_TILE_JOBS * (8 + 4 + 4) + 8 * sizeof(int);
}

template <int PN, typename V>
__global__ __launch_bounds__(SCHED_THREADS) void k_gather_sched(
    const uint2* __restrict__ stream, const float* __restrict__ aHs, const float* __restrict__ wa, int nw,
    const float* __restrict__ lw_g, const float* __restrict__ T, int64_t tstride, int nphif, int nb, int nd,
    int ntau, const int32_t* __restrict__ jm, const int32_t* __restrict__ jtab, const double* __restrict__ joffb,
    const int32_t* __restrict__ tiles, const int32_t* __restrict__ first_last, const uint32_t* __restrict__ desc,
    int64_t cap, const int32_t* __restrict__ rowoff, double rdb, double offd, double rdd, double* __restrict__ om,
    double* __restrict__ os, uint8_t* __restrict__ flags, int accumulate) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int rowlen = nphif * nd, slot = row_slot(rowlen);
    float* rows = (float*)lds;                                              // 3 ring slots
    double* s_acc = (double*)(lds + piece_res_offset(rowlen));              // each read's sum
    float* s_M = (float*)(s_acc + SCHED_TILE_JOBS);                         // each read's reference
    int* s_m = (int*)(s_M + SCHED_TILE_JOBS);                               // each read's template
    int* s_ctr = s_m + SCHED_TILE_JOBS;                                     // [0..2] chunks taken, [4..5] band
    const int r = blockIdx.y, t = blockIdx.x;
    const int j0 = tiles[2 * t], nj = tiles[2 * t + 1];
    const int tid = threadIdx.x, lane = tid & 63;
    const float* Tr = T + (int64_t)jtab[j0] * tstride + (int64_t)r * nb * rowlen;
    const int32_t* tm = jm + j0;                                            // the tile's templates
    const double* toffb = joffb + j0;                                       // and ln-a offsets
    const BandGeom geo{rdb, offd, rdd, rdd / rdb, nb - 1 - 1e-6, nd - 1 - 1e-6, nd};
    if (tid == 0) {
        s_ctr[4] = INT_MAX;
        s_ctr[5] = INT_MIN;
    }
    __syncthreads();
    for (int jl = tid; jl < nj; jl += SCHED_THREADS) {
        const int64_t j = (int64_t)j0 + jl;
        const int first = first_last[2 * j], last = first_last[2 * j + 1];
        const int m = tm[jl];
        const double offb = toffb[jl];
        const float lw = lw_g[m];
        s_acc[jl] = 0.0;
        s_m[jl] = m;
        // the reference: the value at the loudest walked time (row `last`, the exact lookup from memory)
        const int64_t e = ((int64_t)m * ntau + r) * nw + nw - 1;
        const float* R0 = Tr + (int64_t)last * rowlen;
        const double M = exact_lookup<V>(R0, R0 + rowlen, geo, (double)last, offb, (double)lw,
                                         wa[(int64_t)m * nw + nw - 1], stream[e], aHs[e]);
        s_M[jl] = (M > -1e30 && M < 1e30) ? (float)M : 0.0f;
        atomicMin(&s_ctr[4], first);
        atomicMax(&s_ctr[5], last);
    }
    __syncthreads();
    // every row of the band: rows 0 and nb - 2, where beta may be clamped, with the exact lookup
    const int b_start = max(s_ctr[4], 0), b_end = min(s_ctr[5], nb - 2);
    if (b_start <= b_end) {
        const int b_top = min(b_end + 1, nb - 1);
        const int wid = tid >> 6;
        // the chunk counters of rows b_start .. b_start + 2 (every slot: with row 0 swept apart from the
        // interior loop, no loop iteration restarts row 2's)
        if (tid < 3) s_ctr[tid] = 0;
        for (int q = 0; q < 3 && b_start + q <= b_top; ++q)
            row_to_lds(Tr + (int64_t)(b_start + q) * rowlen, rows + ((b_start + q) % 3) * slot, slot, wid, lane,
                       SCHED_THREADS);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        const int32_t* ro = rowoff + (int64_t)t * (nb + 1);
        const uint32_t* dt = desc + (int64_t)j0 * cap;
        // a chunk: 64 consecutive pieces of one row, one per lane (lanes past the row's last piece use
        // none of their slots)
        struct Chunk {
            int row;
            int jl, nuse, pa;
            double offb;
            float lw, M;
            float aL[PN];
            uint2 ev[PN];
        };
        // the next chunk of row b (its index from the row's counter), or -1 once the row is dealt
        auto take = [&](int b) -> int {
            int c = 0;
            if (lane == 0) c = atomicAdd(&s_ctr[b % 3], 1);
            c = __builtin_amdgcn_readfirstlane(c);
            return c * 64 < ro[b + 1] - ro[b] ? c : -1;
        };
        auto fetch = [&](int b, int c, Chunk& P) {
            const int Q = ro[b + 1] - ro[b];
            const int i = c * 64 + lane;
            const bool act = i < Q;
            const uint32_t d = act ? dt[ro[b] + i] : 0u;
            P.row = b;
            P.jl = (int)(d & SD_READ_MASK);
            P.nuse = act ? (int)((d >> SD_USE_SHIFT) & 7u) + 1 : 0;
            const int pa = (int)(d >> SD_POS_SHIFT);
            P.pa = pa;
            const int m = s_m[P.jl];
            P.offb = toffb[P.jl];
            P.lw = lw_g[m];
            P.M = s_M[P.jl];
            load_piece<PN>(wa + (int64_t)m * nw + pa, stream + ((int64_t)m * ntau + r) * nw + pa, P.aL, P.ev);
        };
        // one chunk's lookups: each piece's (max, sum), rescaled to its read's reference and added to the
        // read's sum
        auto compute = [&](const Chunk& P) __attribute__((always_inline)) {
            const int b = P.row;
            const int sb = b % 3;
            const float* R0 = rows + sb * slot;
            const float* R1 = rows + (sb + 1 == 3 ? 0 : sb + 1) * slot;
            const double cb = P.offb - (double)b;
            const int nuse = P.nuse;
            double mx, sm = 0.0;
            if constexpr (sizeof(V) == 4) {
                float av[PN];
#pragma unroll
                for (int k = 0; k < PN; ++k) {
                    const float v = lean_lookup<float>(R0, R1, rdb, cb, P.aL[k], P.ev[k]);
                    av[k] = k < nuse ? v : -INFINITY;
                }
                float mf = av[0];
#pragma unroll
                for (int k = 1; k < PN; ++k) mf = fmaxf(mf, av[k]);
                constexpr float LOG2E = 1.4426950408889634f;
                const float ew = __expf(P.lw), ml = mf * LOG2E;
#pragma unroll
                for (int k = 0; k < PN; ++k) {
                    const float ex = __builtin_amdgcn_exp2f(fmaf(av[k], LOG2E, -ml));
                    sm += (double)((P.ev[k].y & REC_DRAWN) ? ex * ew : ex);
                }
                mx = (double)mf;
            } else {
                const double lw = (double)P.lw;
                double av[PN];
#pragma unroll
                for (int k = 0; k < PN; ++k) {
                    const double a = (double)lean_lookup<V>(R0, R1, rdb, cb, P.aL[k], P.ev[k]) +
                                     ((P.ev[k].y & REC_DRAWN) ? lw : 0.0);
                    av[k] = k < nuse ? a : -1e300;
                }
                float mf = (float)av[0];
#pragma unroll
                for (int k = 1; k < PN; ++k) mf = fmaxf(mf, (float)av[k]);
                mx = (double)mf;
#pragma unroll
                for (int k = 0; k < PN; ++k) sm += (double)__expf((float)(av[k] - mx));
            }
            if (nuse > 0) atomicAdd(&s_acc[P.jl], sm * exp_wide<V>(mx - (double)P.M));
        };

        // rows 0 and nb - 2, where beta may be clamped (about 0.1 % of the lookups): every chunk of row b,
        // each lane's piece by the exact lookup (Delta recomputed with the clamp from ln|z_L| and ln|z_H|)
        // in an online float64 log-sum-exp; rows b and b + 1 are in the ring
        auto edge_row = [&](int b) {
            const float* R0 = rows + (b % 3) * slot;
            const float* R1 = rows + ((b + 1) % 3) * slot;
            const int Q = ro[b + 1] - ro[b];
            for (int c = take(b); c >= 0; c = take(b)) {
                const int i = c * 64 + lane;
                if (i < Q) {
                    const uint32_t d = dt[ro[b] + i];
                    const int jl = (int)(d & SD_READ_MASK), nuse = (int)((d >> SD_USE_SHIFT) & 7u) + 1;
                    const int pa = (int)(d >> SD_POS_SHIFT);
                    const int m = s_m[jl];
                    const double offb = toffb[jl], lw = (double)lw_g[m];
                    const int64_t e0 = ((int64_t)m * ntau + r) * nw + pa;
                    const float* wp = wa + (int64_t)m * nw + pa;
                    double mx = -1e300, sm = 0.0;
                    for (int k = 0; k < nuse; ++k) {
                        const double a = exact_lookup<V>(R0, R1, geo, (double)b, offb, lw, wp[k], stream[e0 + k],
                                                         aHs[e0 + k]);
                        const double dd = a - mx;
                        const double ex = exp_wide<double>(-fabs(dd));
                        sm = dd > 0.0 ? fma(sm, ex, 1.0) : sm + ex;
                        mx = fmax(mx, a);
                    }
                    atomicAdd(&s_acc[jl], sm * exp_wide<V>(mx - (double)s_M[jl]));
                }
            }
        };
        if (b_start == 0) {
            edge_row(0);
            __syncthreads();
        }
        // the interior rows 1 .. nb - 3
        const int bi_start = max(b_start, 1), bi_end = min(b_end, nb - 3);
        Chunk cur, nxt;
        bool have = false;                      // cur holds a chunk of row b
        for (int b = bi_start; b <= bi_end; ++b) {
            // row b + 2 into the slot row b - 1 left; it lands while row b is read.  Row b + 2's counter
            // (row b - 1's, last used before the barrier that ended row b - 1) restarts for it.
            if (b > b_start && b + 2 <= b_top)
                row_to_lds(Tr + (int64_t)(b + 2) * rowlen, rows + ((b + 2) % 3) * slot, slot, wid, lane,
                           SCHED_THREADS);
            if (tid == 0) s_ctr[(b + 2) % 3] = 0;
            if (!have) {
                const int c = take(b);
                if (c >= 0) {
                    fetch(b, c, cur);
                    have = true;
                }
            }
            while (have) {
                // the next chunk (this row's, or once it is dealt, row b + 1's when it is interior) is
                // requested before this one is computed
                int nb_ = b, c = take(b);
                if (c < 0 && b + 1 <= bi_end) {
                    nb_ = b + 1;
                    c = take(b + 1);
                }
                if (c >= 0) fetch(nb_, c, nxt);
                compute(cur);
                have = c >= 0;
                if (have) cur = nxt;
                if (have && nb_ != b) break;    // cur already holds row b + 1's first chunk
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
        if (b_end == nb - 2 && nb - 2 > 0) {
            // rows nb - 2 and nb - 1 are in the ring: preloaded, or loaded by the interior sweep's last
            // iteration and landed at its barrier.  Its counter was restarted at the start or by the
            // iteration of row nb - 4, and the interior sweep takes no chunk of it.
            edge_row(nb - 2);
            __syncthreads();                    // every read's sum complete
        }
    }
    for (int jl = tid; jl < nj; jl += SCHED_THREADS) {
        const int64_t o = ((int64_t)j0 + jl) * ntau + r;
        const double S = s_acc[jl];
        const bool bad = !(S < INFINITY);       // a term beyond float64 (or a NaN input): recomputed
        // the sum as (M + ln S, 1): the log-sum itself, so that the merges after it (edge rows, slices,
        // lags) rescale terms of comparable size
        double mx = -1e300, sm = 0.0;
        if (!bad && S > 0.0) {
            mx = (double)s_M[jl] + log(S);
            sm = 1.0;
        }
        if (accumulate) lse_merge_exact(mx, sm, om[o], os[o]);
        om[o] = mx;
        os[o] = sm;
        flags[o] = bad ? 1 : 0;
    }
}

// The first and the last row pair of a table (rows 0 and nb - 2), where beta may be clamped: each lookup
// recomputes its Delta coordinate from ln|z_L| and ln|z_H| (exact_lookup).  With edge_rows (after the
// piece and wave forms, which sweep the interior rows only), one thread per (read, lag) walks the read's
// times in those rows, reading the table values from memory (about 0.1 % of the lookups fall there; most
// threads find none and return at once), and merges its (max, sum) into om, os, which the interior rows'
// gather has written.  A (read, lag) the scheduled form flagged (flags, when given; that form sweeps
// every row itself) has all its rows recomputed here the same way.
template <typename V>
__global__ __launch_bounds__(256) void k_gather_edge(
    const uint2* __restrict__ stream, const float* __restrict__ aHs, const float* __restrict__ wa, int nw,
    const float* __restrict__ lw_g, const float* __restrict__ T, int64_t tstride, int nphif, int nb, int nd,
    int ntau, const int32_t* __restrict__ jm, const int32_t* __restrict__ jtab, const double* __restrict__ joffb,
    int64_t njobs, const int32_t* __restrict__ first_last, const uint16_t* __restrict__ bounds, double rdb,
    double offd, double rdd, double* __restrict__ om, double* __restrict__ os, const uint8_t* __restrict__ flags,
    int edge_rows) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int r = blockIdx.y;
    if (j >= njobs) return;
    const int first = first_last[2 * j], last = first_last[2 * j + 1];
    const bool low = edge_rows && first == 0, high = edge_rows && last >= nb - 2;
    const bool redo = flags != nullptr && flags[j * ntau + r] != 0;
    if (!low && !high && !redo) return;
    const int rowlen = nphif * nd;
    const float* Tr = T + (int64_t)jtab[j] * tstride + (int64_t)r * nb * rowlen;
    const BandGeom geo{rdb, offd, rdd, rdd / rdb, nb - 1 - 1e-6, nd - 1 - 1e-6, nd};
    const int m = jm[j];
    const double offb = joffb[j], lw = (double)lw_g[m];
    const uint16_t* bj = bounds + j * (nb + 1);
    const int64_t rec = ((int64_t)m * ntau + r) * nw;
    double amx = -1e300, asum = 0.0;
    // rows 0 and nb - 2, then (flagged) every interior row of the read
    const int n_int = redo ? max(0, min(last, nb - 3) - max(first, 1) + 1) : 0;
    for (int pass = 0; pass < 2 + n_int; ++pass) {
        const int b = pass == 0 ? 0 : pass == 1 ? nb - 2 : max(first, 1) + pass - 2;
        if (b < first || b > last || (pass == 1 && nb - 2 == 0)) continue;
        const float* R0 = Tr + (int64_t)b * rowlen;
        for (int p = bj[b - first]; p < bj[b + 1 - first]; ++p) {
            const double a = exact_lookup<V>(R0, R0 + rowlen, geo, (double)b, offb, lw, wa[(int64_t)m * nw + p],
                                             stream[rec + p], aHs[rec + p]);
            const double d = a - amx;
            const double e = exp_wide<double>(-fabs(d));
            asum = d > 0.0 ? fma(asum, e, 1.0) : asum + e;
            amx = fmax(amx, a);
        }
    }
    if (amx > -1e300) {
        const int64_t o = j * ntau + r;
        lse_merge_exact(amx, asum, om[o], os[o]);
        om[o] = amx;
        os[o] = asum;
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ dst, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

// walk records of one segment into wa_dst (n_tmpl, nw) and stream_dst (n_tmpl, ntau, nw): positions
// off .. off + nw - 1 of every template's selection (times), or the arrival times off .. off + nw - 1
// (times NULL)
int launch_walk(Traj* traj, Times* times, int64_t off, int nw, Store* st, float* wa_dst, uint2* stream_dst,
                float* aH_dst) {
    Context& c = ctx();
    const int nm = traj->n_tmpl;
    EXM_CHECK(nw > 0 && nw <= 65535, EXM_ERR_ARG, "walk of %d times exceeds the 16-bit row bounds", nw);
    EXM_CHECK(off >= 0 && off + nw <= (times ? (int64_t)times->n_walk : traj->n_time), EXM_ERR_ARG,
              "walk slice [%lld, %lld) outside the walk", (long long)off, (long long)(off + nw));
    int nwp = 1;
    while (nwp < nw) nwp <<= 1;
    const size_t lds = sizeof(uint64_t) * nwp;
    EXM_CHECK(lds <= 160 * 1024, EXM_ERR_ARG, "walk slice of %d times exceeds the LDS sort", nw);
    const size_t walk_bytes = ((size_t)nm * nw * 4 + 255) & ~(size_t)255;
    void* wbuf = nullptr;
    EXM_TRY(scratch(5, 2 * walk_bytes, &wbuf));
    float* wl = (float*)wbuf;
    int32_t* wt = (int32_t*)((char*)wbuf + walk_bytes);
    ProfileScope prof("walk");
    EXM_HIP(hipFuncSetAttribute((const void*)k_walk_sort, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_walk_sort, dim3(nm), dim3(WALK_THREADS), lds, c.stream, traj->lqL(), traj->n_time,
                       times ? times->d_sel : nullptr, times ? times->d_lw : nullptr,
                       times ? (int64_t)times->n_walk : (int64_t)0, off, nw, nwp, wa_dst, wl, wt);
    EXM_HIP(hipGetLastError());
    const size_t ws_lds = walk_stream_lds_bytes(st->ntau);
    EXM_CHECK(ws_lds <= 160 * 1024, EXM_ERR_ARG, "walk stream of %d lags exceeds the LDS", st->ntau);
    EXM_CHECK((int64_t)st->nd * NPHIF <= (int64_t)REC_O_MASK + 1, EXM_ERR_ARG,
              "table rows of %d x %d values exceed the walk record's 14-bit offset", st->nd, NPHIF);
    EXM_HIP(hipFuncSetAttribute((const void*)k_walk_stream, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ws_lds));
    hipLaunchKernelGGL(k_walk_stream, dim3(nm, (nw + WS_TILE - 1) / WS_TILE), dim3(256), ws_lds, c.stream, traj->lqH(), traj->phH(),
                       traj->phL(), traj->n_time, st->d_tau, st->ntau, wt, wl, wa_dst, nw, 1.0 / st->ddelta,
                       -st->delta0 / st->ddelta, st->nd, stream_dst, aH_dst);
    EXM_HIP(hipGetLastError());
    // the padding after the records holds valid (zero) records: a piece may read up to a piece past the
    // last walked time of a read
    const int64_t n_rec = (int64_t)nm * st->ntau * nw;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c.stream, (uint32_t*)(stream_dst + n_rec),
                       (int64_t)(WALK_PAD_BYTES / 4), 0u);
    EXM_HIP(hipGetLastError());
    return EXM_OK;
}

int launch_walk(Traj* traj, Times* times, Store* st, float* wa_dst, uint2* stream_dst, float* aH_dst) {
    return launch_walk(traj, times, 0, times->n_walk, st, wa_dst, stream_dst, aH_dst);
}

// piece gather of every read of js; template ids index wa, stream and lw_g
// band kernel form of launch_piece: the scheduled form (the default) unless the piece or wave form is
// asked for (exm_set_gather_kernel 2 or 3)
enum class BandForm { Sched, Wave, Piece };
static BandForm band_form() {
    const int mode = gather_mode();
    return mode == 2 ? BandForm::Piece : mode == 3 ? BandForm::Wave : BandForm::Sched;
}

// walked times per piece of the form that launch_piece runs (defaults measured at 2000 templates x 2048
// walked times x 53 nodes: the scheduled form 46.7 ms with 8 against 47.4 ms with 6 and 52.5 ms with 4;
// the wave form 69.2 ms with 6 against 72.9 ms with 4 and 73.6 ms with 8)
static int band_piece_len(BandForm f) {
    const int p = ctx().piece;
    return p ? p : (f == BandForm::Piece ? PIECE : f == BandForm::Sched ? 8 : 6);
}

// Work memory of launch_piece beside the walk: first/last rows and row bounds of every read, and for the
// scheduled form the recompute flags (one byte per read and lag), the pieces per (tile, row), the row
// offsets and the piece descriptors (at most sched_cap_per_read per read).
struct PieceWork {
    size_t fl, bnd, flags, cnt, rowoff, desc, total;
};
static PieceWork piece_work(const Store* st, const JobSet& js, int nw) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    PieceWork w{};
    const size_t nj = (size_t)js.njobs;
    w.fl = 0;
    w.bnd = al(sizeof(int32_t) * 2 * nj);
    size_t end = w.bnd + al(sizeof(uint16_t) * nj * (st->nb + 1));
    const BandForm f = band_form();
    if (f == BandForm::Sched) {
        const size_t rows = (size_t)js.ntiles * (st->nb + 1);
        w.flags = end;
        w.cnt = w.flags + al(nj * st->ntau);
        w.rowoff = w.cnt + al(sizeof(int32_t) * rows);
        w.desc = w.rowoff + al(sizeof(int32_t) * rows);
        end = w.desc + al(sizeof(uint32_t) * nj * (size_t)sched_cap_per_read(nw, band_piece_len(f), st->nb));
    }
    w.total = end;
    return w;
}

size_t piece_bounds_bytes(const Store* st, const JobSet& js, int nw) { return piece_work(st, js, nw).total; }

int launch_piece(Store* st, const uint2* stream, const float* aHs, const float* wa, int nw, const float* lw_g,
                 const JobSet& js, double* om, double* os, void* bounds_buf, bool accumulate) {
    Context& c = ctx();
    const hipStream_t cs = cur_stream();
    const PieceWork pw = piece_work(st, js, nw);
    void* bbuf = bounds_buf;
    if (!bbuf) EXM_TRY(scratch(6, pw.total, &bbuf));
    char* wb = (char*)bbuf;
    int32_t* first_last = (int32_t*)(wb + pw.fl);
    uint16_t* bounds = (uint16_t*)(wb + pw.bnd);
    const double rdb = 2.0 / st->dbeta;
    {
        ProfileScope prof("bounds");
        hipLaunchKernelGGL(k_read_bounds, dim3((unsigned)((js.njobs + 3) / 4)), dim3(256), 0, cs, wa, nw,
                           js.d_m, js.d_offb, js.njobs, rdb, st->nb, first_last, bounds);
        EXM_HIP(hipGetLastError());
    }
    EXM_CHECK(piece_res_offset(st->nphif * st->nd) % 16 == 0, EXM_ERR_ARG, "piece results misaligned");
    EXM_CHECK(row_slot(st->nphif * st->nd) - st->nphif * st->nd <= TABLE_PAD_FLOATS, EXM_ERR_ARG,
              "row copy overruns the table padding");
    const BandForm form = band_form();
    static_assert(TILE_JOBS <= WAVE_THREADS && TILE_JOBS <= BAND_THREADS &&
                      SCHED_TILE_JOBS <= SCHED_READS_PER_THREAD * SCHED_THREADS &&
                      SCHED_TILE_JOBS <= (int)SD_READ_MASK + 1,
                  "a workgroup holds every read of its tile");
    const int rowlen = st->nphif * st->nd;
    const size_t lds = form == BandForm::Sched  ? sched_lds_bytes(rowlen)
                       : form == BandForm::Wave ? piece_res_offset(rowlen) + 16
                                                : piece_lds_bytes(rowlen);
    EXM_CHECK(lds <= 160 * 1024, EXM_ERR_ARG, "three table rows and the per-read state (%zu bytes) exceed the LDS", lds);
    const int pn = band_piece_len(form);
    EXM_CHECK(form == BandForm::Piece || pn != 16, EXM_ERR_ARG,
              "the scheduled and wave forms of the band kernel take pieces of 4, 6 or 8 walked times");
    EXM_CHECK(form != BandForm::Piece || pn != 6, EXM_ERR_ARG,
              "the piece form of the band kernel takes pieces of 4, 8 or 16 walked times");
    const double offd = -st->delta0 / st->ddelta, rdd = 1.0 / st->ddelta;
    uint8_t* flags = nullptr;
    if (form == BandForm::Sched) {
        // the pieces of every (tile, row), listed once for all lags
        EXM_CHECK(st->nb <= SCHED_THREADS, EXM_ERR_ARG, "the schedule takes at most %d beta rows", SCHED_THREADS);
        EXM_CHECK(nw < (1 << (32 - SD_POS_SHIFT)), EXM_ERR_ARG, "walk of %d times exceeds the piece descriptor", nw);
        const int64_t cap = sched_cap_per_read(nw, pn, st->nb);
        EXM_CHECK(cap * SCHED_TILE_JOBS < INT32_MAX, EXM_ERR_ARG, "a tile's pieces exceed the 32-bit row offsets");
        flags = (uint8_t*)(wb + pw.flags);
        int32_t* cnt = (int32_t*)(wb + pw.cnt);
        int32_t* rowoff = (int32_t*)(wb + pw.rowoff);
        uint32_t* desc = (uint32_t*)(wb + pw.desc);
        if (js.ntiles > 0) {
            ProfileScope prof("schedule");
            const dim3 grid(js.ntiles, st->nb);
            auto kc = pn == 4 ? k_sched_count<4> : pn == 6 ? k_sched_count<6> : k_sched_count<8>;
            auto ke = pn == 4 ? k_sched_emit<4> : pn == 6 ? k_sched_emit<6> : k_sched_emit<8>;
            hipLaunchKernelGGL(kc, grid, dim3(SCHED_THREADS), 0, cs, js.d_tiles, first_last, bounds, st->nb, cnt);
            EXM_HIP(hipGetLastError());
            hipLaunchKernelGGL(ke, grid, dim3(SCHED_THREADS), 0, cs, js.d_tiles, first_last, bounds, st->nb, cnt, cap,
                               rowoff, desc);
            EXM_HIP(hipGetLastError());
        }
        using SKern = decltype(&k_gather_sched<6, double>);
        SKern kern = c.f32_interp ? (pn == 4 ? k_gather_sched<4, float> : pn == 6 ? k_gather_sched<6, float>
                                                                                 : k_gather_sched<8, float>)
                                  : (pn == 4 ? k_gather_sched<4, double> : pn == 6 ? k_gather_sched<6, double>
                                                                                   : k_gather_sched<8, double>);
        EXM_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        ProfileScope prof("gather");
        hipLaunchKernelGGL(kern, dim3(js.ntiles, st->ntau), dim3(SCHED_THREADS), lds, cs, stream, aHs, wa, nw, lw_g,
                           st->d_T, st->tstride, st->nphif, st->nb, st->nd, st->ntau, js.d_m, js.d_tab, js.d_offb,
                           js.d_tiles, first_last, desc, cap, rowoff, rdb, offd, rdd, om, os, flags,
                           accumulate ? 1 : 0);
        EXM_HIP(hipGetLastError());
    } else {
        const bool wave = form == BandForm::Wave;
        using Kern = decltype(&k_gather_piece<8, double>);
        Kern kern;
        if (wave)
            kern = c.f32_interp ? (pn == 4 ? k_gather_wave<4, float> : pn == 6 ? k_gather_wave<6, float> : k_gather_wave<8, float>)
                                : (pn == 4 ? k_gather_wave<4, double> : pn == 6 ? k_gather_wave<6, double>
                                                                                : k_gather_wave<8, double>);
        else
            kern = c.f32_interp ? (pn == 4 ? k_gather_piece<4, float> : pn == 16 ? k_gather_piece<16, float>
                                                                                 : k_gather_piece<8, float>)
                                : (pn == 4 ? k_gather_piece<4, double> : pn == 16 ? k_gather_piece<16, double>
                                                                                  : k_gather_piece<8, double>);
        EXM_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        ProfileScope prof("gather");
        // tiles of at most TILE_JOBS reads: one per lane of the wave form, one per thread of the piece form
        hipLaunchKernelGGL(kern, dim3(js.ntiles_small, st->ntau), dim3(wave ? WAVE_THREADS : BAND_THREADS), lds, cs,
                           stream, wa, nw, lw_g, st->d_T, st->tstride, st->nphif, st->nb, st->nd, st->ntau, js.d_m,
                           js.d_tab, js.d_offb, js.d_tiles_small, first_last, bounds, rdb, offd, rdd, om, os,
                           accumulate ? 1 : 0);
        EXM_HIP(hipGetLastError());
    }
    // the piece and wave forms: rows 0 and nb - 2 (beta possibly clamped), merged into the interior rows'
    // (max, sum); the scheduled form: every row of any (read, lag) it flagged
    auto edge = c.f32_interp ? k_gather_edge<float> : k_gather_edge<double>;
    hipLaunchKernelGGL(edge, dim3((unsigned)((js.njobs + 255) / 256), st->ntau), dim3(256), 0, cs, stream, aHs, wa, nw,
                       lw_g, st->d_T, st->tstride, st->nphif, st->nb, st->nd, st->ntau, js.d_m, js.d_tab, js.d_offb,
                       js.njobs, first_last, bounds, rdb, offd, rdd, om, os, flags, form == BandForm::Sched ? 0 : 1);
    EXM_HIP(hipGetLastError());
    return EXM_OK;
}

__global__ void k_fill_f32(float* __restrict__ dst, int64_t n, float v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

__global__ void k_fill_f64(double* __restrict__ dst, int64_t n, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

int launch_fill(float* dst, int64_t n, float v) {
    if (n <= 0) return EXM_OK;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, cur_stream(), dst, n, v);
    EXM_HIP(hipGetLastError());
    return EXM_OK;
}

int launch_fill(double* dst, int64_t n, double v) {
    if (n <= 0) return EXM_OK;
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, cur_stream(), dst, n, v);
    EXM_HIP(hipGetLastError());
    return EXM_OK;
}

// walked times per launch of the band gather: the whole walk when it is short; longer walks (every
// arrival time, or selections of more than 8192 times) in slices of a power of two of at most 8192 times
// whose walk records take at most 4 GiB
int band_slice(const Traj* traj, const Store* st, const Times* times) {
    const int64_t total = times ? (int64_t)times->n_walk : traj->n_time;
    if (total <= 8192) return (int)total;
    const int64_t per = (int64_t)traj->n_tmpl * st->ntau * (int64_t)sizeof(uint2);
    int c = 8192;
    while (c > 256 && per * c > ((int64_t)4 << 30)) c >>= 1;
    return c;
}

// scratch of launch_gather_band: slot 4 (walk records), slot 5 (walk temporaries), slot 6 (row bounds)
size_t band_scratch_bytes(const Traj* traj, const Store* st, const JobSet& js, const Times* times) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t nm = (size_t)traj->n_tmpl, nw = (size_t)band_slice(traj, st, times);
    const size_t slot4 = al(nm * nw * 4 + WALK_PAD_BYTES) + al(nm * 4) + al(sizeof(uint2) * nm * st->ntau * nw + WALK_PAD_BYTES) +
                         sizeof(float) * nm * st->ntau * nw;
    const size_t slot5 = 2 * al(nm * nw * 4);
    return al(slot4) + al(slot5) + al(piece_bounds_bytes(st, js, (int)nw));
}

// the band gather of one segment: the walk (the selection, or every arrival time when times is NULL) in
// slices of band_slice() times, each slice's (max, sum) per read and lag merged into om, os
int launch_gather_band(Traj* traj, Store* st, const JobSet& js, Times* times, double* om, double* os) {
    const int nm = traj->n_tmpl;
    const int64_t total = times ? (int64_t)times->n_walk : traj->n_time;
    const int nw = band_slice(traj, st, times);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t wa_bytes = al((size_t)nm * nw * 4 + WALK_PAD_BYTES), lw_bytes = al((size_t)nm * 4);
    const size_t st_bytes = al(sizeof(uint2) * (size_t)nm * st->ntau * nw + WALK_PAD_BYTES);
    const size_t ah_bytes = sizeof(float) * (size_t)nm * st->ntau * nw;
    void* buf = nullptr;
    EXM_TRY(scratch(4, wa_bytes + lw_bytes + st_bytes + ah_bytes, &buf));
    float* wa = (float*)buf;
    float* lw_g = (float*)((char*)buf + wa_bytes);
    uint2* stream = (uint2*)((char*)buf + wa_bytes + lw_bytes);
    float* aHs = (float*)((char*)buf + wa_bytes + lw_bytes + st_bytes);
    EXM_TRY(launch_fill(lw_g, nm, times ? times->lw_rand : 0.0f));
    for (int64_t off = 0; off < total; off += nw) {
        const int n = (int)std::min<int64_t>(nw, total - off);
        EXM_TRY(launch_walk(traj, times, off, n, st, wa, stream, aHs));
        EXM_TRY(launch_piece(st, stream, aHs, wa, n, lw_g, js, om, os, nullptr, off > 0));
    }
    return EXM_OK;
}

}  // namespace exm
