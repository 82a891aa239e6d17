// Matched filter of one segment against every template (src/exmarg/search/bank.py:56-78):
//
//   w_I = conj(d_I) / sig2_I,  pad[m, nu] = W[m, nu - lo] w_I[nu - lo] on [lo, lo + n_freq), 0 elsewhere
//   z   = FFT_n_time(pad)      (forward, unnormalized)
//   lq  = 1/2 ln max(|z|^2, 1e-38),  ph = atan2(Im z, Re z)
//
// Weighting, transform and modulus/phase run in float64; the trajectories are rounded once to float32
// on output.  The default transform is the band transform below (two passes that read only the band
// and write the trajectories directly); exm_set_filter_transform(1), and lengths it does not take, use
// hipFFT (Z2Z, C2C in the float32 form) batched over a chunk of templates, out of place from an input
// buffer whose zero padding outside [lo, lo + n_freq) is written once (and again only when the geometry
// changes), so each chunk writes only its n_freq weighted bins.  The float32 form
// (exm_set_filter_precision(1), part of the precision form of the batched analysis) keeps the
// weighting and the modulus/phase in float64 but transforms in float32 (C2C), as the reference's GPU
// path does (complex64, src/exmarg/search/bank.py:66-75).
#include <cmath>
#include <vector>

#include "exm_common.h"
#include "exm_fmath.h"

#pragma clang fp contract(off)

namespace exm {

// the n_freq weighted bins of template m0 + j at [lo, lo + n_freq) of row j
__global__ void k_filter_pack(const float2* __restrict__ W, const double2* __restrict__ d,
                              const double* __restrict__ s2, int n_freq, int lo, int64_t n_time, int m0,
                              double2* __restrict__ pad) {
    const int j = blockIdx.y;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_freq) return;
    const float2 wt = W[(int64_t)(m0 + j) * n_freq + f];
    const double2 dd = d[f];
    const double s = s2[f];
    // w = conj(d) / sig2, rounded as numpy rounds a complex divided by a real: times 1/sig2
    const double inv = 1.0 / s;
    const double wr = dd.x * inv, wi = (-dd.y) * inv;
    const double ar = (double)wt.x, ai = (double)wt.y;
    pad[(int64_t)j * n_time + lo + f] = make_double2(ar * wr - ai * wi, ar * wi + ai * wr);
}

// the same weighting rounded to complex64 for the float32 transform
__global__ void k_filter_pack32(const float2* __restrict__ W, const double2* __restrict__ d,
                                const double* __restrict__ s2, int n_freq, int lo, int64_t n_time, int m0,
                                float2* __restrict__ pad) {
    const int j = blockIdx.y;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_freq) return;
    const float2 wt = W[(int64_t)(m0 + j) * n_freq + f];
    const double2 dd = d[f];
    const double inv = 1.0 / s2[f];
    const double wr = dd.x * inv, wi = (-dd.y) * inv;
    const double ar = (double)wt.x, ai = (double)wt.y;
    pad[(int64_t)j * n_time + lo + f] = make_float2((float)(ar * wr - ai * wi), (float)(ar * wi + ai * wr));
}

__global__ void k_zero_u2(uint2* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_uint2(0u, 0u);
}

__global__ void k_filter_post32(const float2* __restrict__ z, int64_t n_time, int m0, float* __restrict__ lq,
                                float* __restrict__ ph) {
    const int j = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_time) return;
    const float2 v = z[(int64_t)j * n_time + t];
    const double x = (double)v.x, y = (double)v.y;
    const double p = x * x + y * y;
    const int64_t o = (int64_t)(m0 + j) * n_time + t;
    lq[o] = (float)(0.5 * log(p > 1e-38 ? p : 1e-38));
    ph[o] = (float)atan2(y, x);
}

__global__ void k_filter_post(const double2* __restrict__ z, int64_t n_time, int m0, float* __restrict__ lq,
                              float* __restrict__ ph) {
    const int j = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_time) return;
    const double2 v = z[(int64_t)j * n_time + t];
    const double p = v.x * v.x + v.y * v.y;
    const int64_t o = (int64_t)(m0 + j) * n_time + t;
    lq[o] = (float)(0.5 * log(p > 1e-38 ? p : 1e-38));
    ph[o] = (float)atan2(v.y, v.x);
}

// ---------------------------------------------------------------- the band transform
// The transform's input is zero outside the band [lo, lo + n_freq) (n_freq / n_time = 12 % at the
// headline geometry), so the default path does not write the zero-padded input nor read it:
// with n_time = N1 * N2, N2 = 512, n = N2 n1 + n2 and k = k1 + N1 k2,
//   A[n2][k1]     = w_N^{n2 k1} sum_{n1} x[N2 n1 + n2] w_N1^{n1 k1}     (pass A, one column per n2)
//   z[k1 + N1 k2] = sum_{n2} A[n2][k1] w_N2^{n2 k2}                     (pass B, one column per k1)
// with w_L = e^{-2 pi i / L}.  Pass A reads only the band bins (the weighting of k_filter_pack fused
// into its load) and writes A (one complex value per sample of every template); pass B reads A and
// writes ln|z| and arg z (the arithmetic of k_filter_post).  The transform runs in float64 (double2),
// or in float32 (float2) for the float32 form.  Each pass transforms its columns in LDS by Stockham
// stages (radix 8, a last radix-4 or radix-2 stage for the remaining factor), several columns
// interleaved so that consecutive lanes touch consecutive addresses; every stage reads its values into
// registers, waits for the workgroup, then writes its outputs in place.  Twiddles come from one table
// of w_N^t, t < N, built on the host to ~1 ulp.
constexpr int BT_N2 = 512;
constexpr int BT_TILE_A = 2048;                       // complex values per pass-A workgroup
constexpr int BT_COLS_B = 16;                         // k1 columns per pass-B workgroup
constexpr int BT_THREADS_A = BT_TILE_A / 8;           // every thread holds 8 values per stage
constexpr int BT_THREADS_B = BT_COLS_B * BT_N2 / 8;
constexpr int BT_MIN_N1 = BT_COLS_B, BT_MAX_N1 = BT_TILE_A;
constexpr double BT_TWO_PI = 6.283185307179586476925286766559;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// a complex value of the transform's type C (double2, or float2 for the float32 form) from two doubles
template <typename C>
__device__ __forceinline__ C cmake(double re, double im) {
    if constexpr (sizeof(C) == 16) return make_double2(re, im);
    else return make_float2((float)re, (float)im);
}

// One Stockham stage of radix R over ncol interleaved transforms of length L, buf[idx * ncol + c]:
// thread t takes butterflies t + q nthr (q < 8 / R), butterfly b of column b % ncol and index
// i = b / ncol < L / R; k = i mod p, inputs x_r = buf[i + r L / R] times w_L^{k r L / (R p)}, outputs
// at j + r p with j = (i - k) R + k (p: the length of the transforms already formed).  ncol and p are
// powers of two and L ncol = 8 nthr.
template <int R, int L, int NCOL, int NTHR, typename C>
__device__ __forceinline__ void stockham_stage(C* buf, const C* tw, int p, int tid) {
    static_assert(L * NCOL == 8 * NTHR, "every thread holds 8 values per stage");
    constexpr int NB = 8 / R, T = L / R;
    C v[8];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int b = tid + q * NTHR;
        const int c = b % NCOL, i = b / NCOL;
#pragma unroll
        for (int r = 0; r < R; ++r) v[q * R + r] = buf[(i + r * T) * NCOL + c];
    }
    __syncthreads();
    const int tstride = L / (R * p);
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int b = tid + q * NTHR;
        const int c = b % NCOL, i = b / NCOL;
        const int k = i & (p - 1);
        C* x = v + q * R;
        if (p > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) x[r] = cmul(x[r], tw[k * r * tstride]);
        }
        const int j = (i - k) * R + k;
        if constexpr (R == 8) {
            // a_k = x_k + x_{k+4}, b_k = (x_k - x_{k+4}) w_8^k; the even outputs are DFT4(a), the odd DFT4(b)
            using T = decltype(x[0].x);
            const T r2 = (T)0.70710678118654752440;
            const C a0 = cadd(x[0], x[4]), a1 = cadd(x[1], x[5]), a2 = cadd(x[2], x[6]), a3 = cadd(x[3], x[7]);
            const C e1 = csub(x[1], x[5]), e2 = csub(x[2], x[6]), e3 = csub(x[3], x[7]);
            const C b0 = csub(x[0], x[4]);
            const C b1 = C{(e1.x + e1.y) * r2, (e1.y - e1.x) * r2};     // times (1 - i) / sqrt 2
            const C b2 = C{e2.y, -e2.x};                               // times -i
            const C b3 = C{(e3.y - e3.x) * r2, -(e3.x + e3.y) * r2};   // times (-1 - i) / sqrt 2
            const C sa02 = cadd(a0, a2), da02 = csub(a0, a2), sa13 = cadd(a1, a3), da13 = csub(a1, a3);
            const C sb02 = cadd(b0, b2), db02 = csub(b0, b2), sb13 = cadd(b1, b3), db13 = csub(b1, b3);
            buf[j * NCOL + c] = cadd(sa02, sa13);
            buf[(j + p) * NCOL + c] = cadd(sb02, sb13);
            buf[(j + 2 * p) * NCOL + c] = C{da02.x + da13.y, da02.y - da13.x};
            buf[(j + 3 * p) * NCOL + c] = C{db02.x + db13.y, db02.y - db13.x};
            buf[(j + 4 * p) * NCOL + c] = csub(sa02, sa13);
            buf[(j + 5 * p) * NCOL + c] = csub(sb02, sb13);
            buf[(j + 6 * p) * NCOL + c] = C{da02.x - da13.y, da02.y + da13.x};
            buf[(j + 7 * p) * NCOL + c] = C{db02.x - db13.y, db02.y + db13.x};
        } else if constexpr (R == 4) {
            const C s02 = cadd(x[0], x[2]), d02 = csub(x[0], x[2]);
            const C s13 = cadd(x[1], x[3]), d13 = csub(x[1], x[3]);
            buf[j * NCOL + c] = cadd(s02, s13);
            buf[(j + p) * NCOL + c] = C{d02.x + d13.y, d02.y - d13.x};      // d02 - i d13
            buf[(j + 2 * p) * NCOL + c] = csub(s02, s13);
            buf[(j + 3 * p) * NCOL + c] = C{d02.x - d13.y, d02.y + d13.x};  // d02 + i d13
        } else {
            buf[j * NCOL + c] = cadd(x[0], x[1]);
            buf[(j + p) * NCOL + c] = csub(x[0], x[1]);
        }
    }
    __syncthreads();
}

// the forward transforms of length L of the NCOL interleaved columns in buf
template <int L, int NCOL, int NTHR, typename C>
__device__ __forceinline__ void stockham(C* buf, const C* tw, int tid) {
    int p = 1;
#pragma unroll
    for (; p * 8 <= L; p *= 8) stockham_stage<8, L, NCOL, NTHR>(buf, tw, p, tid);
    if (p * 4 == L) stockham_stage<4, L, NCOL, NTHR>(buf, tw, p, tid);
    else if (p * 2 == L) stockham_stage<2, L, NCOL, NTHR>(buf, tw, p, tid);
}

// pass A: columns n2_0 .. n2_0 + TA - 1 (TA = BT_TILE_A / N1) of template m0 + blockIdx.y, in the
// arithmetic of C (double2; float2 for the float32 form, whose weighted bins are rounded to complex64 as
// k_filter_pack32 rounds them)
template <int N1, typename C>
__global__ __launch_bounds__(BT_THREADS_A) void k_bt_pass_a(const float2* __restrict__ W, const double2* __restrict__ d,
                                                            const double* __restrict__ s2, int n_freq, int lo,
                                                            int64_t n_time, int m0, const double2* __restrict__ twN,
                                                            C* __restrict__ A) {
    constexpr int TA = BT_TILE_A / N1;
    __shared__ C buf[BT_TILE_A];                       // [n1][TA]
    __shared__ C tw[N1];                               // w_N1^t = w_N^{512 t}
    __shared__ C twl[BT_N2];                           // w_N^t, t < 512
    const int j = blockIdx.y, tid = threadIdx.x;
    const int n2_0 = blockIdx.x * TA;
    for (int t = tid; t < N1; t += BT_THREADS_A) {
        const double2 w = twN[(int64_t)t * BT_N2];
        tw[t] = cmake<C>(w.x, w.y);
    }
    for (int t = tid; t < BT_N2; t += BT_THREADS_A) twl[t] = cmake<C>(twN[t].x, twN[t].y);
    const float2* Wm = W + (int64_t)(m0 + j) * n_freq;
    for (int i = tid; i < BT_TILE_A; i += BT_THREADS_A) {
        const int c = i % TA, n1 = i / TA;
        const int f = n1 * BT_N2 + n2_0 + c - lo;
        C v = cmake<C>(0.0, 0.0);
        if (f >= 0 && f < n_freq) {
            // w = conj(d) / sig2 (times 1/sig2, as numpy divides a complex by a real), times W
            const float2 wt = Wm[f];
            const double2 dd = d[f];
            const double inv = 1.0 / s2[f];
            const double wr = dd.x * inv, wi = (-dd.y) * inv;
            const double ar = (double)wt.x, ai = (double)wt.y;
            v = cmake<C>(ar * wr - ai * wi, ar * wi + ai * wr);
        }
        buf[i] = v;
    }
    __syncthreads();
    stockham<N1, TA, BT_THREADS_A>(buf, tw, tid);
    // times w_N^{n2 k1} = w_N^{512 u} w_N^{v} with n2 k1 = 512 u + v (n2 k1 < 512 N1)
    C* Aj = A + (int64_t)j * n_time;
    for (int i = tid; i < BT_TILE_A; i += BT_THREADS_A) {
        const int c = i / N1, k1 = i % N1;
        const int n2 = n2_0 + c;
        const int e = n2 * k1;
        Aj[(int64_t)n2 * N1 + k1] = cmul(buf[k1 * TA + c], cmul(tw[e / BT_N2], twl[e % BT_N2]));
    }
}

// pass B: columns k1_0 .. k1_0 + 15 of template m0 + blockIdx.y, then ln|z| and arg z (float64 from
// the transform's values, as k_filter_post / k_filter_post32 form them)
template <typename C>
__global__ __launch_bounds__(BT_THREADS_B) void k_bt_pass_b(const C* __restrict__ A, int n1len, int64_t n_time, int m0,
                                                            const double2* __restrict__ twN, float* __restrict__ lq,
                                                            float* __restrict__ ph) {
    extern __shared__ __align__(16) unsigned char bsh_raw[];
    C* buf = (C*)bsh_raw;                              // [n2][16]
    C* tw = buf + BT_COLS_B * BT_N2;                   // w_512^t
    const int j = blockIdx.y, tid = threadIdx.x;
    const int k1_0 = blockIdx.x * BT_COLS_B;
    const int64_t tstep = n_time / BT_N2;
    for (int t = tid; t < BT_N2; t += BT_THREADS_B) tw[t] = cmake<C>(twN[t * tstep].x, twN[t * tstep].y);
    const C* Aj = A + (int64_t)j * n_time;
    for (int i = tid; i < BT_COLS_B * BT_N2; i += BT_THREADS_B) {
        const int c = i & (BT_COLS_B - 1), n2 = i / BT_COLS_B;
        buf[i] = Aj[(int64_t)n2 * n1len + k1_0 + c];
    }
    __syncthreads();
    stockham<BT_N2, BT_COLS_B, BT_THREADS_B>(buf, tw, tid);
    const int64_t row = (int64_t)(m0 + j) * n_time;
    for (int i = tid; i < BT_COLS_B * BT_N2; i += BT_THREADS_B) {
        const int c = i & (BT_COLS_B - 1), k2 = i / BT_COLS_B;
        const C v = buf[i];
        const double x = (double)v.x, y = (double)v.y;
        const double p = x * x + y * y;
        const int64_t o = row + k1_0 + c + (int64_t)n1len * k2;
        lq[o] = (float)(0.5 * ln_pos(p > 1e-38 ? p : 1e-38));
        ph[o] = (float)atan2_f(y, x);
    }
}

template <typename C>
constexpr size_t bt_lds_b() { return sizeof(C) * (size_t)(BT_COLS_B * BT_N2 + BT_N2); }

// whether the band transform takes transforms of length n_time (N1 = n_time / 512 a power of two in
// [16, 2048])
static bool band_transform_fits(int64_t n_time) {
    if (n_time % BT_N2) return false;
    const int64_t n1 = n_time / BT_N2;
    return n1 >= BT_MIN_N1 && n1 <= BT_MAX_N1 && (n1 & (n1 - 1)) == 0;
}

// e^{-2 pi i t / n} for t < n (n a multiple of 8): cos and sin of angles <= pi/4 mapped by the
// symmetries of the octants, so every entry is within ~1 ulp
static void twiddle_table(int64_t n, std::vector<double2>& w) {
    w.resize((size_t)n);
    const int64_t q4 = n / 4, q8 = n / 8;
    for (int64_t t = 0; t < n; ++t) {
        const int64_t quad = t / q4, r = t - quad * q4;
        double c, s;                                   // cos and sin of 2 pi r / n, r < n / 4
        if (r <= q8) {
            const double a = BT_TWO_PI * ((double)r / (double)n);
            c = std::cos(a);
            s = std::sin(a);
        } else {
            const double a = BT_TWO_PI * ((double)(q4 - r) / (double)n);
            c = std::sin(a);
            s = std::cos(a);
        }
        double cc, ss;                                 // of 2 pi t / n = quad pi / 2 + 2 pi r / n
        switch (quad) {
            case 0: cc = c; ss = s; break;
            case 1: cc = -s; ss = c; break;
            case 2: cc = -c; ss = -s; break;
            default: cc = s; ss = -c; break;
        }
        w[(size_t)t] = make_double2(cc, -ss);
    }
}

// the device twiddle table of length n, built once per length and kept
static int twiddles(int64_t n, const double2** out) {
    Context& c = ctx();
    auto it = c.bt_twiddles.find(n);
    if (it != c.bt_twiddles.end()) {
        *out = (const double2*)it->second;
        return EXM_OK;
    }
    std::vector<double2> w;
    twiddle_table(n, w);
    void* d = nullptr;
    EXM_HIP(hipMalloc(&d, sizeof(double2) * (size_t)n));
    EXM_HIP(hipMemcpy(d, w.data(), sizeof(double2) * (size_t)n, hipMemcpyHostToDevice));
    c.bt_twiddles[n] = d;
    *out = (const double2*)d;
    return EXM_OK;
}

static int fft_plan(int64_t n_time, int batch, bool f32, hipfftHandle* out) {
    Context& c = ctx();
    const uint64_t key = ((uint64_t)n_time << 21) | ((uint64_t)f32 << 20) | (uint64_t)batch;
    auto it = c.fft_plans.find(key);
    if (it != c.fft_plans.end()) {
        *out = it->second;
        return EXM_OK;
    }
    hipfftHandle p;
    EXM_FFT(hipfftCreate(&p));
    int n[1] = {(int)n_time};
    size_t ws = 0;
    EXM_FFT(hipfftMakePlanMany(p, 1, n, nullptr, 1, (int)n_time, nullptr, 1, (int)n_time, f32 ? HIPFFT_C2C : HIPFFT_Z2Z,
                               batch, &ws));
    EXM_FFT(hipfftSetStream(p, c.stream));
    c.fft_plans[key] = p;
    *out = p;
    return EXM_OK;
}

// the band transform of every template for one segment's weighted data (d, s2 on the device), in the
// arithmetic of C
template <typename C>
static int filter_band(Bank* bank, const double2* const d[2], const double* const s2[2], Traj* traj) {
    Context& c = ctx();
    const int64_t n_time = bank->n_time;
    const int nf = bank->n_freq;
    const int n1len = (int)(n_time / BT_N2);
    // templates per pass pair: A (sizeof(C) n_time bytes per template) in the workspace, 64 MiB by default
    // so that pass B reads it from the last-level cache.  Consecutive batches alternate between the
    // library's stream and a second one, each with its own A, so that one batch's pass A overlaps the
    // other's pass B (each pass alone leaves the CUs idle during its load and barrier phases).
    const int64_t ws = c.filter_bytes > 0 ? c.filter_bytes : (64ll << 20);
    const int chunk = (int)std::min<int64_t>(bank->n_tmpl, std::max<int64_t>(1, ws / ((int64_t)sizeof(C) * n_time)));
    const size_t abytes = (sizeof(C) * (size_t)chunk * n_time + 255) & ~(size_t)255;
    void* Abuf = nullptr;
    EXM_TRY(scratch(0, 2 * abytes, &Abuf));
    if (!c.filter_side) EXM_HIP(hipStreamCreateWithFlags(&c.filter_side, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k)
        if (!c.filter_ev[k]) EXM_HIP(hipEventCreateWithFlags(&c.filter_ev[k], hipEventDisableTiming));
    hipStream_t streams[2] = {c.stream, c.filter_side};
    C* As[2] = {(C*)Abuf, (C*)((char*)Abuf + abytes)};
    EXM_HIP(hipEventRecord(c.filter_ev[0], c.stream));          // the side stream starts after the uploads
    EXM_HIP(hipStreamWaitEvent(c.filter_side, c.filter_ev[0], 0));
    const double2* twN = nullptr;
    EXM_TRY(twiddles(n_time, &twN));
    const size_t lds_b = bt_lds_b<C>();
    EXM_HIP(hipFuncSetAttribute((const void*)k_bt_pass_b<C>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_b));
    void (*pass_a)(const float2*, const double2*, const double*, int, int, int64_t, int, const double2*, C*) = nullptr;
    switch (n1len) {
        case 16: pass_a = k_bt_pass_a<16, C>; break;
        case 32: pass_a = k_bt_pass_a<32, C>; break;
        case 64: pass_a = k_bt_pass_a<64, C>; break;
        case 128: pass_a = k_bt_pass_a<128, C>; break;
        case 256: pass_a = k_bt_pass_a<256, C>; break;
        case 512: pass_a = k_bt_pass_a<512, C>; break;
        case 1024: pass_a = k_bt_pass_a<1024, C>; break;
        case 2048: pass_a = k_bt_pass_a<2048, C>; break;
        default: EXM_CHECK(false, EXM_ERR_ARG, "band transform of %d x 512 points", n1len);
    }
    int q = 0;                                                  // batch count: stream and buffer q % 2
    for (int det = 0; det < 2; ++det) {
        float* lq = det == 0 ? traj->lqH() : traj->lqL();
        float* ph = det == 0 ? traj->phH() : traj->phL();
        for (int m0 = 0; m0 < bank->n_tmpl; m0 += chunk, ++q) {
            const int nb = std::min(chunk, bank->n_tmpl - m0);
            hipStream_t s = streams[q & 1];
            C* A = As[q & 1];
            hipLaunchKernelGGL(pass_a, dim3((unsigned)(BT_N2 * n1len / BT_TILE_A), (unsigned)nb), dim3(BT_THREADS_A), 0,
                               s, bank->d_W, d[det], s2[det], nf, bank->lo, n_time, m0, twN, A);
            EXM_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_bt_pass_b<C>, dim3((unsigned)(n1len / BT_COLS_B), (unsigned)nb), dim3(BT_THREADS_B),
                               lds_b, s, (const C*)A, n1len, n_time, m0, twN, lq, ph);
            EXM_HIP(hipGetLastError());
        }
    }
    EXM_HIP(hipEventRecord(c.filter_ev[1], c.filter_side));      // later work on the library's stream
    EXM_HIP(hipStreamWaitEvent(c.stream, c.filter_ev[1], 0));    // sees every batch
    return EXM_OK;
}

int launch_filter(Bank* bank, const double* dH, const double* dL, const double* s2H, const double* s2L, Traj* traj) {
    Context& c = ctx();
    const int64_t n_time = bank->n_time;
    const int nf = bank->n_freq;
    EXM_CHECK(bank->lo >= 0 && (int64_t)bank->lo + nf <= n_time, EXM_ERR_ARG, "band [%d, %d) outside %lld samples",
              bank->lo, bank->lo + nf, (long long)n_time);
    const bool f32 = c.filter_f32 != 0;
    double* seg = nullptr;
    EXM_TRY(scratch(2, sizeof(double) * 6 * (size_t)nf + 256, (void**)&seg));
    double2* ddH = (double2*)seg;
    double2* ddL = ddH + nf;
    double* ds2H = (double*)(ddL + nf);
    double* ds2L = ds2H + nf;
    EXM_HIP(hipMemcpyAsync(ddH, dH, sizeof(double2) * nf, hipMemcpyHostToDevice, c.stream));
    EXM_HIP(hipMemcpyAsync(ddL, dL, sizeof(double2) * nf, hipMemcpyHostToDevice, c.stream));
    EXM_HIP(hipMemcpyAsync(ds2H, s2H, sizeof(double) * nf, hipMemcpyHostToDevice, c.stream));
    EXM_HIP(hipMemcpyAsync(ds2L, s2L, sizeof(double) * nf, hipMemcpyHostToDevice, c.stream));
    ProfileScope prof("filter");
    if (c.filter_transform == 0 && band_transform_fits(n_time)) {
        const double2* d[2] = {ddH, ddL};
        const double* s2[2] = {ds2H, ds2L};
        return f32 ? filter_band<float2>(bank, d, s2, traj) : filter_band<double2>(bank, d, s2, traj);
    }
    // hipFFT: chunk of templates whose padded rows fit in the workspace (64 MiB by default: the
    // transform's input and output, 128 MiB together, stay in the 256 MB last-level cache; measured for
    // 2000 templates x 131072 samples: 10.8 ms, against 12.1 ms with 128 MiB and 12.5 ms with 32 MiB)
    const int64_t elem = f32 ? (int64_t)sizeof(float2) : (int64_t)sizeof(double2);
    const int64_t ws = c.filter_bytes > 0 ? c.filter_bytes : (64ll << 20);
    const int chunk_max = (int)std::max<int64_t>(1, ws / (elem * n_time));
    const int chunk = std::min(bank->n_tmpl, chunk_max);
    const size_t ws_bytes = (size_t)elem * chunk_max * n_time;
    void* pad = nullptr;                         // transform input: weighted bins, zero elsewhere
    void* zt = nullptr;                          // transform output
    EXM_TRY(scratch(9, ws_bytes, &pad));
    EXM_TRY(scratch(0, ws_bytes, &zt));
    if (c.filter_in != pad || c.filter_in_nt != n_time || c.filter_in_rows != chunk_max || c.filter_in_lo != bank->lo ||
        c.filter_in_nf != nf || c.filter_in_f32 != (int)f32) {
        const int64_t n8 = (int64_t)(ws_bytes / sizeof(uint2));
        hipLaunchKernelGGL(k_zero_u2, dim3((unsigned)((n8 + 255) / 256)), dim3(256), 0, c.stream, (uint2*)pad, n8);
        EXM_HIP(hipGetLastError());
        c.filter_in = pad;
        c.filter_in_nt = n_time;
        c.filter_in_rows = chunk_max;
        c.filter_in_lo = bank->lo;
        c.filter_in_nf = nf;
        c.filter_in_f32 = (int)f32;
    }
    for (int det = 0; det < 2; ++det) {
        const double2* d = det == 0 ? ddH : ddL;
        const double* s2 = det == 0 ? ds2H : ds2L;
        float* lq = det == 0 ? traj->lqH() : traj->lqL();
        float* ph = det == 0 ? traj->phH() : traj->phL();
        for (int m0 = 0; m0 < bank->n_tmpl; m0 += chunk) {
            const int nb = std::min(chunk, bank->n_tmpl - m0);
            const dim3 pgrid((unsigned)((nf + 255) / 256), (unsigned)nb);
            const dim3 grid((unsigned)((n_time + 255) / 256), (unsigned)nb);
            if (f32)
                hipLaunchKernelGGL(k_filter_pack32, pgrid, dim3(256), 0, c.stream, bank->d_W, d, s2, nf, bank->lo,
                                   n_time, m0, (float2*)pad);
            else
                hipLaunchKernelGGL(k_filter_pack, pgrid, dim3(256), 0, c.stream, bank->d_W, d, s2, nf, bank->lo,
                                   n_time, m0, (double2*)pad);
            EXM_HIP(hipGetLastError());
            // plans exist for power-of-two batches (and the full chunk) only, so that banks of any size
            // share a few plans; rows past nb in the work buffer are transformed and ignored
            int nplan = 1;
            while (nplan < nb) nplan <<= 1;
            nplan = std::min(nplan, chunk_max);
            hipfftHandle plan;
            EXM_TRY(fft_plan(n_time, nplan, f32, &plan));
            if (f32) {
                EXM_FFT(hipfftExecC2C(plan, (hipfftComplex*)pad, (hipfftComplex*)zt, HIPFFT_FORWARD));
                hipLaunchKernelGGL(k_filter_post32, grid, dim3(256), 0, c.stream, (const float2*)zt, n_time, m0, lq,
                                   ph);
            } else {
                EXM_FFT(hipfftExecZ2Z(plan, (hipfftDoubleComplex*)pad, (hipfftDoubleComplex*)zt, HIPFFT_FORWARD));
                hipLaunchKernelGGL(k_filter_post, grid, dim3(256), 0, c.stream, (const double2*)zt, n_time, m0, lq,
                                   ph);
            }
            EXM_HIP(hipGetLastError());
        }
    }
    return EXM_OK;
}

}  // namespace exm
