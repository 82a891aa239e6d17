// Matched filter of one segment against every template (src/exmarg/search/bank.py:56-78):
//
//   w_I = conj(d_I) / sig2_I,  pad[m, nu] = W[m, nu - lo] w_I[nu - lo] on [lo, lo + n_freq), 0 elsewhere
//   z   = FFT_n_time(pad)      (forward, unnormalized)
//   lq  = 1/2 ln max(|z|^2, 1e-38),  ph = atan2(Im z, Re z)
//
// Weighting, transform and modulus/phase run in float64 (hipFFT Z2Z, batched over a chunk of
// templates); the trajectories are rounded once to float32 on output.  The transform runs out of
// place from an input buffer whose zero padding outside [lo, lo + n_freq) is written once (and again
// only when the geometry changes), so each chunk writes only its n_freq weighted bins.  The float32 form
// (exm_set_filter_precision(1), part of the precision form of the batched analysis) keeps the
// weighting and the modulus/phase in float64 but transforms in float32 (C2C), as the reference's GPU
// path does (complex64, src/exmarg/search/bank.py:66-75).
#include <cmath>

#include "exm_common.h"

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

int launch_filter(Bank* bank, const double* dH, const double* dL, const double* s2H, const double* s2L, Traj* traj) {
    Context& c = ctx();
    const int64_t n_time = bank->n_time;
    const int nf = bank->n_freq;
    // chunk of templates whose padded rows fit in the workspace (64 MiB by default: the transform's input
    // and output, 128 MiB together, stay in the 256 MB last-level cache; measured for 2000 templates x
    // 131072 samples: 10.8 ms, against 12.1 ms with 128 MiB and 12.5 ms with 32 MiB)
    const bool f32 = c.filter_f32 != 0;
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
