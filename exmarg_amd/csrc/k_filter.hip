// Matched filter of one segment against every template (src/exmarg/search/bank.py:56-78):
//
//   w_I = conj(d_I) / sig2_I,  pad[m, nu] = W[m, nu - lo] w_I[nu - lo] on [lo, lo + n_freq), 0 elsewhere
//   z   = FFT_n_time(pad)      (forward, unnormalized)
//   lq  = 1/2 ln max(|z|^2, 1e-38),  ph = atan2(Im z, Re z)
//
// Weighting, transform and modulus/phase run in float64 (hipFFT Z2Z, batched over a chunk of
// templates); the trajectories are rounded once to float32 on output.
#include <cmath>

#include "exm_common.h"

#pragma clang fp contract(off)

namespace exm {

__global__ void k_filter_pack(const float2* __restrict__ W, const double2* __restrict__ d,
                              const double* __restrict__ s2, int n_freq, int lo, int64_t n_time, int m0,
                              double2* __restrict__ pad) {
    const int j = blockIdx.y;
    const int64_t nu = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (nu >= n_time) return;
    double2 out = make_double2(0.0, 0.0);
    const int64_t f = nu - lo;
    if (f >= 0 && f < n_freq) {
        const float2 wt = W[(int64_t)(m0 + j) * n_freq + f];
        const double2 dd = d[f];
        const double s = s2[f];
        // w = conj(d) / sig2, rounded as numpy rounds a complex divided by a real: times 1/sig2
        const double inv = 1.0 / s;
        const double wr = dd.x * inv, wi = (-dd.y) * inv;
        const double ar = (double)wt.x, ai = (double)wt.y;
        out.x = ar * wr - ai * wi;
        out.y = ar * wi + ai * wr;
    }
    pad[(int64_t)j * n_time + nu] = out;
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

static int fft_plan(int64_t n_time, int batch, hipfftHandle* out) {
    Context& c = ctx();
    const uint64_t key = ((uint64_t)n_time << 20) | (uint64_t)batch;
    auto it = c.fft_plans.find(key);
    if (it != c.fft_plans.end()) {
        *out = it->second;
        return EXM_OK;
    }
    hipfftHandle p;
    EXM_FFT(hipfftCreate(&p));
    int n[1] = {(int)n_time};
    size_t ws = 0;
    EXM_FFT(hipfftMakePlanMany(p, 1, n, nullptr, 1, (int)n_time, nullptr, 1, (int)n_time, HIPFFT_Z2Z, batch, &ws));
    EXM_FFT(hipfftSetStream(p, c.stream));
    c.fft_plans[key] = p;
    *out = p;
    return EXM_OK;
}

int launch_filter(Bank* bank, const double* dH, const double* dL, const double* s2H, const double* s2L, Traj* traj) {
    Context& c = ctx();
    const int64_t n_time = bank->n_time;
    const int nf = bank->n_freq;
    // chunk of templates whose complex128 padded rows fit in 1 GiB of workspace
    const int chunk_max = (int)std::max<int64_t>(1, (1ll << 30) / (16 * n_time));
    const int chunk = std::min(bank->n_tmpl, chunk_max);
    double2* pad = nullptr;
    EXM_TRY(scratch(0, sizeof(double2) * (size_t)chunk_max * n_time, (void**)&pad));
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
            dim3 grid((unsigned)((n_time + 255) / 256), (unsigned)nb);
            hipLaunchKernelGGL(k_filter_pack, grid, dim3(256), 0, c.stream, bank->d_W, d, s2, nf, bank->lo, n_time,
                               m0, pad);
            EXM_HIP(hipGetLastError());
            // plans exist for power-of-two batches (and the full chunk) only, so that banks of any size
            // share a few plans; rows past nb in the work buffer are transformed and ignored
            int nplan = 1;
            while (nplan < nb) nplan <<= 1;
            nplan = std::min(nplan, chunk_max);
            hipfftHandle plan;
            EXM_TRY(fft_plan(n_time, nplan, &plan));
            EXM_FFT(hipfftExecZ2Z(plan, (hipfftDoubleComplex*)pad, (hipfftDoubleComplex*)pad, HIPFFT_FORWARD));
            hipLaunchKernelGGL(k_filter_post, grid, dim3(256), 0, c.stream, pad, n_time, m0, lq, ph);
            EXM_HIP(hipGetLastError());
        }
    }
    return EXM_OK;
}

}  // namespace exm
