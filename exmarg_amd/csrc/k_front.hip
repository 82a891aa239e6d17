// Time-domain front end of a stream of segments.
//
//   exm_strain_segments  the frequency-domain segment of each row of a time series: d = dt rfft(x) on the
//                        bins nu0 .. nu0 + n_freq - 1 (the segment convention of src/exmarg/pp.py:38-72,
//                        where white noise of one-sided density S gives E|d|^2 = T S / 2)
//   exm_strain_psd       the noise variance per bin estimated from n_blocks neighbouring blocks: the mean
//                        of their periodograms |d|^2 times a scale, N/(N-2) for the estimate of
//                        notebooks/07_estimated_psd.ipynb cell 5
//
// Both transform in float64 with hipFFT (D2Z, batched over rows, in chunks that bound the work memory) and
// read only the band out of the transform.  The periodogram mean adds the blocks in their order, one
// thread per bin, so the estimate does not depend on the chunking.
#include "exm_common.h"

#pragma clang fp contract(off)

namespace exm {

// rows of one transform chunk: input, transform and the rows' band bins within about 256 MiB
static int front_rows(int64_t n) {
    const int64_t per = 8 * n + 16 * (n / 2 + 1);
    const int64_t r = ((int64_t)256 << 20) / per;
    return (int)std::max<int64_t>(1, std::min<int64_t>(r, 4096));
}

static int d2z_plan(int64_t n, int batch, hipfftHandle* out) {
    Context& c = ctx();
    // keys of the complex transforms of the matched filter stay below 2^63
    const uint64_t key = (1ull << 63) | ((uint64_t)n << 21) | (uint64_t)batch;
    auto it = c.fft_plans.find(key);
    if (it != c.fft_plans.end()) {
        *out = it->second;
        return EXM_OK;
    }
    hipfftHandle p;
    EXM_FFT(hipfftCreate(&p));
    int dims[1] = {(int)n};
    size_t ws = 0;
    const hipfftResult r = hipfftMakePlanMany(p, 1, dims, nullptr, 1, (int)n, nullptr, 1, (int)(n / 2 + 1),
                                              HIPFFT_D2Z, batch, &ws);
    if (r != HIPFFT_SUCCESS) {
        (void)hipfftDestroy(p);
        EXM_CHECK(false, EXM_ERR_HIP, "hipfftMakePlanMany(D2Z, n %lld, batch %d): status %d", (long long)n, batch, (int)r);
    }
    c.fft_plans[key] = p;
    *out = p;
    return EXM_OK;
}

// d[row, k] = X[row, nu0 + k] / fs
__global__ void k_front_band(const double2* __restrict__ X, int64_t xstride, int nu0, int n_freq, double inv_fs,
                             double2* __restrict__ d) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = blockIdx.y;
    if (k >= n_freq) return;
    const double2 v = X[(int64_t)row * xstride + nu0 + k];
    d[(int64_t)row * n_freq + k] = make_double2(v.x * inv_fs, v.y * inv_fs);
}

// acc[k] += sum over the chunk's rows, in order, of |X[row, nu0 + k] / fs|^2
__global__ void k_front_power(const double2* __restrict__ X, int64_t xstride, int rows, int nu0, int n_freq,
                              double inv_fs, double* __restrict__ acc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_freq) return;
    double s = acc[k];
    for (int r = 0; r < rows; ++r) {
        const double2 v = X[(int64_t)r * xstride + nu0 + k];
        const double re = v.x * inv_fs, im = v.y * inv_fs;
        s += re * re + im * im;
    }
    acc[k] = s;
}

__global__ void k_front_scale(double* __restrict__ a, int n, double f) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] *= f;
}

__global__ void k_front_zero(double* __restrict__ a, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = 0.0;
}

static int front_check(int64_t rows, int64_t n, double fs, int nu0, int n_freq) {
    EXM_CHECK(rows > 0 && n >= 2 && n <= ((int64_t)1 << 30), EXM_ERR_ARG,
              "need rows > 0 and 2 <= n <= 2^30 samples (got %lld rows of %lld)", (long long)rows, (long long)n);
    EXM_CHECK(fs > 0.0 && std::isfinite(fs), EXM_ERR_ARG, "sampling rate must be positive (got %g)", fs);
    EXM_CHECK(nu0 >= 0 && n_freq > 0 && (int64_t)nu0 + n_freq - 1 <= n / 2, EXM_ERR_ARG,
              "bins %d .. %lld lie outside 0 .. %lld (the Nyquist bin of %lld samples)", nu0,
              (long long)nu0 + n_freq - 1, (long long)(n / 2), (long long)n);
    return EXM_OK;
}

// one pass over the rows in chunks: upload, transform, then per chunk either the band bins (d != null,
// copied back) or the periodogram sum (acc != null, on the device)
static int front_pass(const double* x, int64_t rows, int64_t n, double fs, int nu0, int n_freq, double2* d,
                      double* acc) {
    const hipStream_t s = ctx().stream;
    const int chunk = (int)std::min<int64_t>(rows, front_rows(n));
    const int64_t nx = n / 2 + 1;
    const size_t in_bytes = sizeof(double) * (size_t)chunk * n;
    const size_t x_bytes = sizeof(double2) * (size_t)chunk * nx;
    const size_t d_bytes = d ? sizeof(double2) * (size_t)chunk * n_freq : 0;
    const size_t al_in = (in_bytes + 255) & ~(size_t)255, al_x = (x_bytes + 255) & ~(size_t)255;
    const size_t total = al_in + al_x + d_bytes;
    void* buf = nullptr;
    EXM_TRY(dev_alloc(&buf, total));
    double* din = (double*)buf;
    double2* dX = (double2*)((char*)buf + al_in);
    double2* dd = (double2*)((char*)buf + al_in + al_x);
    int rc = EXM_OK;
    for (int64_t r0 = 0; r0 < rows && rc == EXM_OK; r0 += chunk) {
        const int nr = (int)std::min<int64_t>(chunk, rows - r0);
        hipfftHandle plan;
        rc = d2z_plan(n, nr, &plan);
        if (rc != EXM_OK) break;
        if (hipMemcpyAsync(din, x + r0 * n, sizeof(double) * (size_t)nr * n, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipfftSetStream(plan, s) != HIPFFT_SUCCESS ||
            hipfftExecD2Z(plan, (hipfftDoubleReal*)din, (hipfftDoubleComplex*)dX) != HIPFFT_SUCCESS) {
            set_error("strain transform of rows %lld .. %lld failed", (long long)r0, (long long)(r0 + nr - 1));
            rc = EXM_ERR_HIP;
            break;
        }
        const unsigned gx = (unsigned)((n_freq + 255) / 256);
        if (d) {
            hipLaunchKernelGGL(k_front_band, dim3(gx, nr), dim3(256), 0, s, dX, nx, nu0, n_freq, 1.0 / fs, dd);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(d + r0 * n_freq, dd, sizeof(double2) * (size_t)nr * n_freq, hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                set_error("band of rows %lld .. %lld failed", (long long)r0, (long long)(r0 + nr - 1));
                rc = EXM_ERR_HIP;
            }
        } else {
            hipLaunchKernelGGL(k_front_power, dim3(gx), dim3(256), 0, s, dX, nx, nr, nu0, n_freq, 1.0 / fs, acc);
            if (hipGetLastError() != hipSuccess) {
                set_error("periodogram sum of rows %lld .. %lld failed", (long long)r0, (long long)(r0 + nr - 1));
                rc = EXM_ERR_HIP;
            }
            // the next chunk's upload and transform overwrite the input and the transform only after this
            // chunk's transform and sum have read them: all of it runs in order on the library stream
        }
    }
    dev_free(buf, total);
    return rc;
}

}  // namespace exm

using namespace exm;

extern "C" int exm_strain_segments(const double* x, int n_rows, int64_t n, double fs, int nu0, int n_freq,
                                   double* d) {
    std::lock_guard<std::recursive_mutex> lock(ctx().mu);
    EXM_TRY(ensure_init());
    EXM_CHECK(x != nullptr && d != nullptr, EXM_ERR_ARG, "NULL argument");
    EXM_TRY(front_check(n_rows, n, fs, nu0, n_freq));
    return front_pass(x, n_rows, n, fs, nu0, n_freq, (double2*)d, nullptr);
}

extern "C" int exm_strain_psd(const double* x, int n_blocks, int64_t n, double fs, int nu0, int n_freq, double scale,
                              double* sig2) {
    std::lock_guard<std::recursive_mutex> lock(ctx().mu);
    EXM_TRY(ensure_init());
    EXM_CHECK(x != nullptr && sig2 != nullptr, EXM_ERR_ARG, "NULL argument");
    EXM_TRY(front_check(n_blocks, n, fs, nu0, n_freq));
    EXM_CHECK(scale > 0.0 && std::isfinite(scale), EXM_ERR_ARG, "scale must be positive (got %g)", scale);
    const hipStream_t s = ctx().stream;
    double* acc = nullptr;
    EXM_TRY(dev_alloc((void**)&acc, sizeof(double) * (size_t)n_freq));
    const unsigned gx = (unsigned)((n_freq + 255) / 256);
    hipLaunchKernelGGL(k_front_zero, dim3(gx), dim3(256), 0, s, acc, n_freq);
    int rc = hipGetLastError() == hipSuccess ? EXM_OK : EXM_ERR_HIP;
    if (rc != EXM_OK) set_error("clearing the periodogram sum failed");
    if (rc == EXM_OK) rc = front_pass(x, n_blocks, n, fs, nu0, n_freq, nullptr, acc);
    if (rc == EXM_OK) {
        hipLaunchKernelGGL(k_front_scale, dim3(gx), dim3(256), 0, s, acc, n_freq, scale / (double)n_blocks);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(sig2, acc, sizeof(double) * (size_t)n_freq, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            set_error("periodogram mean failed");
            rc = EXM_ERR_HIP;
        }
    }
    dev_free(acc, sizeof(double) * (size_t)n_freq);
    return rc;
}
