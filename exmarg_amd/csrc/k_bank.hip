// Template factors on the device: W[m, i] = (a / D0) e^{-i psi} of a TaylorF2 waveform on the
// contiguous bins nu = lo + i (f = nu / T), and u[m] = sum_i (a / D0)^2 / sig2_H[i].
//
// This is the device form of the host TaylorF2 (exmarg_amd.synth.taylorf2 / template_factors), the
// waveform the bank carries in place of the reference's IMRPhenomXAS call
// (src/exmarg/search/bank.py:44-53): leading-order SPA amplitude, 3.5PN point-particle phase with
// the 1.5PN, 2.5PN and 3.5PN spin-orbit and 2PN spin-spin terms (Buonanno, Iyer, Ochsner, Pan &
// Sathyaprakash 2009, PRD 80 084043, Eq. 3.18 and Table I; spin terms of Arun, Buonanno, Faye &
// Ochsner 2009, PRD 79 104023 and Bohe, Marsat & Blanchet 2013), zero above the Schwarzschild ISCO
// frequency.  Everything is float64 up to the single rounding of W to complex64; u is a float64
// block reduction.  One workgroup per template.

#include "exm_common.h"

namespace exm {

namespace {

constexpr int BANK_THREADS = 256;
constexpr double PI = 3.141592653589793;
constexpr double MSUN_S = 4.925490947641267e-06;                    // G Msun / c^3 (s)
constexpr double MPC_S = 3.0856775814913674e22 / 299792458.0;        // one megaparsec in light-seconds
constexpr double EULER_GAMMA = 0.5772156649015329;

struct Tf2 {
    double eta, Ms, fisco, ampc, p2, p3, p4, p5c, p6, p6l, p7;
};

__device__ Tf2 tf2_coefficients(double M, double q, double chi, double d0_mpc) {
    Tf2 c;
    const double eta = q / ((1.0 + q) * (1.0 + q));
    const double delta = (1.0 - q) / (1.0 + q);
    const double chis = chi, chia = 0.0;
    c.eta = eta;
    c.Ms = M * MSUN_S;
    const double Mc = c.Ms * pow(eta, 0.6);
    c.fisco = 1.0 / (pow(6.0, 1.5) * PI * c.Ms);
    c.ampc = sqrt(5.0 / 24.0) * pow(PI, -2.0 / 3.0) * pow(Mc, 5.0 / 6.0) / MPC_S / d0_mpc;
    c.p2 = 3715.0 / 756.0 + 55.0 * eta / 9.0;
    c.p3 = -16.0 * PI + (113.0 / 3.0) * delta * chia + (113.0 / 3.0 - 76.0 * eta / 3.0) * chis;
    c.p4 = 15293365.0 / 508032.0 + 27145.0 * eta / 504.0 + 3085.0 * eta * eta / 72.0 +
           (-405.0 / 8.0 + 200.0 * eta) * chia * chia - 405.0 / 4.0 * delta * chia * chis +
           (-405.0 / 8.0 + 5.0 * eta / 2.0) * chis * chis;
    c.p5c = PI * (38645.0 / 756.0 - 65.0 * eta / 9.0) -
            (732985.0 / 2268.0 - 24260.0 * eta / 81.0 - 340.0 * eta * eta / 9.0) * chis -
            (732985.0 / 2268.0 + 140.0 * eta / 9.0) * delta * chia;
    c.p6 = 11583231236531.0 / 4694215680.0 - 640.0 * PI * PI / 3.0 - 6848.0 * EULER_GAMMA / 21.0 +
           (-15737765635.0 / 3048192.0 + 2255.0 * PI * PI / 12.0) * eta + 76055.0 * eta * eta / 1728.0 -
           127825.0 * eta * eta * eta / 1296.0 +
           PI * (2270.0 / 3.0 * delta * chia + (2270.0 / 3.0 - 520.0 * eta) * chis);
    c.p6l = -6848.0 / 21.0;
    c.p7 = PI * (77096675.0 / 254016.0 + 378515.0 * eta / 1512.0 - 74045.0 * eta * eta / 756.0) +
           (-25150083775.0 / 3048192.0 + 10566655595.0 * eta / 762048.0 - 1042165.0 * eta * eta / 3024.0 +
            5345.0 * eta * eta * eta / 36.0) * chis +
           delta * (-25150083775.0 / 3048192.0 + 26804935.0 * eta / 6048.0 - 1985.0 * eta * eta / 48.0) * chia;
    return c;
}

__global__ __launch_bounds__(BANK_THREADS) void k_taylorf2(const double* __restrict__ M, const double* __restrict__ q,
                                                           const double* __restrict__ chi, int n_freq, int lo,
                                                           double T, double d0_mpc, const double* __restrict__ sig2H,
                                                           float2* __restrict__ W, double* __restrict__ u) {
    __shared__ Tf2 c;
    __shared__ double part[BANK_THREADS / 64];
    const int m = blockIdx.x, tid = threadIdx.x;
    if (tid == 0) c = tf2_coefficients(M[m], q[m], chi[m], d0_mpc);
    __syncthreads();
    double acc = 0.0;
    float2* Wm = W + (int64_t)m * n_freq;
    for (int i = tid; i < n_freq; i += BANK_THREADS) {
        const double f = (double)(lo + i) / T;
        float2 w = make_float2(0.0f, 0.0f);
        if (f > 0.0 && f <= c.fisco) {
            const double v = pow(PI * c.Ms * f, 1.0 / 3.0);
            const double lv = log(v);
            const double v2 = v * v, v3 = v2 * v, v4 = v3 * v, v5 = v4 * v, v6 = v5 * v, v7 = v6 * v;
            const double series = 1.0 + c.p2 * v2 + c.p3 * v3 + c.p4 * v4 + c.p5c * (1.0 + 3.0 * lv) * v5 +
                                  (c.p6 + c.p6l * log(4.0 * v)) * v6 + c.p7 * v7;
            const double psi = 3.0 / (128.0 * c.eta * v5) * series;
            const double a = c.ampc * pow(f, -7.0 / 6.0);
            double sn, cs;
            sincos(-psi, &sn, &cs);
            w = make_float2((float)(a * cs), (float)(a * sn));
            acc += a * a / sig2H[i];
        }
        Wm[i] = w;
    }
    // float64 sum of the template power over the block
    for (int d = 32; d > 0; d >>= 1) acc += __shfl_down(acc, d, 64);
    if ((tid & 63) == 0) part[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int k = 0; k < BANK_THREADS / 64; ++k) t += part[k];
        u[m] = t;
    }
}

}  // namespace

int launch_taylorf2(const double* d_M, const double* d_q, const double* d_chi, int n_tmpl, int n_freq, int lo,
                    double T, double d0_mpc, const double* d_sig2H, float2* d_W, double* d_u) {
    hipLaunchKernelGGL(k_taylorf2, dim3(n_tmpl), dim3(BANK_THREADS), 0, ctx().stream, d_M, d_q, d_chi, n_freq, lo, T,
                       d0_mpc, d_sig2H, d_W, d_u);
    EXM_HIP(hipGetLastError());
    return EXM_OK;
}

}  // namespace exm
