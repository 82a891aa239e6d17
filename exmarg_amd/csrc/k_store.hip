// Table residency: reorder host-layout tables into the device layout, and the
// device generator of synthetic store values.
//
// Host (file) layout of one table: [lag][phase][beta][Delta]
//   (scripts/build_tables/build_bd.py:64,79 writes it; src/exmarg/search/store.py:110-120 stacks it).
// Device layout: [table][lag][beta][phase][Delta], so that one beta row of every phase bin is one
// contiguous span of nphif * nd floats: the unit the gather stages through LDS.
#include "exm_common.h"

namespace exm {

// one thread block per (lag, beta, phase) row of nd floats
__global__ void k_transpose_in(const float* __restrict__ src, float* __restrict__ dst, int nphif, int nb, int nd) {
    const int64_t row = blockIdx.x;                 // device order: (r * nb + b) * nphif + c
    const int c = (int)(row % nphif);
    const int64_t rb = row / nphif;
    const int b = (int)(rb % nb);
    const int64_t r = rb / nb;
    const float* s = src + ((r * nphif + c) * nb + b) * (int64_t)nd;
    float* d = dst + row * (int64_t)nd;
    for (int i = threadIdx.x; i < nd; i += blockDim.x) d[i] = s[i];
}

__global__ void k_transpose_out(const float* __restrict__ src, float* __restrict__ dst, int nphif, int nb, int nd) {
    const int64_t row = blockIdx.x;
    const int c = (int)(row % nphif);
    const int64_t rb = row / nphif;
    const int b = (int)(rb % nb);
    const int64_t r = rb / nb;
    const float* s = src + row * (int64_t)nd;
    float* d = dst + ((r * nphif + c) * nb + b) * (int64_t)nd;
    for (int i = threadIdx.x; i < nd; i += blockDim.x) d[i] = s[i];
}

__device__ __forceinline__ uint64_t splitmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// synthetic value at reference-layout flat index idx (integer recipe, exact in float32)
__global__ void k_fill_synthetic(float* __restrict__ dst, const int32_t* __restrict__ tau, uint64_t seed,
                                 int64_t node, int kind, int nphif, int nb, int nd) {
    const int64_t row = blockIdx.x;
    const int c = (int)(row % nphif);
    const int64_t rb = row / nphif;
    const int b = (int)(rb % nb);
    const int64_t r = rb / nb;
    const int64_t ref_row = (r * nphif + c) * nb + b;
    const uint64_t key0 = (seed * 0x100000001B3ull) ^ ((uint64_t)node << 40);
    int64_t ta = tau[r];
    ta = ta < 0 ? -ta : ta;
    const int64_t bb = b - 56 > 0 ? b - 56 : 0;
    for (int d = threadIdx.x; d < nd; d += blockDim.x) {
        const int64_t idx = ref_row * nd + d;
        const uint64_t h = splitmix(key0 ^ (uint64_t)idx);
        float v;
        if (kind == 1) {
            const int64_t hv = (int64_t)(h >> 41);
            v = (float)((double)(hv - (1 << 22)) / (double)(1 << 22));
        } else {
            const int64_t hv = (int64_t)(h >> 55);
            int64_t dd = d - nd / 2;
            dd = dd < 0 ? -dd : dd;
            const int64_t num = 8 * bb * (64 - c) - 64 * dd - 32 * ta + hv;
            v = (float)((double)num / 4096.0);
        }
        dst[row * (int64_t)nd + d] = v;
    }
}

int launch_store_transpose_in(Store* st, int k, const float* d_src) {
    const int64_t rows = (int64_t)st->ntau * st->nb * st->nphif;
    hipLaunchKernelGGL(k_transpose_in, dim3((unsigned)rows), dim3(256), 0, ctx().stream, d_src,
                       st->d_T + (int64_t)k * st->tstride, st->nphif, st->nb, st->nd);
    EXM_HIP(hipGetLastError());
    return EXM_OK;
}

int launch_store_transpose_out(Store* st, int k, float* d_dst) {
    const int64_t rows = (int64_t)st->ntau * st->nb * st->nphif;
    hipLaunchKernelGGL(k_transpose_out, dim3((unsigned)rows), dim3(256), 0, ctx().stream,
                       st->d_T + (int64_t)k * st->tstride, d_dst, st->nphif, st->nb, st->nd);
    EXM_HIP(hipGetLastError());
    return EXM_OK;
}

int launch_fill_synthetic(Store* st, int k, uint64_t seed, int64_t node, int kind) {
    const int64_t rows = (int64_t)st->ntau * st->nb * st->nphif;
    hipLaunchKernelGGL(k_fill_synthetic, dim3((unsigned)rows), dim3(256), 0, ctx().stream,
                       st->d_T + (int64_t)k * st->tstride, st->d_tau, seed, node, kind, st->nphif, st->nb, st->nd);
    EXM_HIP(hipGetLastError());
    return EXM_OK;
}

// plain copy of n float4 values: the kernel of known traffic (16 n bytes read, 16 n written) that
// calibrates the memory-traffic counters
__global__ void k_copy_calibration(float4* __restrict__ dst, const float4* __restrict__ src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int launch_copy_calibration(float4* dst, const float4* src, int64_t n) {
    hipLaunchKernelGGL(k_copy_calibration, dim3(4096), dim3(256), 0, ctx().stream, dst, src, n);
    EXM_HIP(hipGetLastError());
    return EXM_OK;
}

}  // namespace exm
