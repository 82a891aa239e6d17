// Table residency: reorder host-layout tables into the device layout, and the
// device generator of synthetic store values.
//
// Host (file) layout of one table: [lag][phase][beta][Delta]
//   (scripts/build_tables/build_bd.py:64,79 writes it; src/exmarg/search/store.py:110-120 stacks it).
// Device layout: [table][lag][beta][Delta][phase], so that one beta row (every Delta column and phase
// bin) is one contiguous span of nd * nphif floats, the unit the gather stages through LDS, and the two
// phase bins a lookup interpolates between are adjacent floats.
#include "exm_common.h"

namespace exm {

constexpr int TR_DT = 64;        // Delta columns per transpose block

// one block per (lag, beta, 64 Delta columns): the nphif x 64 tile through LDS, read along Delta from
// the host layout and written along the phase bins of the device layout
__global__ __launch_bounds__(256) void k_transpose_in(const float* __restrict__ src, float* __restrict__ dst,
                                                      int nphif, int nb, int nd) {
    __shared__ float tile[NPHIF][TR_DT + 1];
    const int nc = (nd + TR_DT - 1) / TR_DT;
    const int64_t rb = blockIdx.x / nc;             // r * nb + b
    const int d0 = (int)(blockIdx.x - rb * nc) * TR_DT;
    const int b = (int)(rb % nb);
    const int64_t r = rb / nb;
    const int w = min(TR_DT, nd - d0);
    for (int i = threadIdx.x; i < nphif * TR_DT; i += blockDim.x) {
        const int c = i / TR_DT, dl = i - c * TR_DT;
        if (dl < w) tile[c][dl] = src[((r * nphif + c) * nb + b) * (int64_t)nd + d0 + dl];
    }
    __syncthreads();
    float* d = dst + (rb * nd + d0) * (int64_t)nphif;
    for (int i = threadIdx.x; i < w * nphif; i += blockDim.x) {
        const int dl = i / nphif, c = i - dl * nphif;
        d[i] = tile[c][dl];
    }
}

__global__ __launch_bounds__(256) void k_transpose_out(const float* __restrict__ src, float* __restrict__ dst,
                                                       int nphif, int nb, int nd) {
    __shared__ float tile[NPHIF][TR_DT + 1];
    const int nc = (nd + TR_DT - 1) / TR_DT;
    const int64_t rb = blockIdx.x / nc;
    const int d0 = (int)(blockIdx.x - rb * nc) * TR_DT;
    const int b = (int)(rb % nb);
    const int64_t r = rb / nb;
    const int w = min(TR_DT, nd - d0);
    const float* s = src + (rb * nd + d0) * (int64_t)nphif;
    for (int i = threadIdx.x; i < w * nphif; i += blockDim.x) {
        const int dl = i / nphif, c = i - dl * nphif;
        tile[c][dl] = s[i];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nphif * TR_DT; i += blockDim.x) {
        const int c = i / TR_DT, dl = i - c * TR_DT;
        if (dl < w) dst[((r * nphif + c) * nb + b) * (int64_t)nd + d0 + dl] = tile[c][dl];
    }
}

__device__ __forceinline__ uint64_t splitmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// synthetic value at reference-layout flat index idx (integer recipe, exact in float32); one block per
// (lag, beta) row, written in the device order
__global__ void k_fill_synthetic(float* __restrict__ dst, const int32_t* __restrict__ tau, uint64_t seed,
                                 int64_t node, int kind, int nphif, int nb, int nd) {
    const int64_t rb = blockIdx.x;                  // r * nb + b
    const int b = (int)(rb % nb);
    const int64_t r = rb / nb;
    const uint64_t key0 = (seed * 0x100000001B3ull) ^ ((uint64_t)node << 40);
    int64_t ta = tau[r];
    ta = ta < 0 ? -ta : ta;
    const int64_t bb = b - 56 > 0 ? b - 56 : 0;
    float* row = dst + rb * (int64_t)nd * nphif;
    for (int i = threadIdx.x; i < nd * nphif; i += blockDim.x) {
        const int d = i / nphif, c = i - d * nphif;
        const int64_t idx = ((r * nphif + c) * nb + b) * (int64_t)nd + d;
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
        row[i] = v;
    }
}

int launch_store_transpose_in(Store* st, int k, const float* d_src) {
    EXM_CHECK(st->nphif <= NPHIF, EXM_ERR_ARG, "tables of %d phase bins exceed %d", st->nphif, NPHIF);
    const int64_t blocks = (int64_t)st->ntau * st->nb * ((st->nd + TR_DT - 1) / TR_DT);
    hipLaunchKernelGGL(k_transpose_in, dim3((unsigned)blocks), dim3(256), 0, ctx().stream, d_src,
                       st->d_T + (int64_t)k * st->tstride, st->nphif, st->nb, st->nd);
    EXM_HIP(hipGetLastError());
    return EXM_OK;
}

int launch_store_transpose_out(Store* st, int k, float* d_dst) {
    EXM_CHECK(st->nphif <= NPHIF, EXM_ERR_ARG, "tables of %d phase bins exceed %d", st->nphif, NPHIF);
    const int64_t blocks = (int64_t)st->ntau * st->nb * ((st->nd + TR_DT - 1) / TR_DT);
    hipLaunchKernelGGL(k_transpose_out, dim3((unsigned)blocks), dim3(256), 0, ctx().stream,
                       st->d_T + (int64_t)k * st->tstride, d_dst, st->nphif, st->nb, st->nd);
    EXM_HIP(hipGetLastError());
    return EXM_OK;
}

int launch_fill_synthetic(Store* st, int k, uint64_t seed, int64_t node, int kind) {
    const int64_t rows = (int64_t)st->ntau * st->nb;
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
