// The no-allocation form of the plan gather (exm_plan_gather_noalloc): every scratch buffer of the
// gather comes from one statically sized device array instead of the library's pool, so that a call
// never maps device memory (for callers that cannot allocate: tight serving loops, harnesses that
// forbid it).  The array is part of the code object: each GPU the library runs on holds its own copy
// from the module's load until the process ends.  Its address on a device is found once, at exm_init,
// with that device current; host code uses that address, never the array's name (which on the host
// is only a stand-in for the symbol).
//
// Every call on a device uses the same array and finds the previous call's contents there.  Nothing
// the gather reads is expected to start at zero: each buffer is written by the kernels of the call
// before it is read (the walk buffers' padding past a walk's end is read only by lookups whose
// values are discarded).  Calls that overlap on one device race on the array.
#include "exm_common.h"

// the scratch region of the no-allocation entry point, per device
static __device__ __align__(256) unsigned char scratch[EXM_NOALLOC_SCRATCH_BYTES];

namespace exm {

constexpr int MAX_DEVICES = 64;
static void* g_noalloc_base[MAX_DEVICES] = {};

int noalloc_init(int device) {
    EXM_CHECK(device >= 0 && device < MAX_DEVICES, EXM_ERR_ARG, "device %d beyond %d", device, MAX_DEVICES);
    if (g_noalloc_base[device]) return EXM_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    EXM_CHECK(e == hipSuccess && p != nullptr, EXM_ERR_HIP, "address of the no-allocation scratch on device %d: %s",
              device, hipGetErrorString(e));
    g_noalloc_base[device] = p;
    return EXM_OK;
}

// out[order[i]] = in[i]: the table-sorted results back into the caller's read order
__global__ void k_unsort(const double* __restrict__ in, const int64_t* __restrict__ order, int64_t n,
                         double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[order[i]] = in[i];
}

// sets the context's arena for the lifetime of the scope
struct ArenaScope {
    Context::Arena arena;
    ArenaScope(void* base, size_t cap) {
        arena.base = (char*)base;
        arena.cap = cap;
        ctx().arena = &arena;
    }
    ~ArenaScope() { ctx().arena = nullptr; }
};

}  // namespace exm

using namespace exm;

extern "C" int exm_plan_gather_noalloc(exm_handle hplan, exm_handle htraj, exm_handle hstore, exm_handle htimes,
                                       double* out) {
    std::lock_guard<std::recursive_mutex> lock(ctx().mu);
    EXM_TRY(ensure_init());
    Plan* plan = get<Plan>(hplan, K_PLAN);
    Traj* traj = get<Traj>(htraj, K_TRAJ);
    Store* st = get<Store>(hstore, K_STORE);
    EXM_CHECK(plan && traj && st, EXM_ERR_HANDLE, "invalid plan, trajectory or store handle");
    EXM_CHECK(out != nullptr, EXM_ERR_ARG, "out is NULL");
    Times* tm = nullptr;
    EXM_TRY(check_times(traj, htimes, &tm));
    EXM_CHECK(plan->n_tmpl == traj->n_tmpl, EXM_ERR_ARG, "plan is for %d templates, trajectories for %d", plan->n_tmpl,
              traj->n_tmpl);
    void* base = g_noalloc_base[ctx().device];
    EXM_CHECK(base != nullptr, EXM_ERR_STATE, "no-allocation scratch not set up on device %d", ctx().device);
    const int64_t njobs = plan->js.njobs;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    // the results (table-sorted), the caller's order and the results in that order, then the gather's
    // own scratch: all of it must fit, checked before anything is launched
    const size_t own = 2 * al(sizeof(double) * (size_t)njobs) + al(sizeof(int64_t) * (size_t)njobs);
    const size_t need = own + gather_scratch_bytes(traj, st, plan->js, tm);
    EXM_CHECK(need <= (size_t)EXM_NOALLOC_SCRATCH_BYTES, EXM_ERR_NOMEM,
              "gather of %lld reads needs %zu bytes of scratch, above the no-allocation maximum of %zu bytes",
              (long long)njobs, need, (size_t)EXM_NOALLOC_SCRATCH_BYTES);
    char* b = (char*)base;
    double* d_sorted = (double*)b;
    int64_t* d_order = (int64_t*)(b + al(sizeof(double) * (size_t)njobs));
    double* d_out = (double*)(b + al(sizeof(double) * (size_t)njobs) + al(sizeof(int64_t) * (size_t)njobs));
    ArenaScope arena(b + own, (size_t)EXM_NOALLOC_SCRATCH_BYTES - own);
    const hipStream_t s = ctx().stream;
    EXM_HIP(hipMemcpyAsync(d_order, plan->js.order.data(), sizeof(int64_t) * (size_t)njobs, hipMemcpyHostToDevice, s));
    EXM_TRY(launch_gather(traj, st, plan->js, tm, d_sorted));
    if (njobs > 0)
        hipLaunchKernelGGL(k_unsort, dim3((unsigned)((njobs + 255) / 256)), dim3(256), 0, s, d_sorted, d_order, njobs,
                           d_out);
    EXM_HIP(hipGetLastError());
    EXM_HIP(hipMemcpyAsync(out, d_out, sizeof(double) * (size_t)njobs, hipMemcpyDeviceToHost, s));
    EXM_HIP(hipStreamSynchronize(s));
    return EXM_OK;
}
