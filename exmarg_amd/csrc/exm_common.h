// Shared definitions of the exmarg HIP library: error state, the device
// context, object types behind the opaque handles, and device helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/exmarg_hip.h"

namespace exm {

constexpr int NPHI = 96;
constexpr int NPHIF = NPHI / 2 + 1;

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
const char* last_error();

#define EXM_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            exm::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                           __LINE__);                                                       \
            return EXM_ERR_HIP;                                                             \
        }                                                                                   \
    } while (0)

#define EXM_FFT(call)                                                              \
    do {                                                                           \
        hipfftResult r_ = (call);                                                  \
        if (r_ != HIPFFT_SUCCESS) {                                                \
            exm::set_error("%s failed: hipfft status %d (%s:%d)", #call, (int)r_, \
                           __FILE__, __LINE__);                                    \
            return EXM_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

#define EXM_CHECK(cond, code, ...)     \
    do {                               \
        if (!(cond)) {                 \
            exm::set_error(__VA_ARGS__); \
            return code;               \
        }                              \
    } while (0)

#define EXM_TRY(expr)            \
    do {                         \
        int rc_ = (expr);        \
        if (rc_ != EXM_OK) return rc_; \
    } while (0)

// ---------------------------------------------------------------- objects
enum Kind : uint32_t { K_STORE = 1, K_BANK = 2, K_TRAJ = 3, K_TIMES = 4, K_PLAN = 5, K_BATCH = 6, K_GRID = 7 };

struct Object {
    uint32_t kind;
    virtual ~Object() {}
};

// Tables of every node in one device allocation, layout [table][lag][beta][Delta][phase]
// (float32): one beta row of all phases is contiguous, nphif * nd floats.  The allocation ends with
// TABLE_PAD_FLOATS spare floats so that row copies in whole 64-float blocks stay inside it.
constexpr int TABLE_PAD_FLOATS = 256;
// Spare bytes after the walk buffers (walked ln|z_L| and the lag stream): the band kernel's wave form
// reads every piece as PN consecutive records from one address, up to 7 records past a read's range.
constexpr size_t WALK_PAD_BYTES = 64;
// Largest key buffer of one selection launch (8 bytes per template and time): the templates of one launch
// are as many as fit, so a bank of a few thousand templates is selected in one launch at every length.
constexpr int64_t SEL_KEY_BYTES = (int64_t)1 << 34;
struct Store : Object {
    int ntab = 0, ntau = 0, nphif = 0, nb = 0, nd = 0;
    int64_t tstride = 0;   // floats per table
    std::vector<int64_t> tau;
    double beta0 = 0, dbeta = 0, delta0 = 0, ddelta = 0;
    float* d_T = nullptr;      // ntab * tstride floats + TABLE_PAD_FLOATS
    int32_t* d_tau = nullptr;  // lags (samples)
    size_t table_bytes() const { return sizeof(float) * ((size_t)tstride * ntab + TABLE_PAD_FLOATS); }
    ~Store();
};

struct Bank : Object {
    int n_tmpl = 0, n_freq = 0, lo = 0;
    int64_t n_time = 0;
    float2* d_W = nullptr;  // (n_tmpl, n_freq) complex64
    ~Bank();
};

// Matched-filter outputs of one segment: four (n_tmpl, n_time) float32 arrays.
struct Traj : Object {
    int n_tmpl = 0;
    int64_t n_time = 0;
    float* d_buf = nullptr;  // lqH | phH | lqL | phL
    float* lqH() const { return d_buf; }
    float* phH() const { return d_buf + (int64_t)n_tmpl * n_time; }
    float* lqL() const { return d_buf + 2 * (int64_t)n_tmpl * n_time; }
    float* phL() const { return d_buf + 3 * (int64_t)n_tmpl * n_time; }
    ~Traj();
};

// Arrival times walked per template (sorted) and their log weights.
struct Times : Object {
    int n_tmpl = 0, n_walk = 0;
    float lw_rand = 0.0f;   // log weight of every drawn time (the loud times carry 0)
    int32_t* d_sel = nullptr;
    float* d_lw = nullptr;
    ~Times();
};

// A list of (template, table, ln a) reads grouped by table: jobs sorted by table, tiles of at most
// SCHED_TILE_JOBS jobs of one table (the scheduled form of the band kernel), the same tiles split into
// parts of at most TILE_JOBS jobs (the wave and piece forms), and for each sorted job its index in the
// caller's order.
constexpr int TILE_JOBS = 1024;
constexpr int SCHED_TILE_JOBS = 2048;
struct JobSet {
    int64_t njobs = 0;
    int ntiles = 0, ntiles_small = 0;
    int32_t* d_m = nullptr;       // (njobs) template row, table-sorted
    int32_t* d_tab = nullptr;     // (njobs) table index
    double* d_offb = nullptr;     // (njobs) (ln a - beta0) / dbeta
    int32_t* d_tiles = nullptr;   // (ntiles, 2) first job, job count
    int32_t* d_tiles_small = nullptr;   // (ntiles_small, 2) the tiles in parts of at most TILE_JOBS jobs
    std::vector<int64_t> order;   // sorted position -> caller index
    void* d_block = nullptr;
    size_t bytes = 0;
};
int jobset_build(int64_t njobs, const int32_t* m, const int32_t* tab, const double* offb, JobSet* js);
void jobset_free(JobSet* js);

// The float64 assembly of results over reads (k_assemble.hip): per template, the positions of its reads
// among the table-sorted values in the caller's read order (CSR), ln w of every read by sorted position.
struct LseLayout {
    std::vector<int32_t> off;     // (n_tmpl + 1)
    std::vector<int32_t> pos;     // (njobs) sorted positions, template by template
    std::vector<double> lnw;      // (njobs) ln w by sorted position
};
void lse_layout(int64_t njobs, int n_tmpl, const int32_t* m, const std::vector<int64_t>& order, const double* lnw,
                LseLayout* out);
size_t lse_device_bytes(int64_t njobs, int64_t n_tmpl, int nseg);
int check_node_weights(int n_tmpl, int64_t njobs, const double* lnw, const double* cellw);
// the layout on the device, the cell weights, the segments' template ranges and the outputs, carved from
// one block of lse_device_bytes
struct LseDevice {
    int32_t* off = nullptr;
    int32_t* pos = nullptr;
    double* lnw = nullptr;
    double* cell = nullptr;
    int32_t* seg = nullptr;
    double* L = nullptr;
    double* ln_b = nullptr;
    int64_t nm = 0;
    int nseg = 0;
    void carve(void* base, int64_t njobs, int64_t n_tmpl, int nseg);
    int upload(const LseLayout& lay, const double* cellw, const int32_t* segs, hipStream_t s) const;
    // ln Lambda per template from the table-sorted read values val, then ln B per segment
    int launch(const double* val, hipStream_t s) const;
};

// A reusable job set for exm_plan_gather, with the reads in the caller's order kept on the host.
struct Plan : Object {
    int n_tmpl = 0;
    JobSet js;
    std::vector<int32_t> h_m, h_tab;
    std::vector<double> h_offb;
    // node weights (exm_plan_set_weights): ln w per read in the caller's order, cell weight per template
    bool has_weights = false;
    std::vector<double> h_lnw, h_cellw;
    void* d_w = nullptr;
    size_t w_bytes = 0;
    LseDevice lse;
    ~Plan();
};

// The per-call constants of the source-frame step on the device (exm_source_grid_create): node
// coordinates yN (nrow, nm), masses, yq, m1, z and the prior weights, then per segment its node values,
// the surface, the grid and ln B, all in one allocation.
struct SourceGrid : Object {
    int nrow = 0, nm = 0, nm1 = 0, nz = 0;
    bool has_w = false;
    void* d = nullptr;
    size_t bytes = 0;
    double *lamN = nullptr, *yN = nullptr, *M = nullptr, *yq = nullptr, *m1 = nullptr, *z = nullptr,
           *w = nullptr, *surf = nullptr, *out = nullptr, *ln_b = nullptr, *part = nullptr;
    hipEvent_t done[2] = {nullptr, nullptr};     // the copies of the two staging slots have landed
    ~SourceGrid();
};

// Segments analysed together: walk records of every admitted segment in device buffers, the reads of
// every segment gathered by one launch.
struct Batch : Object {
    exm_handle store = 0;
    int nw = 0, ntau = 0;
    int64_t cap = 0, used = 0;          // templates (over all admitted segments)
    uint2* d_stream = nullptr;           // (cap, ntau, nw) walk records
    float* d_aH = nullptr;               // (cap, ntau, nw) ln|z_H| of the records (rows where beta may clamp)
    float* d_wa = nullptr;               // (cap, nw)
    float* d_lw = nullptr;               // (cap) drawn-time log weight per template
    double* d_lnnt = nullptr;            // (cap) ln n_time per template
    size_t bytes = 0;
    std::vector<int32_t> h_m, h_tab;
    std::vector<double> h_offb;
    std::vector<int64_t> seg_reads;
    // asynchronous gather (exm_batch_launch): own stream, completion event, jobs and work memory
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool in_flight = false;
    JobSet js;
    void* d_work = nullptr;              // first/last rows, row bounds, (max, sum) pairs, results
    size_t work_bytes = 0;
    size_t res_off = 0;                  // where the launch put the results in d_work
    // node weights of every admitted segment (from its plan), for exm_batch_collect_lse: ln w per read in
    // admission order, cell weight per template, first template of every segment
    bool weights = true;
    std::vector<double> h_lnw, h_cellw;
    std::vector<int32_t> seg_tmpl;
    LseLayout lay;                       // the launch's uploads read these until the batch is collected
    std::vector<int32_t> h_segs;
    LseDevice lse;
    bool lse_launched = false;
    ~Batch();
};

// gather kernel choice: 0 automatic, 1 reference-order kernel, 2 band kernel
int gather_mode();

// ---------------------------------------------------------------- context
struct Context {
    int device = -1;
    hipStream_t stream = nullptr;         // the library's stream
    hipStream_t launch_stream = nullptr;  // when set, gather-side launches go here (a batch's own stream)
    std::recursive_mutex mu;
    // grow-only scratch buffers
    static constexpr int NSCRATCH = 10;
    void* scratch[NSCRATCH] = {};
    size_t scratch_size[NSCRATCH] = {};
    std::unordered_map<uint64_t, hipfftHandle> fft_plans;  // key: (n_time << 20) | batch
    int piece = 0;                                          // walked times per piece (0: default)
    int f32_interp = 0;                                     // band kernel interpolates in float32
    int tile = 0;                                           // reads per band workgroup (0: SCHED_TILE_JOBS)
    int64_t filter_bytes = 0;                               // matched-filter workspace per transform (0: 64 MiB)
    int filter_f32 = 0;                                     // matched-filter transform in float32
    int filter_transform = 0;                               // 0: band transform where it fits, 1: hipFFT
    std::unordered_map<int64_t, void*> bt_twiddles;         // device tables e^{-2 pi i t / n}, by n
    hipStream_t filter_side = nullptr;                      // second stream of the band transform's batches
    hipEvent_t filter_ev[2] = {nullptr, nullptr};           // fork and join of the two streams
    // geometry of the zero-padded transform input (scratch slot 9) whose padding is already zero
    void* filter_in = nullptr;
    int64_t filter_in_nt = 0, filter_in_rows = 0;
    int filter_in_lo = -1, filter_in_nf = -1, filter_in_f32 = -1;
    int num_cu = 0;
    int64_t total_mem = 0;
    char name[256] = {0};
    // while set (inside the no-allocation entry point), scratch() carves its slots from this
    // caller-independent device region instead of allocating: each slot once per call, 256-byte aligned
    struct Arena {
        char* base = nullptr;
        size_t cap = 0, used = 0;
        void* slot_ptr[NSCRATCH] = {};
        size_t slot_size[NSCRATCH] = {};
    };
    Arena* arena = nullptr;
};

Context& ctx();
int ensure_init();

// stream of the gather-side launches: a batch's own stream inside exm_batch_launch, else the library's
inline hipStream_t cur_stream() {
    Context& c = ctx();
    return c.launch_stream ? c.launch_stream : c.stream;
}

// routes gather-side launches to stream s for the lifetime of the scope
struct StreamScope {
    explicit StreamScope(hipStream_t s) { ctx().launch_stream = s; }
    ~StreamScope() { ctx().launch_stream = nullptr; }
};

// kernel timing: events around named stages on the library stream (enabled by exm_profile_enable)
struct ProfileScope {
    hipEvent_t start = nullptr;
    const char* name;
    explicit ProfileScope(const char* n);
    ~ProfileScope();
};
int scratch(int slot, size_t bytes, void** out);

// handle registry
exm_handle register_object(Object* o);
Object* lookup(exm_handle h, uint32_t kind);
template <class T>
T* get(exm_handle h, uint32_t kind) {
    return static_cast<T*>(lookup(h, kind));
}

// ---------------------------------------------------------------- kernels (host launchers)
int launch_filter(Bank* bank, const double* dH, const double* dL, const double* s2H, const double* s2L, Traj* traj);
// arrival-time selection; the draw comes from the host (pick) or the device (d_pick, pick NULL)
int launch_select(Traj* traj, int lag_lo, int lag_hi, int K, int S, const int64_t* pick, const int64_t* d_pick,
                  float lw_rand, Times* times);
// numpy's default_rng(...).integers(0, high, n_out) from its PCG64 state (state hi, lo, inc hi, lo)
int launch_draw(const uint64_t rng_state[4], int64_t n_out, int64_t high, int64_t* d_out);
// ln Lambda of every job of js into d_out (table-sorted order)
int launch_gather(Traj* traj, Store* store, const JobSet& js, Times* times, double* d_out);
// scratch bytes launch_gather takes for these arguments (every slot rounded up to 256 bytes)
size_t gather_scratch_bytes(const Traj* traj, const Store* store, const JobSet& js, const Times* times);
size_t band_scratch_bytes(const Traj* traj, const Store* store, const JobSet& js, const Times* times);
// the no-allocation scratch region of the current device (exm_noalloc.hip), found at exm_init
int noalloc_init(int device);
// the times object of a handle (NULL for handle 0: every arrival time), checked against the trajectories
int check_times(Traj* traj, exm_handle htimes, Times** tm);
int launch_gather_ref(Traj* traj, Store* store, const JobSet& js, Times* times, double* om, double* os);
int launch_gather_band(Traj* traj, Store* store, const JobSet& js, Times* times, double* om, double* os);
int launch_walk(Traj* traj, Times* times, Store* st, float* wa_dst, uint2* stream_dst, float* aH_dst);
int launch_walk(Traj* traj, Times* times, int64_t off, int nw, Store* st, float* wa_dst, uint2* stream_dst,
                float* aH_dst);
int band_slice(const Traj* traj, const Store* st, const Times* times);
// bounds_buf: piece_bounds_bytes() of device memory for the row bounds and the band kernel's schedule
// (scratch slot 6 when null); its size depends on the band kernel form and piece length set at the call
size_t piece_bounds_bytes(const Store* st, const JobSet& js, int nw);
// accumulate: merge this walk's (max, sum) into om, os instead of writing them
int launch_piece(Store* st, const uint2* stream, const float* aHs, const float* wa, int nw, const float* lw_g,
                 const JobSet& js, double* om, double* os, void* bounds_buf = nullptr, bool accumulate = false);
// LSE over the lags of every read; ln n_time per template (ln_nt_g) or one value for all (ln_nt)
int launch_lse(const JobSet& js, int ntau, const double* om, const double* os, const double* ln_nt_g, double ln_nt,
               double* d_out);
int launch_fill(float* dst, int64_t n, float v);
int launch_copy_calibration(float4* dst, const float4* src, int64_t n);
// tables from store files (file, byte offset of the float32 member, device index) through pinned buffers
// pooled device memory of the library (blocks are kept for reuse until memory runs short)
int dev_alloc(void** p, size_t bytes);
void dev_free(void* p, size_t bytes);
int tables_sum_orders(const float* col, const float* row, int nord, int ntau, int nq, int nphif, const double* lnq,
                      const double* beta, int nb, const int* ladder, int nlad, float* const* out);
int store_load_files(Store* st, int n, const char* const* paths, const int64_t* offsets, const int* tables,
                     int readers, double* seconds);
int launch_source_frame(const double* lamN, const double* yN, int nrow, int nm, const double* M, const double* yq,
                        const double* m1, int nm1, const double* z, int nz, const double* w, double* surf,
                        double* out, double* ln_b, double* part);
int source_frame_limits(int nrow, int nm);
// workgroups of the first stage of ln B on the source-frame grid; part holds 2 doubles per workgroup
constexpr int LOG_MEAN_BLOCKS = 1024;
int launch_taylorf2(const double* d_M, const double* d_q, const double* d_chi, int n_tmpl, int n_freq, int lo,
                    double T, double d0_mpc, const double* d_sig2H, float2* d_W, double* d_u);
int launch_fill(double* dst, int64_t n, double v);
int launch_fill_synthetic(Store* st, int k, uint64_t seed, int64_t node, int kind);
int launch_store_transpose_in(Store* st, int k, const float* d_src);
int launch_store_transpose_out(Store* st, int k, float* d_dst);

// ---------------------------------------------------------------- device helpers
// e^x by Cody-Waite reduction and a degree-13 Taylor polynomial, every operation separately
// rounded: the same operation sequence as the CPU oracle, so arrival-time ranks agree bitwise.
__device__ __forceinline__ double pexp(double x) {
    const double k = __builtin_rint(__dmul_rn(x, 1.4426950408889634));
    const double r = __dsub_rn(__dsub_rn(x, __dmul_rn(k, 6.93147180369123816490e-01)),
                               __dmul_rn(k, 1.90821492927058770002e-10));
    const double c[14] = {1.0, 1.0, 0.5, 0.16666666666666666, 0.041666666666666664, 0.008333333333333333,
                          0.001388888888888889, 0.0001984126984126984, 2.48015873015873e-05,
                          2.7557319223985893e-06, 2.755731922398589e-07, 2.505210838544172e-08,
                          2.08767569878681e-09, 1.6059043836821613e-10};
    double p = c[13];
#pragma unroll
    for (int i = 12; i >= 0; --i) p = __dadd_rn(__dmul_rn(p, r), c[i]);
    return ldexp(p, (int)k);
}

}  // namespace exm
