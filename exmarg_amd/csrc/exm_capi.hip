// C ABI of the exmarg HIP library: context, handles, argument checks and
// the entry points declared in include/exmarg_hip.h.
#include <algorithm>
#include <cstdarg>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>
#include <vector>

#include "exm_common.h"

namespace exm {

// ---------------------------------------------------------------- errors
static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char* last_error() { return g_err; }

// ---------------------------------------------------------------- context
static Context g_ctx;
static int64_t g_mem_in_use = 0;

Context& ctx() { return g_ctx; }

int ensure_init() {
    EXM_CHECK(g_ctx.device >= 0, EXM_ERR_STATE, "exm_init has not been called");
    EXM_HIP(hipSetDevice(g_ctx.device));
    return EXM_OK;
}

// Device memory: blocks are rounded up (2 MiB above 16 MiB) and freed blocks are kept in a pool by
// size, up to a quarter of the device, so that the per-call objects of repeated analyses (trajectories,
// times, read lists) reuse memory instead of mapping and unmapping it every call.  Frees are stream-
// ordered: a freed block carries an event recorded on the stream that issued its last work, and is handed
// out again only once that event has completed, so that a free never waits for unrelated work (a batch's
// gather running on its own stream).  A failed hipMalloc releases the pool and tries once more.
struct PooledBlock {
    void* p;
    hipEvent_t ev;
};
static std::multimap<size_t, PooledBlock> g_pool;        // free blocks by size
static size_t g_pool_bytes = 0;
static std::unordered_map<void*, size_t> g_blocks;        // the size of every block handed out
static std::vector<hipEvent_t> g_spare_events;

static size_t block_size(size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (bytes > ((size_t)16 << 20)) return (bytes + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
    return (bytes + 255) & ~(size_t)255;
}

static void pool_release() {
    (void)hipDeviceSynchronize();
    for (auto& kv : g_pool) {
        (void)hipFree(kv.second.p);
        g_spare_events.push_back(kv.second.ev);
    }
    g_pool.clear();
    g_pool_bytes = 0;
}

// allocation-failure injection (exm_debug_fail_alloc): the allocation after this many more fails once
static int64_t g_fail_after = -1;

int dev_alloc(void** p, size_t bytes) {
    if (g_fail_after >= 0 && g_fail_after-- == 0) {
        set_error("device allocation of %zu bytes failed: injected failure", bytes);
        *p = nullptr;
        return EXM_ERR_NOMEM;
    }
    bytes = block_size(bytes);
    // the smallest pooled block that holds the request: of the same size, or for large requests up to a
    // quarter larger (work buffers whose size varies from call to call then reuse one block instead of
    // allocating anew, which costs ~0.25 s per GB)
    auto it = g_pool.lower_bound(bytes);
    const size_t slack = bytes > ((size_t)16 << 20) ? bytes / 4 : 0;
    if (it != g_pool.end() && it->first <= bytes + slack) {
        const size_t got = it->first;
        const PooledBlock blk = it->second;
        g_pool.erase(it);
        g_pool_bytes -= got;
        // the work that used the block before it was freed has finished
        const hipError_t e = hipEventSynchronize(blk.ev);
        g_spare_events.push_back(blk.ev);
        if (e != hipSuccess) {
            (void)hipFree(blk.p);
            set_error("waiting for a pooled block: %s", hipGetErrorString(e));
            *p = nullptr;
            return EXM_ERR_HIP;
        }
        *p = blk.p;
        g_mem_in_use += (int64_t)got;
        g_blocks[*p] = got;
        return EXM_OK;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess && !g_pool.empty()) {
        (void)hipGetLastError();
        pool_release();
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        *p = nullptr;
        return EXM_ERR_NOMEM;
    }
    g_mem_in_use += (int64_t)bytes;
    g_blocks[*p] = bytes;
    return EXM_OK;
}

void dev_free(void* p, size_t bytes) {
    if (!p) return;
    auto b = g_blocks.find(p);
    if (b != g_blocks.end()) {              // the block's own size (a pooled block may exceed the request)
        bytes = b->second;
        g_blocks.erase(b);
    } else {
        bytes = block_size(bytes);
    }
    g_mem_in_use -= (int64_t)bytes;
    if (g_ctx.total_mem > 0 && g_pool_bytes + bytes <= (size_t)g_ctx.total_mem / 4) {
        // the block's last work ran on the gather stream in use (a batch's inside its launch) or the
        // library stream, which also waits for the filter's side stream at the end of every filter
        hipEvent_t ev = nullptr;
        if (!g_spare_events.empty()) {
            ev = g_spare_events.back();
            g_spare_events.pop_back();
        } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            ev = nullptr;
        }
        if (ev && hipEventRecord(ev, cur_stream()) == hipSuccess) {
            g_pool.emplace(bytes, PooledBlock{p, ev});
            g_pool_bytes += bytes;
            return;
        }
        (void)hipGetLastError();
        if (ev) g_spare_events.push_back(ev);
    }
    (void)hipDeviceSynchronize();
    (void)hipFree(p);
}

int scratch(int slot, size_t bytes, void** out) {
    Context& c = g_ctx;
    if (c.arena) {
        Context::Arena& a = *c.arena;
        if (a.slot_ptr[slot] && a.slot_size[slot] >= bytes) {
            *out = a.slot_ptr[slot];
            return EXM_OK;
        }
        const size_t need = (bytes + 255) & ~(size_t)255;
        EXM_CHECK(need <= a.cap - a.used, EXM_ERR_NOMEM,
                  "the no-allocation scratch of %zu bytes is full (%zu used, %zu more asked)", a.cap, a.used, need);
        a.slot_ptr[slot] = a.base + a.used;
        a.slot_size[slot] = need;
        a.used += need;
        *out = a.slot_ptr[slot];
        return EXM_OK;
    }
    if (c.scratch_size[slot] < bytes) {
        dev_free(c.scratch[slot], c.scratch_size[slot]);
        c.scratch[slot] = nullptr;
        c.scratch_size[slot] = 0;
        EXM_TRY(dev_alloc(&c.scratch[slot], bytes));
        c.scratch_size[slot] = bytes;
    }
    *out = c.scratch[slot];
    return EXM_OK;
}

Store::~Store() {
    dev_free(d_T, table_bytes());
    dev_free(d_tau, sizeof(int32_t) * ntau);
}
Bank::~Bank() { dev_free(d_W, sizeof(float2) * (size_t)n_tmpl * n_freq); }
Traj::~Traj() { dev_free(d_buf, sizeof(float) * 4 * (size_t)n_tmpl * n_time); }
Times::~Times() {
    dev_free(d_sel, sizeof(int32_t) * (size_t)n_tmpl * n_walk);
    dev_free(d_lw, sizeof(float) * (size_t)n_tmpl * n_walk);
}
Plan::~Plan() {
    jobset_free(&js);
    if (d_w) dev_free(d_w, w_bytes);
}
SourceGrid::~SourceGrid() {
    for (hipEvent_t e : done)
        if (e) (void)hipEventDestroy(e);
    if (d) dev_free(d, bytes);
}

// Pinned host staging of source-frame grids (the device-to-host copy at full rate, then a threaded copy to
// the caller's memory): one block, grown on demand and kept for the process; calls hold the library lock.
static void* g_pin = nullptr;
static size_t g_pin_bytes = 0;

static int pinned_staging(size_t bytes, void** p) {
    if (g_pin_bytes < bytes) {
        if (g_pin) (void)hipHostFree(g_pin);
        g_pin = nullptr;
        g_pin_bytes = 0;
        EXM_CHECK(hipHostMalloc(&g_pin, bytes, hipHostMallocDefault) == hipSuccess, EXM_ERR_HIP,
                  "hipHostMalloc of %zu bytes failed", bytes);
        g_pin_bytes = bytes;
    }
    *p = g_pin;
    return EXM_OK;
}

// memcpy of a large host block by up to 8 threads (one when a thread cannot be started)
static void copy_host_parallel(void* dst, const void* src, size_t bytes) {
    constexpr size_t PIECE = 2u << 20;
    const size_t n = std::min<size_t>(8, (bytes + PIECE - 1) / PIECE);
    const size_t per = n ? (bytes + n - 1) / n : 0;
    std::vector<std::thread> th;
    size_t done = per;                  // bytes [0, per) are copied by this thread
    try {
        for (size_t k = 1; k < n && k * per < bytes; ++k) {
            const size_t a = k * per, len = std::min(per, bytes - a);
            th.emplace_back([=] { std::memcpy((char*)dst + a, (const char*)src + a, len); });
            done = a + len;
        }
    } catch (...) {
    }
    std::memcpy(dst, src, std::min(per, bytes));
    for (auto& t : th) t.join();
    if (done < bytes) std::memcpy((char*)dst + done, (const char*)src + done, bytes - done);
}
Batch::~Batch() {
    if (done) {
        (void)hipEventSynchronize(done);
        (void)hipEventDestroy(done);
    }
    if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
    if (js.d_block) dev_free(js.d_block, js.bytes);
    if (d_work) dev_free(d_work, work_bytes);
    dev_free(d_stream, bytes);
}

int jobset_build(int64_t njobs, const int32_t* m, const int32_t* tab, const double* offb, JobSet* js) {
    js->njobs = njobs;
    js->order.resize(njobs);
    // jobs grouped by table, keeping their order within a table: a counting sort over the table ids
    // (a few hundred tables against millions of jobs), a stable comparison sort when the ids spread wide
    int32_t tlo = 0, thi = -1;
    if (njobs > 0) tlo = thi = tab[0];
    for (int64_t i = 1; i < njobs; ++i) {
        tlo = std::min(tlo, tab[i]);
        thi = std::max(thi, tab[i]);
    }
    if (njobs > 0 && (int64_t)thi - tlo < (int64_t)1 << 20) {
        std::vector<int64_t> start((size_t)(thi - tlo) + 2, 0);
        for (int64_t i = 0; i < njobs; ++i) ++start[(size_t)(tab[i] - tlo) + 1];
        for (size_t k = 1; k < start.size(); ++k) start[k] += start[k - 1];
        for (int64_t i = 0; i < njobs; ++i) js->order[start[(size_t)(tab[i] - tlo)]++] = i;
    } else {
        for (int64_t i = 0; i < njobs; ++i) js->order[i] = i;
        std::stable_sort(js->order.begin(), js->order.end(), [&](int64_t a, int64_t b) { return tab[a] < tab[b]; });
    }
    std::vector<int32_t> sm(njobs), st(njobs), tiles;
    std::vector<double> so(njobs);
    for (int64_t i = 0; i < njobs; ++i) {
        sm[i] = m[js->order[i]];
        st[i] = tab[js->order[i]];
        so[i] = offb[js->order[i]];
    }
    // tiles: runs of one table split evenly into pieces of at most SCHED_TILE_JOBS jobs.  Piece q of every
    // table holds about the same templates, so the tiles of one piece read the same walk records:
    // they are queued for one XCD (workgroups go to the 8 XCDs round-robin by workgroup index) and
    // the queues interleaved, so that the workgroups of an XCD share their records in its L2.
    constexpr int NXCD = 8;
    std::vector<int64_t> run_a, run_n, run_p;
    int64_t maxp = 0;
    for (int64_t a = 0; a < njobs;) {
        int64_t b = a;
        while (b < njobs && st[b] == st[a]) ++b;
        const int64_t n = b - a;
        const int64_t tile = g_ctx.tile ? g_ctx.tile : SCHED_TILE_JOBS;
        const int64_t pieces = (n + tile - 1) / tile;
        run_a.push_back(a);
        run_n.push_back(n);
        run_p.push_back(pieces);
        maxp = std::max(maxp, pieces);
        a = b;
    }
    std::vector<std::vector<int32_t>> queue(NXCD);
    for (int64_t q = 0; q < maxp; ++q)
        for (size_t u = 0; u < run_a.size(); ++u) {
            const int64_t pieces = run_p[u], n = run_n[u], a = run_a[u];
            if (q >= pieces) continue;
            int x;
            if (maxp >= NXCD) {
                x = (int)(q % NXCD);
            } else {                                 // XCDs x = q, q + maxp, ... share piece q
                const int nx = (int)((NXCD - 1 - q) / maxp) + 1;
                x = (int)q + (int)maxp * (int)(u % nx);
            }
            const int64_t lo = a + n * q / pieces, hi = a + n * (q + 1) / pieces;
            queue[x].push_back((int32_t)lo);
            queue[x].push_back((int32_t)(hi - lo));
        }
    std::vector<size_t> head(NXCD, 0);
    for (bool more = true; more;) {
        more = false;
        for (int x = 0; x < NXCD; ++x)
            if (head[x] < queue[x].size()) {
                tiles.push_back(queue[x][head[x]]);
                tiles.push_back(queue[x][head[x] + 1]);
                head[x] += 2;
                more = true;
            }
    }
    js->ntiles = (int)(tiles.size() / 2);
    // the same tiles in consecutive parts of at most TILE_JOBS jobs, for the wave and piece forms
    std::vector<int32_t> small;
    for (size_t t = 0; t < tiles.size(); t += 2) {
        const int64_t lo = tiles[t], n = tiles[t + 1], parts = (n + TILE_JOBS - 1) / TILE_JOBS;
        for (int64_t q = 0; q < parts; ++q) {
            small.push_back((int32_t)(lo + n * q / parts));
            small.push_back((int32_t)(n * (q + 1) / parts - n * q / parts));
        }
    }
    js->ntiles_small = (int)(small.size() / 2);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_m = 0, o_t = al(4 * (size_t)njobs), o_o = o_t + al(4 * (size_t)njobs),
                 o_tiles = o_o + al(8 * (size_t)njobs), o_small = o_tiles + al(4 * tiles.size()),
                 total = o_small + al(4 * small.size());
    EXM_TRY(dev_alloc(&js->d_block, total));
    js->bytes = total;
    char* base = (char*)js->d_block;
    js->d_m = (int32_t*)(base + o_m);
    js->d_tab = (int32_t*)(base + o_t);
    js->d_offb = (double*)(base + o_o);
    js->d_tiles = (int32_t*)(base + o_tiles);
    js->d_tiles_small = (int32_t*)(base + o_small);
    EXM_HIP(hipMemcpyAsync(js->d_m, sm.data(), 4 * (size_t)njobs, hipMemcpyHostToDevice, cur_stream()));
    EXM_HIP(hipMemcpyAsync(js->d_tab, st.data(), 4 * (size_t)njobs, hipMemcpyHostToDevice, cur_stream()));
    EXM_HIP(hipMemcpyAsync(js->d_offb, so.data(), 8 * (size_t)njobs, hipMemcpyHostToDevice, cur_stream()));
    EXM_HIP(hipMemcpyAsync(js->d_tiles, tiles.data(), 4 * tiles.size(), hipMemcpyHostToDevice, cur_stream()));
    EXM_HIP(hipMemcpyAsync(js->d_tiles_small, small.data(), 4 * small.size(), hipMemcpyHostToDevice, cur_stream()));
    EXM_HIP(hipStreamSynchronize(cur_stream()));
    return EXM_OK;
}

void jobset_free(JobSet* js) {
    if (js->d_block) {
        (void)hipStreamSynchronize(g_ctx.stream);
        dev_free(js->d_block, js->bytes);
    }
    js->d_block = nullptr;
}

static int g_gather_mode = 0;
int gather_mode() { return g_gather_mode; }

// ---------------------------------------------------------------- profiling
struct ProfileEntry {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    double done_ms = 0.0;
    int64_t count = 0;
};
static bool g_profile = false;
static std::unordered_map<std::string, ProfileEntry> g_prof;

ProfileScope::ProfileScope(const char* n) : name(n) {
    if (!g_profile) return;
    if (hipEventCreate(&start) != hipSuccess) {
        start = nullptr;
        return;
    }
    (void)hipEventRecord(start, cur_stream());
}

ProfileScope::~ProfileScope() {
    if (!start) return;
    hipEvent_t stop;
    if (hipEventCreate(&stop) != hipSuccess) {
        (void)hipEventDestroy(start);
        return;
    }
    (void)hipEventRecord(stop, cur_stream());
    g_prof[name].events.emplace_back(start, stop);
}

static void profile_collect() {
    for (auto& kv : g_prof) {
        for (auto& ev : kv.second.events) {
            float ms = 0.0f;
            (void)hipEventSynchronize(ev.second);
            if (hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess) kv.second.done_ms += ms;
            kv.second.count += 1;
            (void)hipEventDestroy(ev.first);
            (void)hipEventDestroy(ev.second);
        }
        kv.second.events.clear();
    }
}

// ---------------------------------------------------------------- registry
static std::unordered_map<exm_handle, Object*> g_objects;
static uint64_t g_next = 1;

exm_handle register_object(Object* o) {
    exm_handle h = (g_next++ << 8) | o->kind;
    g_objects[h] = o;
    return h;
}

Object* lookup(exm_handle h, uint32_t kind) {
    auto it = g_objects.find(h);
    if (it == g_objects.end() || it->second->kind != kind) return nullptr;
    return it->second;
}

static int get_plain(exm_handle h, Object** o) {
    auto it = g_objects.find(h);
    EXM_CHECK(it != g_objects.end(), EXM_ERR_HANDLE, "unknown handle %llu", (unsigned long long)h);
    *o = it->second;
    return EXM_OK;
}

int check_times(Traj* traj, exm_handle htimes, Times** tm) {
    *tm = nullptr;
    if (htimes == 0) return EXM_OK;
    Times* t = get<Times>(htimes, K_TIMES);
    EXM_CHECK(t != nullptr, EXM_ERR_HANDLE, "handle %llu is not a Times", (unsigned long long)htimes);
    EXM_CHECK(t->n_tmpl == traj->n_tmpl, EXM_ERR_ARG, "times are for %d templates, trajectories for %d",
              t->n_tmpl, traj->n_tmpl);
    *tm = t;
    return EXM_OK;
}

}  // namespace exm

using namespace exm;

#define LOCK std::lock_guard<std::recursive_mutex> lock_(exm::ctx().mu)
#define NEED(type, var, h, kind)                                                                 \
    type* var = exm::get<type>((h), (kind));                                                     \
    EXM_CHECK(var != nullptr, EXM_ERR_HANDLE, "handle %llu is not a " #type, (unsigned long long)(h))

extern "C" {

int exm_init(int device) {
    LOCK;
    Context& c = ctx();
    if (c.device >= 0) {
        EXM_CHECK(c.device == device, EXM_ERR_ARG, "library already initialised on device %d", c.device);
        return EXM_OK;
    }
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    EXM_CHECK(e == hipSuccess && n > 0, EXM_ERR_STATE, "no HIP device visible (%s)", hipGetErrorString(e));
    EXM_CHECK(device >= 0 && device < n, EXM_ERR_ARG, "device %d out of range [0, %d)", device, n);
    EXM_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    EXM_HIP(hipGetDeviceProperties(&prop, device));
    EXM_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    c.num_cu = prop.multiProcessorCount;
    c.total_mem = (int64_t)prop.totalGlobalMem;
    snprintf(c.name, sizeof(c.name), "%s", prop.gcnArchName);
    EXM_TRY(noalloc_init(device));
    c.device = device;
    return EXM_OK;
}

int exm_last_error(char* buf, size_t n) {
    if (buf && n) {
        snprintf(buf, n, "%s", last_error());
    }
    return EXM_OK;
}

int exm_device_info(char* name, size_t n, int64_t* total_mem, int* num_cu) {
    LOCK;
    EXM_TRY(ensure_init());
    if (name && n) snprintf(name, n, "%s", ctx().name);
    if (total_mem) *total_mem = ctx().total_mem;
    if (num_cu) *num_cu = ctx().num_cu;
    return EXM_OK;
}

int exm_free(exm_handle h) {
    LOCK;
    EXM_TRY(ensure_init());
    Object* o = nullptr;
    EXM_TRY(get_plain(h, &o));
    EXM_HIP(hipStreamSynchronize(ctx().stream));
    g_objects.erase(h);
    delete o;
    return EXM_OK;
}

int exm_mem_info(int64_t* free_bytes, int64_t* total_bytes) {
    LOCK;
    EXM_TRY(ensure_init());
    size_t f = 0, t = 0;
    EXM_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = (int64_t)(f + g_pool_bytes);   // pooled blocks are free for this library
    if (total_bytes) *total_bytes = (int64_t)t;
    return EXM_OK;
}

int exm_mem_in_use(int64_t* bytes) {
    LOCK;
    *bytes = g_mem_in_use;
    return EXM_OK;
}

// ---------------------------------------------------------------- tables
int exm_store_create(int ntab, int ntau, int nphif, int nb, int nd, const int64_t* tau_idx, double beta0,
                     double dbeta, double delta0, double ddelta, exm_handle* out) {
    LOCK;
    EXM_TRY(ensure_init());
    EXM_CHECK(ntab > 0 && ntau > 0 && nb >= 2 && nd >= 2, EXM_ERR_ARG,
              "bad table geometry ntab=%d ntau=%d nb=%d nd=%d", ntab, ntau, nb, nd);
    EXM_CHECK(nphif == NPHIF, EXM_ERR_ARG, "tables must hold %d phase bins, got %d", NPHIF, nphif);
    EXM_CHECK(dbeta > 0 && ddelta > 0, EXM_ERR_ARG, "grid steps must be positive");
    EXM_CHECK(tau_idx != nullptr, EXM_ERR_ARG, "tau_idx is NULL");
    Store* st = new Store();
    st->kind = K_STORE;
    st->ntab = ntab;
    st->ntau = ntau;
    st->nphif = nphif;
    st->nb = nb;
    st->nd = nd;
    st->tstride = (int64_t)ntau * nphif * nb * nd;
    st->tau.assign(tau_idx, tau_idx + ntau);
    st->beta0 = beta0;
    st->dbeta = dbeta;
    st->delta0 = delta0;
    st->ddelta = ddelta;
    std::vector<int32_t> t32(ntau);
    for (int i = 0; i < ntau; ++i) {
        if (tau_idx[i] > INT32_MAX / 2 || tau_idx[i] < -INT32_MAX / 2) {
            delete st;
            set_error("lag %lld out of range", (long long)tau_idx[i]);
            return EXM_ERR_ARG;
        }
        t32[i] = (int32_t)tau_idx[i];
    }
    int rc = dev_alloc((void**)&st->d_T, st->table_bytes());
    if (rc == EXM_OK) rc = dev_alloc((void**)&st->d_tau, sizeof(int32_t) * ntau);
    if (rc != EXM_OK) {
        delete st;
        return rc;
    }
    EXM_HIP(hipMemcpyAsync(st->d_tau, t32.data(), sizeof(int32_t) * ntau, hipMemcpyHostToDevice, ctx().stream));
    EXM_HIP(hipStreamSynchronize(ctx().stream));
    *out = register_object(st);
    return EXM_OK;
}

int exm_store_set_table(exm_handle h, int k, const float* table) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Store, st, h, K_STORE);
    EXM_CHECK(k >= 0 && k < st->ntab, EXM_ERR_ARG, "table index %d out of range [0, %d)", k, st->ntab);
    EXM_CHECK(table != nullptr, EXM_ERR_ARG, "table is NULL");
    // stage in slabs of one lag through scratch, then reorder on the device
    size_t lag_floats = (size_t)st->nphif * st->nb * st->nd;
    float* stage = nullptr;
    EXM_TRY(scratch(3, sizeof(float) * (size_t)st->tstride, (void**)&stage));
    EXM_HIP(hipMemcpyAsync(stage, table, sizeof(float) * lag_floats * st->ntau, hipMemcpyHostToDevice,
                           ctx().stream));
    EXM_TRY(launch_store_transpose_in(st, k, stage));
    EXM_HIP(hipStreamSynchronize(ctx().stream));
    return EXM_OK;
}

int exm_store_load_files(exm_handle h, int n, const char* const* paths, const int64_t* offsets, const int* tables,
                         int readers, double* seconds) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Store, st, h, K_STORE);
    EXM_CHECK(n >= 0 && (n == 0 || (paths && offsets && tables)), EXM_ERR_ARG, "NULL argument");
    for (int k = 0; k < n; ++k) {
        EXM_CHECK(tables[k] >= 0 && tables[k] < st->ntab, EXM_ERR_ARG, "table index %d out of range [0, %d)",
                  tables[k], st->ntab);
        EXM_CHECK(paths[k] != nullptr && offsets[k] >= 0, EXM_ERR_ARG, "bad file or offset for table %d", k);
    }
    if (n == 0) return EXM_OK;
    return store_load_files(st, n, paths, offsets, tables, readers, seconds);
}

int exm_store_fill_synthetic(exm_handle h, int k, uint64_t seed, int64_t node, int kind) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Store, st, h, K_STORE);
    EXM_CHECK(k >= 0 && k < st->ntab, EXM_ERR_ARG, "table index %d out of range [0, %d)", k, st->ntab);
    EXM_CHECK(kind == 0 || kind == 1, EXM_ERR_ARG, "kind must be 0 (smooth) or 1 (rough)");
    EXM_TRY(launch_fill_synthetic(st, k, seed, node, kind));
    EXM_HIP(hipStreamSynchronize(ctx().stream));
    return EXM_OK;
}

int exm_store_read_table(exm_handle h, int k, float* table) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Store, st, h, K_STORE);
    EXM_CHECK(k >= 0 && k < st->ntab, EXM_ERR_ARG, "table index %d out of range [0, %d)", k, st->ntab);
    float* stage = nullptr;
    EXM_TRY(scratch(3, sizeof(float) * (size_t)st->tstride, (void**)&stage));
    EXM_TRY(launch_store_transpose_out(st, k, stage));
    EXM_HIP(hipMemcpyAsync(table, stage, sizeof(float) * (size_t)st->tstride, hipMemcpyDeviceToHost,
                           ctx().stream));
    EXM_HIP(hipStreamSynchronize(ctx().stream));
    return EXM_OK;
}

// ---------------------------------------------------------------- bank, filter, trajectories
int exm_bank_create(const float* W, int n_tmpl, int n_freq, int lo, int64_t n_time, exm_handle* out) {
    LOCK;
    EXM_TRY(ensure_init());
    EXM_CHECK(n_tmpl > 0 && n_freq > 0 && lo >= 0, EXM_ERR_ARG, "bad bank shape n_tmpl=%d n_freq=%d lo=%d",
              n_tmpl, n_freq, lo);
    EXM_CHECK(n_time >= (int64_t)lo + n_freq, EXM_ERR_ARG, "n_time %lld shorter than the band end %d",
              (long long)n_time, lo + n_freq);
    EXM_CHECK(n_time <= (1ll << 30), EXM_ERR_ARG, "n_time %lld too large", (long long)n_time);
    Bank* b = new Bank();
    b->kind = K_BANK;
    b->n_tmpl = n_tmpl;
    b->n_freq = n_freq;
    b->lo = lo;
    b->n_time = n_time;
    int rc = dev_alloc((void**)&b->d_W, sizeof(float2) * (size_t)n_tmpl * n_freq);
    if (rc != EXM_OK) {
        delete b;
        return rc;
    }
    EXM_HIP(hipMemcpyAsync(b->d_W, W, sizeof(float2) * (size_t)n_tmpl * n_freq, hipMemcpyHostToDevice,
                           ctx().stream));
    EXM_HIP(hipStreamSynchronize(ctx().stream));
    *out = register_object(b);
    return EXM_OK;
}

int exm_bank_taylorf2(const double* M, const double* q, const double* chi, int n_tmpl, double T, int lo, int n_freq,
                      int64_t n_time, const double* sig2_H, double d0_mpc, double* u, exm_handle* out) {
    LOCK;
    EXM_TRY(ensure_init());
    EXM_CHECK(M && q && chi && sig2_H && u && out, EXM_ERR_ARG, "NULL argument");
    EXM_CHECK(n_tmpl > 0 && n_freq > 0 && lo >= 0 && T > 0 && d0_mpc > 0, EXM_ERR_ARG,
              "bad bank n_tmpl=%d n_freq=%d lo=%d T=%g D0=%g", n_tmpl, n_freq, lo, T, d0_mpc);
    EXM_CHECK(n_time >= (int64_t)lo + n_freq && n_time <= (1ll << 30), EXM_ERR_ARG, "bad n_time %lld",
              (long long)n_time);
    for (int i = 0; i < n_tmpl; ++i)
        EXM_CHECK(M[i] > 0 && q[i] > 0 && q[i] <= 1 && fabs(chi[i]) < 1, EXM_ERR_ARG,
                  "template %d: M=%g q=%g chi=%g outside M > 0, 0 < q <= 1, |chi| < 1", i, M[i], q[i], chi[i]);
    Bank* b = new Bank();
    b->kind = K_BANK;
    b->n_tmpl = n_tmpl;
    b->n_freq = n_freq;
    b->lo = lo;
    b->n_time = n_time;
    int rc = dev_alloc((void**)&b->d_W, sizeof(float2) * (size_t)n_tmpl * n_freq);
    if (rc != EXM_OK) {
        delete b;
        return rc;
    }
    // inputs and u through the small-upload scratch slot: (M, q, chi, u) per template, then sig2_H
    const size_t nb = sizeof(double) * (4 * (size_t)n_tmpl + n_freq);
    void* buf = nullptr;
    rc = scratch(2, nb, &buf);
    if (rc != EXM_OK) {
        delete b;
        return rc;
    }
    double* d = (double*)buf;
    hipStream_t s = g_ctx.stream;
    EXM_HIP(hipMemcpyAsync(d, M, sizeof(double) * n_tmpl, hipMemcpyHostToDevice, s));
    EXM_HIP(hipMemcpyAsync(d + n_tmpl, q, sizeof(double) * n_tmpl, hipMemcpyHostToDevice, s));
    EXM_HIP(hipMemcpyAsync(d + 2 * (size_t)n_tmpl, chi, sizeof(double) * n_tmpl, hipMemcpyHostToDevice, s));
    EXM_HIP(hipMemcpyAsync(d + 4 * (size_t)n_tmpl, sig2_H, sizeof(double) * n_freq, hipMemcpyHostToDevice, s));
    rc = launch_taylorf2(d, d + n_tmpl, d + 2 * (size_t)n_tmpl, n_tmpl, n_freq, lo, T, d0_mpc, d + 4 * (size_t)n_tmpl,
                         b->d_W, d + 3 * (size_t)n_tmpl);
    if (rc != EXM_OK) {
        delete b;
        return rc;
    }
    EXM_HIP(hipMemcpyAsync(u, d + 3 * (size_t)n_tmpl, sizeof(double) * n_tmpl, hipMemcpyDeviceToHost, s));
    EXM_HIP(hipStreamSynchronize(s));
    *out = register_object(b);
    return EXM_OK;
}

int exm_bank_read(exm_handle h, float* W) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Bank, b, h, K_BANK);
    EXM_CHECK(W != nullptr, EXM_ERR_ARG, "W is NULL");
    EXM_HIP(hipMemcpyAsync(W, b->d_W, sizeof(float2) * (size_t)b->n_tmpl * b->n_freq, hipMemcpyDeviceToHost,
                           g_ctx.stream));
    EXM_HIP(hipStreamSynchronize(g_ctx.stream));
    return EXM_OK;
}

static int new_traj(int n_tmpl, int64_t n_time, Traj** out) {
    Traj* t = new Traj();
    t->kind = K_TRAJ;
    t->n_tmpl = n_tmpl;
    t->n_time = n_time;
    int rc = dev_alloc((void**)&t->d_buf, sizeof(float) * 4 * (size_t)n_tmpl * n_time);
    if (rc != EXM_OK) {
        delete t;
        return rc;
    }
    *out = t;
    return EXM_OK;
}

int exm_filter(exm_handle hb, const double* dH, const double* dL, const double* s2H, const double* s2L,
               exm_handle* traj_out) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Bank, bank, hb, K_BANK);
    EXM_CHECK(dH && dL && s2H && s2L, EXM_ERR_ARG, "segment arrays must not be NULL");
    Traj* t = nullptr;
    EXM_TRY(new_traj(bank->n_tmpl, bank->n_time, &t));
    int rc = launch_filter(bank, dH, dL, s2H, s2L, t);
    if (rc == EXM_OK) {
        hipError_t e = hipStreamSynchronize(ctx().stream);
        if (e != hipSuccess) {
            set_error("filter: %s", hipGetErrorString(e));
            rc = EXM_ERR_HIP;
        }
    }
    if (rc != EXM_OK) {
        delete t;
        return rc;
    }
    *traj_out = register_object(t);
    return EXM_OK;
}

int exm_traj_create(int n_tmpl, int64_t n_time, const float* lqH, const float* phH, const float* lqL,
                    const float* phL, exm_handle* out) {
    LOCK;
    EXM_TRY(ensure_init());
    EXM_CHECK(n_tmpl > 0 && n_time > 0, EXM_ERR_ARG, "bad trajectory shape");
    Traj* t = nullptr;
    EXM_TRY(new_traj(n_tmpl, n_time, &t));
    size_t nb = sizeof(float) * (size_t)n_tmpl * n_time;
    const float* src[4] = {lqH, phH, lqL, phL};
    float* dst[4] = {t->lqH(), t->phH(), t->lqL(), t->phL()};
    for (int i = 0; i < 4; ++i) {
        hipError_t e = hipMemcpyAsync(dst[i], src[i], nb, hipMemcpyHostToDevice, ctx().stream);
        if (e != hipSuccess) {
            delete t;
            set_error("trajectory upload: %s", hipGetErrorString(e));
            return EXM_ERR_HIP;
        }
    }
    EXM_HIP(hipStreamSynchronize(ctx().stream));
    *out = register_object(t);
    return EXM_OK;
}

int exm_traj_read(exm_handle h, float* lqH, float* phH, float* lqL, float* phL) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Traj, t, h, K_TRAJ);
    size_t nb = sizeof(float) * (size_t)t->n_tmpl * t->n_time;
    float* dst[4] = {lqH, phH, lqL, phL};
    const float* src[4] = {t->lqH(), t->phH(), t->lqL(), t->phL()};
    for (int i = 0; i < 4; ++i)
        if (dst[i]) EXM_HIP(hipMemcpyAsync(dst[i], src[i], nb, hipMemcpyDeviceToHost, ctx().stream));
    EXM_HIP(hipStreamSynchronize(ctx().stream));
    return EXM_OK;
}

// ---------------------------------------------------------------- times
// the draw of the random times on the device into scratch slot 2
static int device_draw(const uint64_t* rng_state, int64_t n, int64_t high, int64_t** out) {
    EXM_CHECK(rng_state != nullptr, EXM_ERR_ARG, "rng_state is NULL");
    void* buf = nullptr;
    EXM_TRY(scratch(2, sizeof(int64_t) * (size_t)std::max<int64_t>(n, 1), &buf));
    *out = (int64_t*)buf;
    return launch_draw(rng_state, n, high, *out);
}

static int select_times(exm_handle ht, int lag_lo, int lag_hi, int K, int S, const int64_t* pick,
                        const uint64_t* rng_state, float lw_rand, exm_handle* times_out) {
    NEED(Traj, traj, ht, K_TRAJ);
    EXM_CHECK(lag_lo <= lag_hi, EXM_ERR_ARG, "lag range [%d, %d] is empty", lag_lo, lag_hi);
    EXM_CHECK((int64_t)(lag_hi - lag_lo) < traj->n_time, EXM_ERR_ARG, "lag range wider than the segment");
    EXM_CHECK(K > 0 && K < traj->n_time, EXM_ERR_ARG, "n_loud %d must lie in (0, n_time=%lld)", K,
              (long long)traj->n_time);
    EXM_CHECK(S > 0, EXM_ERR_ARG, "n_random must be positive, got %d", S);
    EXM_CHECK(K + S <= 16384, EXM_ERR_ARG, "n_loud + n_random = %d exceeds 16384", K + S);
    const int64_t nrest = traj->n_time - K;
    int64_t* d_pick = nullptr;
    if (pick) {
        for (int64_t i = 0; i < (int64_t)traj->n_tmpl * S; ++i)
            EXM_CHECK(pick[i] >= 0 && pick[i] < nrest, EXM_ERR_ARG, "pick[%lld] = %lld outside [0, %lld)",
                      (long long)i, (long long)pick[i], (long long)nrest);
    } else {
        EXM_TRY(device_draw(rng_state, (int64_t)traj->n_tmpl * S, nrest, &d_pick));
    }
    Times* tm = new Times();
    tm->kind = K_TIMES;
    tm->n_tmpl = traj->n_tmpl;
    tm->n_walk = K + S;
    tm->lw_rand = lw_rand;
    int rc = dev_alloc((void**)&tm->d_sel, sizeof(int32_t) * (size_t)tm->n_tmpl * tm->n_walk);
    if (rc == EXM_OK) rc = dev_alloc((void**)&tm->d_lw, sizeof(float) * (size_t)tm->n_tmpl * tm->n_walk);
    if (rc == EXM_OK) rc = launch_select(traj, lag_lo, lag_hi, K, S, pick, d_pick, lw_rand, tm);
    if (rc == EXM_OK) {
        hipError_t e = hipStreamSynchronize(ctx().stream);
        if (e != hipSuccess) {
            set_error("select_times: %s", hipGetErrorString(e));
            rc = EXM_ERR_HIP;
        }
    }
    if (rc != EXM_OK) {
        delete tm;
        return rc;
    }
    *times_out = register_object(tm);
    return EXM_OK;
}

int exm_select_times(exm_handle ht, int lag_lo, int lag_hi, int K, int S, const int64_t* pick, float lw_rand,
                     exm_handle* times_out) {
    LOCK;
    EXM_TRY(ensure_init());
    EXM_CHECK(pick != nullptr, EXM_ERR_ARG, "pick is NULL");
    return select_times(ht, lag_lo, lag_hi, K, S, pick, nullptr, lw_rand, times_out);
}

int exm_select_times_seeded(exm_handle ht, int lag_lo, int lag_hi, int K, int S, const uint64_t* rng_state,
                            float lw_rand, exm_handle* times_out) {
    LOCK;
    EXM_TRY(ensure_init());
    return select_times(ht, lag_lo, lag_hi, K, S, nullptr, rng_state, lw_rand, times_out);
}

int exm_draw_integers(const uint64_t* rng_state, int64_t n, int64_t high, int64_t* out) {
    LOCK;
    EXM_TRY(ensure_init());
    EXM_CHECK(out != nullptr || n == 0, EXM_ERR_ARG, "out is NULL");
    int64_t* d = nullptr;
    EXM_TRY(device_draw(rng_state, n, high, &d));
    if (n > 0) EXM_HIP(hipMemcpyAsync(out, d, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToHost, g_ctx.stream));
    EXM_HIP(hipStreamSynchronize(g_ctx.stream));
    return EXM_OK;
}

int exm_times_read(exm_handle h, int32_t* sel, float* lw) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Times, tm, h, K_TIMES);
    size_t n = (size_t)tm->n_tmpl * tm->n_walk;
    if (sel) EXM_HIP(hipMemcpyAsync(sel, tm->d_sel, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx().stream));
    if (lw) EXM_HIP(hipMemcpyAsync(lw, tm->d_lw, sizeof(float) * n, hipMemcpyDeviceToHost, ctx().stream));
    EXM_HIP(hipStreamSynchronize(ctx().stream));
    return EXM_OK;
}

// ---------------------------------------------------------------- gather

int exm_gather(exm_handle htraj, exm_handle hstore, const int64_t* mass_ids, int64_t nsel, const double* lna,
               const int64_t* tab_ids, int ns, exm_handle htimes, double* out) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Traj, traj, htraj, K_TRAJ);
    NEED(Store, st, hstore, K_STORE);
    Times* tm = nullptr;
    EXM_TRY(check_times(traj, htimes, &tm));
    EXM_CHECK(nsel > 0 && ns > 0, EXM_ERR_ARG, "empty gather (nsel=%lld ns=%d)", (long long)nsel, ns);
    EXM_CHECK(mass_ids && lna && tab_ids && out, EXM_ERR_ARG, "NULL argument");
    int64_t njobs = nsel * ns;
    // job e = n * nsel + j  ->  out[n, j]
    std::vector<int32_t> m(njobs), tab(njobs);
    std::vector<double> offb(njobs);
    for (int64_t j = 0; j < nsel; ++j) {
        EXM_CHECK(mass_ids[j] >= 0 && mass_ids[j] < traj->n_tmpl, EXM_ERR_ARG,
                  "mass_ids[%lld] = %lld outside [0, %d)", (long long)j, (long long)mass_ids[j], traj->n_tmpl);
        for (int n = 0; n < ns; ++n) {
            int64_t t = tab_ids[j * ns + n];
            EXM_CHECK(t >= 0 && t < st->ntab, EXM_ERR_ARG, "tab_ids[%lld, %d] = %lld outside [0, %d)",
                      (long long)j, n, (long long)t, st->ntab);
            int64_t e = (int64_t)n * nsel + j;
            m[e] = (int32_t)mass_ids[j];
            tab[e] = (int32_t)t;
            offb[e] = (lna[j * ns + n] - st->beta0) / st->dbeta;
        }
    }
    JobSet js;
    EXM_TRY(jobset_build(njobs, m.data(), tab.data(), offb.data(), &js));
    double* dout = nullptr;
    int rc = scratch(1, sizeof(double) * (size_t)njobs, (void**)&dout);
    if (rc == EXM_OK) rc = launch_gather(traj, st, js, tm, dout);
    std::vector<double> sorted(njobs);
    if (rc == EXM_OK && hipMemcpyAsync(sorted.data(), dout, sizeof(double) * njobs, hipMemcpyDeviceToHost,
                                       ctx().stream) != hipSuccess) {
        set_error("gather: copy of the results failed");
        rc = EXM_ERR_HIP;
    }
    if (rc == EXM_OK && hipStreamSynchronize(ctx().stream) != hipSuccess) {
        set_error("gather: kernel failed");
        rc = EXM_ERR_HIP;
    }
    if (rc == EXM_OK)
        for (int64_t i = 0; i < njobs; ++i) out[js.order[i]] = sorted[i];
    jobset_free(&js);
    return rc;
}

int exm_plan_create(exm_handle hstore, int n_tmpl, int64_t njobs, const int32_t* job_m, const int32_t* job_tab,
                    const double* job_lna, exm_handle* out) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Store, st, hstore, K_STORE);
    EXM_CHECK(njobs > 0 && n_tmpl > 0, EXM_ERR_ARG, "empty plan");
    EXM_CHECK(job_m && job_tab && job_lna, EXM_ERR_ARG, "NULL argument");
    std::vector<double> offb(njobs);
    for (int64_t e = 0; e < njobs; ++e) {
        EXM_CHECK(job_m[e] >= 0 && job_m[e] < n_tmpl, EXM_ERR_ARG, "job %lld: template %d outside [0, %d)",
                  (long long)e, job_m[e], n_tmpl);
        EXM_CHECK(job_tab[e] >= 0 && job_tab[e] < st->ntab, EXM_ERR_ARG, "job %lld: table %d outside [0, %d)",
                  (long long)e, job_tab[e], st->ntab);
        offb[e] = (job_lna[e] - st->beta0) / st->dbeta;
    }
    Plan* p = new Plan();
    p->kind = K_PLAN;
    p->n_tmpl = n_tmpl;
    p->h_m.assign(job_m, job_m + njobs);
    p->h_tab.assign(job_tab, job_tab + njobs);
    p->h_offb = offb;
    int rc = jobset_build(njobs, job_m, job_tab, offb.data(), &p->js);
    if (rc != EXM_OK) {
        delete p;
        return rc;
    }
    *out = register_object(p);
    return EXM_OK;
}

int exm_plan_gather(exm_handle hplan, exm_handle htraj, exm_handle hstore, exm_handle htimes, double* out) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Plan, plan, hplan, K_PLAN);
    NEED(Traj, traj, htraj, K_TRAJ);
    NEED(Store, st, hstore, K_STORE);
    Times* tm = nullptr;
    EXM_TRY(check_times(traj, htimes, &tm));
    EXM_CHECK(plan->n_tmpl == traj->n_tmpl, EXM_ERR_ARG, "plan is for %d templates, trajectories for %d",
              plan->n_tmpl, traj->n_tmpl);
    const int64_t njobs = plan->js.njobs;
    double* dout = nullptr;
    EXM_TRY(scratch(1, sizeof(double) * (size_t)njobs, (void**)&dout));
    EXM_TRY(launch_gather(traj, st, plan->js, tm, dout));
    std::vector<double> sorted(njobs);
    EXM_HIP(hipMemcpyAsync(sorted.data(), dout, sizeof(double) * njobs, hipMemcpyDeviceToHost, ctx().stream));
    EXM_HIP(hipStreamSynchronize(ctx().stream));
    for (int64_t i = 0; i < njobs; ++i) out[plan->js.order[i]] = sorted[i];
    return EXM_OK;
}

// ---------------------------------------------------------------- batches
int exm_batch_create(exm_handle hstore, int64_t max_templates, int n_walk, exm_handle* out) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Store, st, hstore, K_STORE);
    EXM_CHECK(max_templates > 0 && n_walk > 0 && n_walk <= 8192, EXM_ERR_ARG,
              "batch needs max_templates > 0 and 0 < n_walk <= 8192 (got %lld, %d)", (long long)max_templates, n_walk);
    Batch* b = new Batch();
    b->kind = K_BATCH;
    b->store = hstore;
    b->nw = n_walk;
    b->ntau = st->ntau;
    b->cap = max_templates;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t s_bytes = al(sizeof(uint2) * (size_t)max_templates * st->ntau * n_walk + WALK_PAD_BYTES);
    const size_t wa_bytes = al(sizeof(float) * (size_t)max_templates * n_walk + WALK_PAD_BYTES);
    const size_t lw_bytes = al(sizeof(float) * (size_t)max_templates);
    const size_t nt_bytes = al(sizeof(double) * (size_t)max_templates);
    const size_t ah_bytes = al(sizeof(float) * (size_t)max_templates * st->ntau * n_walk);
    b->bytes = s_bytes + wa_bytes + lw_bytes + nt_bytes + ah_bytes;
    int rc = dev_alloc((void**)&b->d_stream, b->bytes);
    if (rc != EXM_OK) {
        delete b;
        return rc;
    }
    char* base = (char*)b->d_stream;
    b->d_wa = (float*)(base + s_bytes);
    b->d_lw = (float*)(base + s_bytes + wa_bytes);
    b->d_lnnt = (double*)(base + s_bytes + wa_bytes + lw_bytes);
    b->d_aH = (float*)(base + s_bytes + wa_bytes + lw_bytes + nt_bytes);
    *out = register_object(b);
    return EXM_OK;
}

// n segments of one bank (and plan) admitted together: dH, dL (n, n_freq) complex128, one noise variance
// s2H, s2L for all of them; the random draw from pick (n, n_tmpl, S) or from rng_state (n, 4), one
// generator per segment.  Each segment is filtered apart; the selection, the walk and the fills run once
// over the n x n_tmpl templates.
static int batch_add(exm_handle hbatch, exm_handle hbank, exm_handle hplan, int n, const double* dH,
                     const double* dL, const double* s2H, const double* s2L, int lag_lo, int lag_hi, int K, int S,
                     const int64_t* pick, const uint64_t* rng_state, float lw_rand, int* slot) {
    NEED(Batch, b, hbatch, K_BATCH);
    NEED(Bank, bank, hbank, K_BANK);
    NEED(Plan, plan, hplan, K_PLAN);
    NEED(Store, st, b->store, K_STORE);
    EXM_CHECK(dH && dL && s2H && s2L && (pick || rng_state), EXM_ERR_ARG, "NULL argument");
    EXM_CHECK(n >= 1, EXM_ERR_ARG, "no segment to admit (n = %d)", n);
    EXM_CHECK(!b->in_flight, EXM_ERR_STATE, "batch in flight; collect it before adding segments");
    EXM_CHECK(K + S == b->nw, EXM_ERR_ARG, "n_loud + n_random = %d, the batch walks %d times", K + S, b->nw);
    EXM_CHECK(plan->n_tmpl == bank->n_tmpl, EXM_ERR_ARG, "plan is for %d templates, bank has %d", plan->n_tmpl,
              bank->n_tmpl);
    const int nm = bank->n_tmpl;
    const int64_t ntot = (int64_t)n * nm;
    EXM_CHECK(ntot <= INT32_MAX && b->used + ntot <= b->cap, EXM_ERR_NOMEM,
              "batch full: %lld + %d x %d templates exceed %lld", (long long)b->used, n, nm, (long long)b->cap);
    EXM_CHECK(lag_lo <= lag_hi && (int64_t)(lag_hi - lag_lo) < bank->n_time, EXM_ERR_ARG, "bad lag range");
    EXM_CHECK(K > 0 && K < bank->n_time && S > 0, EXM_ERR_ARG, "n_loud %d / n_random %d invalid for n_time %lld", K,
              S, (long long)bank->n_time);
    const int64_t nrest = bank->n_time - K;
    if (pick)
        for (int64_t i = 0; i < ntot * S; ++i)
            EXM_CHECK(pick[i] >= 0 && pick[i] < nrest, EXM_ERR_ARG, "pick[%lld] outside [0, %lld)", (long long)i,
                      (long long)nrest);
    // transient trajectories and times of the n segments in scratch slot 7
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t tr_bytes = al(sizeof(float) * 4 * (size_t)ntot * bank->n_time);
    const size_t sel_bytes = al(sizeof(int32_t) * (size_t)ntot * b->nw);
    void* tbuf = nullptr;
    EXM_TRY(scratch(7, tr_bytes + 2 * sel_bytes, &tbuf));
    Traj tr;
    tr.n_tmpl = (int)ntot;
    tr.n_time = bank->n_time;
    tr.d_buf = (float*)tbuf;
    Times tm;
    tm.n_tmpl = (int)ntot;
    tm.n_walk = b->nw;
    tm.lw_rand = lw_rand;
    tm.d_sel = (int32_t*)((char*)tbuf + tr_bytes);
    tm.d_lw = (float*)((char*)tbuf + tr_bytes + sel_bytes);
    int rc = EXM_OK;
    for (int s = 0; s < n && rc == EXM_OK; ++s) {
        // segment s writes template rows s nm .. s nm + nm - 1 of each of the four arrays: a view with the
        // whole set's row count and its first row moved
        Traj view;
        view.n_tmpl = (int)ntot;
        view.n_time = bank->n_time;
        view.d_buf = tr.d_buf + (int64_t)s * nm * bank->n_time;
        rc = launch_filter(bank, dH + 2 * (size_t)s * bank->n_freq, dL + 2 * (size_t)s * bank->n_freq, s2H, s2L, &view);
        view.d_buf = nullptr;
    }
    int64_t* d_pick = nullptr;
    if (rc == EXM_OK && !pick) {
        void* pk = nullptr;
        rc = scratch(2, sizeof(int64_t) * (size_t)std::max<int64_t>(ntot * S, 1), &pk);
        d_pick = (int64_t*)pk;
        for (int s = 0; s < n && rc == EXM_OK; ++s)
            rc = launch_draw(rng_state + 4 * (size_t)s, (int64_t)nm * S, nrest, d_pick + (int64_t)s * nm * S);
    }
    if (rc == EXM_OK) rc = launch_select(&tr, lag_lo, lag_hi, K, S, pick, d_pick, lw_rand, &tm);
    if (rc == EXM_OK)
        rc = launch_walk(&tr, &tm, st, b->d_wa + b->used * b->nw, b->d_stream + b->used * b->ntau * b->nw,
                         b->d_aH + b->used * b->ntau * b->nw);
    if (rc == EXM_OK) rc = launch_fill(b->d_lw + b->used, ntot, lw_rand);
    if (rc == EXM_OK) rc = launch_fill(b->d_lnnt + b->used, ntot, std::log((double)bank->n_time));
    // the host-side trajectory and times objects only borrow scratch memory
    tr.d_buf = nullptr;
    tm.d_sel = nullptr;
    tm.d_lw = nullptr;
    if (rc != EXM_OK) return rc;
    EXM_HIP(hipStreamSynchronize(ctx().stream));
    const int first = (int)b->seg_reads.size();
    for (int s = 0; s < n; ++s) {
        for (size_t e = 0; e < plan->h_m.size(); ++e) {
            b->h_m.push_back((int32_t)(plan->h_m[e] + b->used));
            b->h_tab.push_back(plan->h_tab[e]);
            b->h_offb.push_back(plan->h_offb[e]);
        }
        // the plan's node weights go with its reads (a segment admitted without them leaves the batch
        // without a device assembly: exm_batch_collect still gives the read values)
        b->weights = b->weights && plan->has_weights;
        if (b->weights) {
            b->h_lnw.insert(b->h_lnw.end(), plan->h_lnw.begin(), plan->h_lnw.end());
            b->h_cellw.insert(b->h_cellw.end(), plan->h_cellw.begin(), plan->h_cellw.end());
        }
        b->seg_tmpl.push_back((int32_t)b->used);
        b->seg_reads.push_back((int64_t)plan->h_m.size());
        b->used += nm;
    }
    if (slot) *slot = first;
    return EXM_OK;
}


int exm_batch_add(exm_handle hbatch, exm_handle hbank, exm_handle hplan, const double* dH, const double* dL,
                  const double* s2H, const double* s2L, int lag_lo, int lag_hi, int K, int S, const int64_t* pick,
                  float lw_rand, int* slot) {
    LOCK;
    EXM_TRY(ensure_init());
    EXM_CHECK(pick != nullptr, EXM_ERR_ARG, "pick is NULL");
    return batch_add(hbatch, hbank, hplan, 1, dH, dL, s2H, s2L, lag_lo, lag_hi, K, S, pick, nullptr, lw_rand, slot);
}

int exm_batch_add_seeded(exm_handle hbatch, exm_handle hbank, exm_handle hplan, const double* dH, const double* dL,
                         const double* s2H, const double* s2L, int lag_lo, int lag_hi, int K, int S,
                         const uint64_t* rng_state, float lw_rand, int* slot) {
    LOCK;
    EXM_TRY(ensure_init());
    return batch_add(hbatch, hbank, hplan, 1, dH, dL, s2H, s2L, lag_lo, lag_hi, K, S, nullptr, rng_state, lw_rand,
                     slot);
}

int exm_batch_add_seeded_many(exm_handle hbatch, exm_handle hbank, exm_handle hplan, int n, const double* dH,
                              const double* dL, const double* s2H, const double* s2L, int lag_lo, int lag_hi, int K,
                              int S, const uint64_t* rng_states, float lw_rand, int* first_slot) {
    LOCK;
    EXM_TRY(ensure_init());
    return batch_add(hbatch, hbank, hplan, n, dH, dL, s2H, s2L, lag_lo, lag_hi, K, S, nullptr, rng_states, lw_rand,
                     first_slot);
}

static void batch_release_jobs(Batch* b) {
    if (b->js.d_block) dev_free(b->js.d_block, b->js.bytes);
    b->js = JobSet();
}

int exm_reserve(int n_tmpl, int64_t n_time, int n_freq, int n_walk) {
    LOCK;
    EXM_TRY(ensure_init());
    EXM_CHECK(n_tmpl > 0 && n_time > 0 && n_freq > 0 && n_walk > 0, EXM_ERR_ARG, "bad reserve shape");
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t nm = (size_t)n_tmpl, nt = (size_t)n_time, nw = (size_t)n_walk, nf = (size_t)n_freq;
    void* p = nullptr;
    // filter and selection work (the selection's keys: 8 n_time bytes per template, at most SEL_KEY_BYTES; the
    // band transform's two buffers of one workspace each; plus one row), the filter's zero-padded input
    // (its workspace, at least one row), uploads, walk temporaries, batch admission
    const size_t fws = g_ctx.filter_bytes > 0 ? (size_t)g_ctx.filter_bytes : ((size_t)64 << 20);
    const size_t keys = std::min<size_t>((size_t)SEL_KEY_BYTES, 8 * nm * nt);
    const size_t bands = 2 * (std::max(std::min(fws, 16 * nm * nt), 16 * nt) + 256);
    EXM_TRY(scratch(0, std::max(std::max(keys, bands), 16 * nt) + 16 * nt, &p));
    EXM_TRY(scratch(9, std::max(fws, 16 * nt), &p));
    EXM_TRY(scratch(2, 8 * std::max(std::max(6 * nf + 32, nm * nw), 4 * nm + nf), &p));
    EXM_TRY(scratch(5, 2 * al(4 * nm * nw), &p));
    EXM_TRY(scratch(7, al(16 * nm * nt) + 2 * al(4 * nm * nw), &p));
    return EXM_OK;
}

int exm_batch_launch(exm_handle hbatch) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Batch, b, hbatch, K_BATCH);
    NEED(Store, st, b->store, K_STORE);
    EXM_CHECK(!b->in_flight, EXM_ERR_STATE, "batch already launched; collect it first");
    if (!b->stream) EXM_HIP(hipStreamCreate(&b->stream));
    if (!b->done) EXM_HIP(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    const int64_t njobs = (int64_t)b->h_m.size();
    if (njobs > 0) {
        StreamScope scope(b->stream);
        EXM_TRY(jobset_build(njobs, b->h_m.data(), b->h_tab.data(), b->h_offb.data(), &b->js));
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        const size_t bnd = al(piece_bounds_bytes(st, b->js, b->nw));
        const size_t pairs = al(sizeof(double) * 2 * (size_t)njobs * st->ntau);
        const size_t lse_at = bnd + pairs + ((sizeof(double) * (size_t)njobs + 255) & ~(size_t)255);
        const int nseg = (int)b->seg_tmpl.size();
        const size_t need = lse_at + (b->weights ? lse_device_bytes(njobs, b->used, nseg) : 0);
        if (need > b->work_bytes) {
            if (b->d_work) dev_free(b->d_work, b->work_bytes);
            b->d_work = nullptr;
            b->work_bytes = 0;
            int rc = dev_alloc(&b->d_work, need);
            if (rc != EXM_OK) {
                batch_release_jobs(b);
                return rc;
            }
            b->work_bytes = need;
        }
        char* w = (char*)b->d_work;
        double* om = (double*)(w + bnd);
        double* os = om + (size_t)njobs * st->ntau;
        double* dout = (double*)(w + bnd + pairs);
        b->res_off = bnd + pairs;
        int rc = launch_piece(st, b->d_stream, b->d_aH, b->d_wa, b->nw, b->d_lw, b->js, om, os, w);
        if (rc == EXM_OK) rc = launch_lse(b->js, st->ntau, om, os, b->d_lnnt, 0.0, dout);
        // with every segment's node weights: ln Lambda per template and ln B per segment on the device
        b->lse_launched = false;
        if (rc == EXM_OK && b->weights) {
            lse_layout(njobs, (int)b->used, b->h_m.data(), b->js.order, b->h_lnw.data(), &b->lay);
            b->h_segs.assign(b->seg_tmpl.begin(), b->seg_tmpl.end());
            b->h_segs.push_back((int32_t)b->used);
            b->lse.carve(w + lse_at, njobs, b->used, nseg);
            rc = b->lse.upload(b->lay, b->h_cellw.data(), b->h_segs.data(), b->stream);
            if (rc == EXM_OK) rc = b->lse.launch(dout, b->stream);
            b->lse_launched = rc == EXM_OK;
        }
        if (rc != EXM_OK) {
            (void)hipStreamSynchronize(b->stream);
            batch_release_jobs(b);
            return rc;
        }
    }
    EXM_HIP(hipEventRecord(b->done, b->stream));
    b->in_flight = true;
    return EXM_OK;
}

int exm_batch_poll(exm_handle hbatch, int* done) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Batch, b, hbatch, K_BATCH);
    EXM_CHECK(done != nullptr, EXM_ERR_ARG, "done is NULL");
    EXM_CHECK(b->in_flight, EXM_ERR_STATE, "batch not launched");
    const hipError_t e = hipEventQuery(b->done);
    if (e == hipErrorNotReady) {
        *done = 0;
        return EXM_OK;
    }
    EXM_HIP(e);
    *done = 1;
    return EXM_OK;
}

int exm_batch_collect(exm_handle hbatch, double* out) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Batch, b, hbatch, K_BATCH);
    NEED(Store, st, b->store, K_STORE);
    EXM_CHECK(b->in_flight, EXM_ERR_STATE, "batch not launched");
    const int64_t njobs = (int64_t)b->h_m.size();
    EXM_CHECK(out != nullptr || njobs == 0, EXM_ERR_ARG, "out is NULL");
    b->in_flight = false;
    int rc = hipEventSynchronize(b->done) == hipSuccess ? EXM_OK : EXM_ERR_HIP;
    if (rc != EXM_OK) set_error("batch gather failed");
    if (rc == EXM_OK && njobs > 0) {
        auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
        const size_t off = b->res_off;
        std::vector<double> sorted(njobs);
        if (hipMemcpyAsync(sorted.data(), (char*)b->d_work + off, sizeof(double) * njobs, hipMemcpyDeviceToHost,
                           b->stream) != hipSuccess ||
            hipStreamSynchronize(b->stream) != hipSuccess) {
            set_error("batch result copy failed");
            rc = EXM_ERR_HIP;
        } else {
            for (int64_t i = 0; i < njobs; ++i) out[b->js.order[i]] = sorted[i];
        }
    }
    b->lse_launched = false;
    batch_release_jobs(b);
    return rc;
}

int exm_batch_collect_lse(exm_handle hbatch, double* ln_lambda, double* ln_b) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Batch, b, hbatch, K_BATCH);
    EXM_CHECK(b->in_flight, EXM_ERR_STATE, "batch not launched");
    const int64_t njobs = (int64_t)b->h_m.size();
    EXM_CHECK(b->lse_launched || njobs == 0, EXM_ERR_STATE,
              "a segment of the batch was admitted with a plan without node weights (exm_plan_set_weights); "
              "collect the read values with exm_batch_collect");
    EXM_CHECK((ln_lambda && ln_b) || b->used == 0, EXM_ERR_ARG, "NULL argument");
    b->in_flight = false;
    int rc = hipEventSynchronize(b->done) == hipSuccess ? EXM_OK : EXM_ERR_HIP;
    if (rc != EXM_OK) set_error("batch gather failed");
    if (rc == EXM_OK && njobs > 0) {
        if (hipMemcpyAsync(ln_lambda, b->lse.L, sizeof(double) * (size_t)b->used, hipMemcpyDeviceToHost, b->stream) !=
                hipSuccess ||
            hipMemcpyAsync(ln_b, b->lse.ln_b, sizeof(double) * (size_t)b->lse.nseg, hipMemcpyDeviceToHost, b->stream) !=
                hipSuccess ||
            hipStreamSynchronize(b->stream) != hipSuccess) {
            set_error("batch result copy failed");
            rc = EXM_ERR_HIP;
        }
    }
    b->lse_launched = false;
    batch_release_jobs(b);
    return rc;
}

int exm_debug_fail_alloc(int64_t after) {
    LOCK;
    g_fail_after = after < 0 ? -1 : after;
    return EXM_OK;
}

int exm_batch_run(exm_handle hbatch, double* out) {
    LOCK;
    EXM_TRY(exm_batch_launch(hbatch));
    return exm_batch_collect(hbatch, out);
}

int exm_batch_reset(exm_handle hbatch) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Batch, b, hbatch, K_BATCH);
    EXM_CHECK(!b->in_flight, EXM_ERR_STATE, "batch in flight; collect it first");
    b->used = 0;
    b->h_m.clear();
    b->h_tab.clear();
    b->h_offb.clear();
    b->seg_reads.clear();
    b->weights = true;
    b->h_lnw.clear();
    b->h_cellw.clear();
    b->seg_tmpl.clear();
    b->lse_launched = false;
    return EXM_OK;
}

int exm_batch_info(exm_handle hbatch, int64_t* templates_used, int64_t* reads, int* segments) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(Batch, b, hbatch, K_BATCH);
    if (templates_used) *templates_used = b->used;
    if (reads) *reads = (int64_t)b->h_m.size();
    if (segments) *segments = (int)b->seg_reads.size();
    return EXM_OK;
}

int exm_set_gather_kernel(int which) {
    LOCK;
    EXM_CHECK(which >= 0 && which <= 4, EXM_ERR_ARG,
              "gather kernel must be 0 (auto), 1 (reference), 2 (band, piece form), 3 (band, wave form) or 4 "
              "(band, scheduled form)");
    g_gather_mode = which;
    return EXM_OK;
}

int exm_set_gather_piece(int n) {
    LOCK;
    EXM_CHECK(n == 0 || n == 4 || n == 6 || n == 8 || n == 16, EXM_ERR_ARG,
              "piece length must be 0 (default), 4, 6, 8 or 16");
    g_ctx.piece = n;
    return EXM_OK;
}

int exm_set_gather_tile(int n) {
    LOCK;
    EXM_CHECK(n == 0 || n == 256 || n == 512 || n == TILE_JOBS || n == SCHED_TILE_JOBS, EXM_ERR_ARG,
              "reads per workgroup must be 0 (default), 256, 512, %d or %d", TILE_JOBS, SCHED_TILE_JOBS);
    g_ctx.tile = n;
    return EXM_OK;
}

int exm_set_filter_workspace(int64_t bytes) {
    LOCK;
    EXM_CHECK(bytes == 0 || (bytes >= (1ll << 20) && bytes <= (1ll << 32)), EXM_ERR_ARG,
              "filter workspace must be 0 (default, 64 MiB) or 1 MiB .. 4 GiB, got %lld", (long long)bytes);
    g_ctx.filter_bytes = bytes;
    return EXM_OK;
}

int exm_set_filter_precision(int p) {
    LOCK;
    EXM_CHECK(p == 0 || p == 1, EXM_ERR_ARG, "filter precision must be 0 (float64) or 1 (float32 transform)");
    g_ctx.filter_f32 = p;
    return EXM_OK;
}

int exm_set_filter_transform(int t) {
    LOCK;
    EXM_CHECK(t == 0 || t == 1, EXM_ERR_ARG, "filter transform must be 0 (band transform where it fits) or 1 (hipFFT)");
    g_ctx.filter_transform = t;
    return EXM_OK;
}

int exm_get_gather_tuning(int* kernel, int* piece, int* tile, int* precision) {
    LOCK;
    if (kernel) *kernel = g_gather_mode;
    if (piece) *piece = g_ctx.piece;
    if (tile) *tile = g_ctx.tile;
    if (precision) *precision = g_ctx.f32_interp;
    return EXM_OK;
}

int exm_set_gather_precision(int p) {
    LOCK;
    EXM_CHECK(p == 0 || p == 1, EXM_ERR_ARG, "gather precision must be 0 (float64) or 1 (float32 interpolation)");
    g_ctx.f32_interp = p;
    return EXM_OK;
}

int exm_profile_enable(int on) {
    LOCK;
    EXM_TRY(ensure_init());
    g_profile = on != 0;
    return EXM_OK;
}

int exm_profile_reset(void) {
    LOCK;
    EXM_TRY(ensure_init());
    profile_collect();
    g_prof.clear();
    return EXM_OK;
}

int exm_profile_get(const char* name, double* total_ms, int64_t* count) {
    LOCK;
    EXM_TRY(ensure_init());
    profile_collect();
    auto it = g_prof.find(name ? name : "");
    *total_ms = it == g_prof.end() ? 0.0 : it->second.done_ms;
    *count = it == g_prof.end() ? 0 : it->second.count;
    return EXM_OK;
}

int exm_source_frame(const double* lamN, const double* yN, int nrow, int nm, const double* Mdet, const double* yq,
                     const double* m1, int nm1, const double* z, int nz, const double* w, double* out, double* ln_b) {
    LOCK;
    EXM_TRY(ensure_init());
    EXM_CHECK(lamN && yN && Mdet && yq && m1 && z, EXM_ERR_ARG, "NULL argument");
    EXM_CHECK(nm1 > 0 && nz > 0, EXM_ERR_ARG, "empty (m1, z) grid");
    EXM_CHECK(w == nullptr || ln_b != nullptr, EXM_ERR_ARG, "ln_b is NULL");
    EXM_TRY(source_frame_limits(nrow, nm));
    for (int m = 1; m < nm; ++m)
        EXM_CHECK(Mdet[m] > Mdet[m - 1], EXM_ERR_ARG, "detector-frame masses must ascend (template %d)", m);
    // device copies: lamN, yN, M, yq, m1, z, w | surf | out | ln_b (slot 8) | ln B partials
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t n_in = 2 * (size_t)nrow * nm + nm + nz + nm1 + nz;
    const size_t n_grid = (size_t)nm1 * nz;
    const size_t o_w = al(8 * n_in), o_surf = o_w + (w ? al(8 * n_grid) : 0), o_out = o_surf + al(8 * (size_t)nz * nm),
                 o_lnb = o_out + al(8 * n_grid), o_part = o_lnb + 256, total = o_part + al(16 * LOG_MEAN_BLOCKS);
    void* buf = nullptr;
    EXM_TRY(scratch(8, total, &buf));
    char* b = (char*)buf;
    double* d = (double*)b;
    double *dL = d, *dY = dL + (size_t)nrow * nm, *dM = dY + (size_t)nrow * nm, *dyq = dM + nm, *dm1 = dyq + nz,
           *dz = dm1 + nm1;
    const hipStream_t s = g_ctx.stream;
    EXM_HIP(hipMemcpyAsync(dL, lamN, 8 * (size_t)nrow * nm, hipMemcpyHostToDevice, s));
    EXM_HIP(hipMemcpyAsync(dY, yN, 8 * (size_t)nrow * nm, hipMemcpyHostToDevice, s));
    EXM_HIP(hipMemcpyAsync(dM, Mdet, 8 * (size_t)nm, hipMemcpyHostToDevice, s));
    EXM_HIP(hipMemcpyAsync(dyq, yq, 8 * (size_t)nz, hipMemcpyHostToDevice, s));
    EXM_HIP(hipMemcpyAsync(dm1, m1, 8 * (size_t)nm1, hipMemcpyHostToDevice, s));
    EXM_HIP(hipMemcpyAsync(dz, z, 8 * (size_t)nz, hipMemcpyHostToDevice, s));
    double* dw = w ? (double*)(b + o_w) : nullptr;
    if (w) EXM_HIP(hipMemcpyAsync(dw, w, 8 * n_grid, hipMemcpyHostToDevice, s));
    double* dout = (double*)(b + o_out);
    double* dlnb = (double*)(b + o_lnb);
    {
        ProfileScope prof("source_frame");
        EXM_TRY(launch_source_frame(dL, dY, nrow, nm, dM, dyq, dm1, nm1, dz, nz, dw, (double*)(b + o_surf), dout, dlnb,
                                    (double*)(b + o_part)));
    }
    void* pin = nullptr;
    if (out) {
        EXM_TRY(pinned_staging(8 * n_grid, &pin));
        EXM_HIP(hipMemcpyAsync(pin, dout, 8 * n_grid, hipMemcpyDeviceToHost, s));
    }
    if (w) EXM_HIP(hipMemcpyAsync(ln_b, dlnb, 8, hipMemcpyDeviceToHost, s));
    EXM_HIP(hipStreamSynchronize(s));
    if (out) copy_host_parallel(out, pin, 8 * n_grid);
    return EXM_OK;
}

int exm_source_grid_create(const double* yN, int nrow, int nm, const double* Mdet, const double* yq, const double* m1,
                           int nm1, const double* z, int nz, const double* w, exm_handle* out) {
    LOCK;
    EXM_TRY(ensure_init());
    EXM_CHECK(yN && Mdet && yq && m1 && z && out, EXM_ERR_ARG, "NULL argument");
    EXM_CHECK(nm1 > 0 && nz > 0, EXM_ERR_ARG, "empty (m1, z) grid");
    EXM_TRY(source_frame_limits(nrow, nm));
    for (int m = 1; m < nm; ++m)
        EXM_CHECK(Mdet[m] > Mdet[m - 1], EXM_ERR_ARG, "detector-frame masses must ascend (template %d)", m);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t n_node = (size_t)nrow * nm, n_grid = (size_t)nm1 * nz;
    // lamN | yN | M | yq | m1 | z | w | surf | out | ln_b | partials
    const size_t o_y = al(8 * n_node), o_m = o_y + al(8 * n_node), o_yq = o_m + al(8 * (size_t)nm),
                 o_m1 = o_yq + al(8 * (size_t)nz), o_z = o_m1 + al(8 * (size_t)nm1), o_w = o_z + al(8 * (size_t)nz),
                 o_surf = o_w + (w ? al(8 * n_grid) : 0), o_out = o_surf + al(8 * (size_t)nz * nm),
                 o_lnb = o_out + al(8 * n_grid), o_part = o_lnb + 256, total = o_part + al(16 * LOG_MEAN_BLOCKS);
    SourceGrid* g = new SourceGrid();
    g->kind = K_GRID;
    g->nrow = nrow, g->nm = nm, g->nm1 = nm1, g->nz = nz, g->has_w = w != nullptr;
    int rc = dev_alloc(&g->d, total);
    if (rc != EXM_OK) {
        delete g;
        return rc;
    }
    g->bytes = total;
    char* b = (char*)g->d;
    g->lamN = (double*)b, g->yN = (double*)(b + o_y), g->M = (double*)(b + o_m), g->yq = (double*)(b + o_yq);
    g->m1 = (double*)(b + o_m1), g->z = (double*)(b + o_z), g->w = w ? (double*)(b + o_w) : nullptr;
    g->surf = (double*)(b + o_surf), g->out = (double*)(b + o_out), g->ln_b = (double*)(b + o_lnb);
    g->part = (double*)(b + o_part);
    for (hipEvent_t& e : g->done)
        if (rc == EXM_OK && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            e = nullptr;
            set_error("hipEventCreate failed");
            rc = EXM_ERR_HIP;
        }
    const hipStream_t s = g_ctx.stream;
    const struct { void* dst; const void* src; size_t n; } up[] = {
        {g->yN, yN, 8 * n_node}, {g->M, Mdet, 8 * (size_t)nm}, {g->yq, yq, 8 * (size_t)nz},
        {g->m1, m1, 8 * (size_t)nm1}, {g->z, z, 8 * (size_t)nz}, {g->w, w, w ? 8 * n_grid : 0}};
    for (const auto& u : up)
        if (u.n && rc == EXM_OK && hipMemcpyAsync(u.dst, u.src, u.n, hipMemcpyHostToDevice, s) != hipSuccess) {
            set_error("hipMemcpyAsync of the source-frame grid failed");
            rc = EXM_ERR_HIP;
        }
    if (rc == EXM_OK && hipStreamSynchronize(s) != hipSuccess) {
        set_error("hipStreamSynchronize failed");
        rc = EXM_ERR_HIP;
    }
    if (rc != EXM_OK) {
        delete g;
        return rc;
    }
    *out = register_object(g);
    return EXM_OK;
}

int exm_source_grid_run_many(exm_handle hgrid, int n, const double* lamN, double* out, double* ln_b) {
    LOCK;
    EXM_TRY(ensure_init());
    NEED(SourceGrid, g, hgrid, K_GRID);
    EXM_CHECK(n >= 0 && (n == 0 || lamN != nullptr), EXM_ERR_ARG, "NULL argument");
    EXM_CHECK(!g->has_w || n == 0 || ln_b != nullptr, EXM_ERR_ARG, "ln_b is NULL");
    if (n == 0) return EXM_OK;
    const hipStream_t s = g_ctx.stream;
    const size_t lam_n = (size_t)g->nrow * g->nm, grid_n = (size_t)g->nm1 * g->nz;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    // pinned staging, two slots: the grid, the node values, ln B
    const size_t o_lam = 2 * al(8 * grid_n), o_b = o_lam + 2 * al(8 * lam_n);
    char* pin = nullptr;
    EXM_TRY(pinned_staging(o_b + 2 * 256, (void**)&pin));
    auto slot_out = [&](int k) { return pin + (k & 1) * al(8 * grid_n); };
    auto slot_lam = [&](int k) { return pin + o_lam + (k & 1) * al(8 * lam_n); };
    auto slot_b = [&](int k) { return (double*)(pin + o_b + (k & 1) * 256); };
    // segment k is staged, computed and copied back while the host copies segment k - 1's grid out of the
    // other slot; a slot is reused only after its previous segment's event has been waited for
    int rc = EXM_OK;
    for (int k = 0; k <= n && rc == EXM_OK; ++k) {
        if (k < n) {
            std::memcpy(slot_lam(k), lamN + (size_t)k * lam_n, 8 * lam_n);
            if (hipMemcpyAsync(g->lamN, slot_lam(k), 8 * lam_n, hipMemcpyHostToDevice, s) != hipSuccess) {
                set_error("hipMemcpyAsync of the node values failed");
                rc = EXM_ERR_HIP;
            }
            if (rc == EXM_OK) {
                ProfileScope prof("source_frame");
                rc = launch_source_frame(g->lamN, g->yN, g->nrow, g->nm, g->M, g->yq, g->m1, g->nm1, g->z, g->nz,
                                         g->w, g->surf, g->out, g->ln_b, g->part);
            }
            if (rc == EXM_OK && out &&
                hipMemcpyAsync(slot_out(k), g->out, 8 * grid_n, hipMemcpyDeviceToHost, s) != hipSuccess) {
                set_error("hipMemcpyAsync of the grid failed");
                rc = EXM_ERR_HIP;
            }
            if (rc == EXM_OK && g->has_w &&
                hipMemcpyAsync(slot_b(k), g->ln_b, 8, hipMemcpyDeviceToHost, s) != hipSuccess) {
                set_error("hipMemcpyAsync of ln B failed");
                rc = EXM_ERR_HIP;
            }
            if (rc == EXM_OK && hipEventRecord(g->done[k & 1], s) != hipSuccess) {
                set_error("hipEventRecord failed");
                rc = EXM_ERR_HIP;
            }
        }
        if (rc == EXM_OK && k >= 1) {
            const int p = k - 1;
            if (hipEventSynchronize(g->done[p & 1]) != hipSuccess) {
                set_error("hipEventSynchronize failed");
                rc = EXM_ERR_HIP;
                break;
            }
            if (out) copy_host_parallel(out + (size_t)p * grid_n, slot_out(p), 8 * grid_n);
            if (g->has_w) ln_b[p] = *slot_b(p);
        }
    }
    if (rc != EXM_OK) (void)hipStreamSynchronize(s);   // no copy left in flight into the staging block
    return rc;
}

int exm_source_grid_run(exm_handle hgrid, const double* lamN, double* out, double* ln_b) {
    return exm_source_grid_run_many(hgrid, 1, lamN, out, ln_b);
}

int exm_copy_calibration(int64_t bytes, double* ms) {
    LOCK;
    EXM_TRY(ensure_init());
    EXM_CHECK(bytes > 0 && bytes % 16 == 0, EXM_ERR_ARG, "copy size must be a positive multiple of 16 bytes");
    void *a = nullptr, *b = nullptr;
    EXM_TRY(dev_alloc(&a, (size_t)bytes));
    int rc = dev_alloc(&b, (size_t)bytes);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == EXM_OK) rc = launch_fill((float*)a, bytes / 4, 1.0f);
    if (rc == EXM_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
        set_error("hipEventCreate failed");
        rc = EXM_ERR_HIP;
    }
    if (rc == EXM_OK) rc = hipEventRecord(e0, g_ctx.stream) == hipSuccess ? EXM_OK : EXM_ERR_HIP;
    if (rc == EXM_OK) rc = launch_copy_calibration((float4*)b, (const float4*)a, bytes / 16);
    if (rc == EXM_OK) rc = hipEventRecord(e1, g_ctx.stream) == hipSuccess ? EXM_OK : EXM_ERR_HIP;
    if (rc == EXM_OK) rc = hipEventSynchronize(e1) == hipSuccess ? EXM_OK : EXM_ERR_HIP;
    float t = 0.0f;
    if (rc == EXM_OK) rc = hipEventElapsedTime(&t, e0, e1) == hipSuccess ? EXM_OK : EXM_ERR_HIP;
    if (ms) *ms = t;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(g_ctx.stream);
    dev_free(a, (size_t)bytes);
    if (b) dev_free(b, (size_t)bytes);
    return rc;
}

int exm_tables_sum_orders(const float* col, const float* row, int nord, int ntau, int nq, int nphif,
                          const double* lnq, const double* beta, int nb, const int* ladder, int nlad, float* const* out) {
    LOCK;
    EXM_TRY(ensure_init());
    return tables_sum_orders(col, row, nord, ntau, nq, nphif, lnq, beta, nb, ladder, nlad, out);
}

}  // extern "C"
