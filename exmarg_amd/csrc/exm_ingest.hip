// Table ingest straight from the store files: pinned staging, reader threads, asynchronous copies and
// the device layout transform, overlapped.
//
// The reference reads every bd_*.npz member into (optionally pinned) host memory with 6 reader threads
// and copies the stack to the device (src/exmarg/search/store.py:55-87, pe.py:146-154).  Here the
// caller passes, for each table, the file and the byte offset of its float32 C-order member (the npz
// members are stored uncompressed).  Each table is cut into chunks; reader threads pread the chunks in
// order into a ring of pinned buffers; the calling thread copies each chunk to a device staging slab
// on a copy stream as soon as it is read, and when a table's slab is complete the layout transform to
// [lag][beta][Delta][phase] runs on the library's stream while the next table streams into the other
// slab.  Host memory never holds more than the ring.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#include "exm_common.h"

namespace exm {

namespace {

constexpr size_t CHUNK = (size_t)32 << 20;     // bytes per pinned buffer
constexpr int RING = 8;                          // pinned buffers

struct Chunk {
    int file;           // index into the caller's files
    int table;          // device table index
    int slot;           // staging slab
    size_t offset;      // byte offset of the chunk inside the table
    size_t bytes;
};

}  // namespace

int store_load_files(Store* st, int n, const char* const* paths, const int64_t* offsets, const int* tables,
                     int readers, double* seconds) {
    const size_t table_bytes = sizeof(float) * (size_t)st->tstride;
    const size_t nchunk = (table_bytes + CHUNK - 1) / CHUNK;
    readers = std::max(1, std::min(readers, 32));
    void* slabs_v = nullptr;
    EXM_TRY(scratch(3, 2 * table_bytes, &slabs_v));
    char* slab[2] = {(char*)slabs_v, (char*)slabs_v + table_bytes};

    std::vector<Chunk> chunks;
    for (int k = 0; k < n; ++k)
        for (size_t c = 0; c < nchunk; ++c)
            chunks.push_back(Chunk{k, tables[k], k & 1, c * CHUNK, std::min(CHUNK, table_bytes - c * CHUNK)});
    const size_t nj = chunks.size();

    int rc = EXM_OK;
    std::vector<void*> ring(RING, nullptr);
    std::vector<hipEvent_t> copied(RING, nullptr);
    hipEvent_t landed[2] = {nullptr, nullptr}, transposed[2] = {nullptr, nullptr};
    hipStream_t copy = nullptr;
    std::vector<int> fds(n, -1);
    for (int i = 0; i < RING && rc == EXM_OK; ++i)
        if (hipHostMalloc(&ring[i], CHUNK, hipHostMallocDefault) != hipSuccess) {
            ring[i] = nullptr;
            set_error("hipHostMalloc of %zu bytes failed", CHUNK);
            rc = EXM_ERR_NOMEM;
        }
    if (rc == EXM_OK && hipStreamCreate(&copy) != hipSuccess) rc = EXM_ERR_HIP;
    for (int i = 0; i < RING && rc == EXM_OK; ++i)
        if (hipEventCreateWithFlags(&copied[i], hipEventDisableTiming) != hipSuccess) rc = EXM_ERR_HIP;
    for (int i = 0; i < 2 && rc == EXM_OK; ++i)
        if (hipEventCreateWithFlags(&landed[i], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&transposed[i], hipEventDisableTiming) != hipSuccess)
            rc = EXM_ERR_HIP;
    for (int k = 0; k < n && rc == EXM_OK; ++k) {
        fds[k] = open(paths[k], O_RDONLY);
        if (fds[k] < 0) {
            set_error("cannot open %s", paths[k]);
            rc = EXM_ERR_ARG;
        }
    }

    // readers: chunk j goes to buffer j % RING once the copy of chunk j - RING has run
    std::mutex mu;
    std::condition_variable cv;
    std::vector<int> state(nj, 0);                 // 0 pending, 1 read, -1 failed
    size_t released = 0;                           // chunks below this have left their buffers
    bool stop = false;
    std::atomic<size_t> next{0};
    auto reader = [&]() {
        for (;;) {
            const size_t j = next.fetch_add(1);
            if (j >= nj) return;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || j < released + RING; });
                if (stop) return;
            }
            const Chunk& c = chunks[j];
            char* dst = (char*)ring[j % RING];
            size_t got = 0;
            while (got < c.bytes) {
                const ssize_t r = pread(fds[c.file], dst + got, c.bytes - got, (off_t)(offsets[c.file] + c.offset + got));
                if (r <= 0) break;
                got += (size_t)r;
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                state[j] = got == c.bytes ? 1 : -1;
            }
            cv.notify_all();
        }
    };
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    if (rc == EXM_OK)
        for (int i = 0; i < readers; ++i) pool.emplace_back(reader);

    for (size_t j = 0; j < nj && rc == EXM_OK; ++j) {
        const Chunk& c = chunks[j];
        const int b = (int)(j % RING);
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return state[j] != 0; });
        }
        if (state[j] < 0) {
            set_error("short read of %s", paths[c.file]);
            rc = EXM_ERR_ARG;
            break;
        }
        // a slab is rewritten only after the transform of the table it held
        if (c.offset == 0 && c.file >= 2 && hipStreamWaitEvent(copy, transposed[c.slot], 0) != hipSuccess) {
            rc = EXM_ERR_HIP;
            break;
        }
        if (hipMemcpyAsync(slab[c.slot] + c.offset, ring[b], c.bytes, hipMemcpyHostToDevice, copy) != hipSuccess ||
            hipEventRecord(copied[b], copy) != hipSuccess) {
            rc = EXM_ERR_HIP;
            break;
        }
        if (c.offset + c.bytes == table_bytes) {
            // the table is on the device: its transform runs on the library stream
            if (hipEventRecord(landed[c.slot], copy) != hipSuccess ||
                hipStreamWaitEvent(ctx().stream, landed[c.slot], 0) != hipSuccess) {
                rc = EXM_ERR_HIP;
                break;
            }
            rc = launch_store_transpose_in(st, c.table, (const float*)slab[c.slot]);
            if (rc == EXM_OK && hipEventRecord(transposed[c.slot], ctx().stream) != hipSuccess) rc = EXM_ERR_HIP;
        }
        // half a ring behind, wait for a copy to run and hand its buffers back to the readers
        if (j + 1 >= RING / 2) {
            const size_t done = j + 1 - RING / 2;
            if (hipEventSynchronize(copied[done % RING]) != hipSuccess) {
                rc = EXM_ERR_HIP;
                break;
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                released = done + 1;
            }
            cv.notify_all();
        }
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
    }
    cv.notify_all();
    for (auto& th : pool) th.join();
    if (copy) (void)hipStreamSynchronize(copy);
    (void)hipStreamSynchronize(ctx().stream);
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc == EXM_ERR_HIP) set_error("table ingest: %s", hipGetErrorString(hipGetLastError()));
    for (int fd : fds)
        if (fd >= 0) close(fd);
    for (auto e : copied)
        if (e) (void)hipEventDestroy(e);
    for (int i = 0; i < 2; ++i) {
        if (landed[i]) (void)hipEventDestroy(landed[i]);
        if (transposed[i]) (void)hipEventDestroy(transposed[i]);
    }
    if (copy) (void)hipStreamDestroy(copy);
    for (auto p : ring)
        if (p) (void)hipHostFree(p);
    return rc;
}

}  // namespace exm
