// Large transfers between pageable host memory and the device for the
// host-buffer entry points (what a MATLAB MEX gateway hands over: mxArray data
// is ordinary pageable memory).  A plain hipMemcpy from pageable memory goes
// through the runtime's own staging: 30 GB/s host -> device and 17 GB/s device
// -> host on 206 MB here (47 and 48 GB/s this way).  copy_h2d /
// copy_d2h split a buffer over kIoThreads host threads; each copies its share
// in kIoChunk pieces through two pinned slots of its own and its own stream,
// so the host memcpy of one piece overlaps the DMA of the previous one and the
// threads' DMAs run side by side.  Buffers below kIoMin bytes take the plain
// path.  One transfer runs at a time per process (the pool is shared).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "sid_internal.h"

namespace sid {

namespace {

constexpr int kIoThreads = 8;
constexpr size_t kIoChunk = size_t(4) << 20;
constexpr size_t kIoMin = size_t(16) << 20;

struct IoLane {
    void *slot[2] = {nullptr, nullptr};   // pinned host staging
    hipEvent_t done[2] = {nullptr, nullptr};
    hipStream_t stream = nullptr;
};

struct IoPool {
    int device = -1;
    IoLane lane[kIoThreads];
    bool ok = false;
};

std::mutex g_io_mutex;
std::vector<IoPool *> g_pools;   // one per device, kept until the process ends

// The pool of the current device, created on first use (nullptr on failure:
// the caller then takes the plain path).
IoPool *io_pool()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    for (IoPool *q : g_pools)
        if (q->device == dev) return q->ok ? q : nullptr;
    IoPool *q = new IoPool;
    q->device = dev;
    bool ok = true;
    for (auto &l : q->lane) {
        for (int s = 0; s < 2 && ok; ++s) {
            ok = hipHostMalloc(&l.slot[s], kIoChunk, hipHostMallocPortable) == hipSuccess &&
                 hipEventCreateWithFlags(&l.done[s], hipEventDisableTiming) == hipSuccess;
        }
        ok = ok && hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking) == hipSuccess;
    }
    (void)hipGetLastError();
    q->ok = ok;
    g_pools.push_back(q);
    return ok ? q : nullptr;
}

// Byte range of lane t (64-byte aligned boundaries).
void lane_range(size_t bytes, int t, size_t *b0, size_t *b1)
{
    const size_t per = (bytes / kIoThreads + 63) / 64 * 64;
    *b0 = std::min(bytes, per * (size_t)t);
    *b1 = t + 1 == kIoThreads ? bytes : std::min(bytes, per * (size_t)(t + 1));
}

}  // namespace

int copy_h2d(void *dst_dev, const void *src_host, size_t bytes, hipStream_t st)
{
    if (bytes < kIoMin) {
        SID_TRY_HIP(hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice, st));
        return SID_OK;
    }
    std::lock_guard<std::mutex> lock(g_io_mutex);
    IoPool *pool = io_pool();
    if (!pool) {
        SID_TRY_HIP(hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice, st));
        return SID_OK;
    }
    // the lanes' copies may not overtake work already queued on st that reads dst
    hipEvent_t start = nullptr;
    SID_TRY_HIP(hipEventCreateWithFlags(&start, hipEventDisableTiming));
    SID_TRY_HIP(hipEventRecord(start, st));
    int dev = pool->device;
    std::vector<int> err(kIoThreads, 0);
    std::vector<std::thread> th;
    auto lane_h2d = [&](int t) {
        if (hipSetDevice(dev) != hipSuccess) { err[t] = 1; return; }
        IoLane &l = pool->lane[t];
        if (hipStreamWaitEvent(l.stream, start, 0) != hipSuccess) { err[t] = 1; return; }
        size_t b0, b1;
        lane_range(bytes, t, &b0, &b1);
        int k = 0;
        for (size_t off = b0; off < b1; off += kIoChunk, ++k) {
            const size_t len = std::min(kIoChunk, b1 - off);
            const int s = k & 1;
            if (hipEventSynchronize(l.done[s]) != hipSuccess) { err[t] = 1; return; }   // slot free
            std::memcpy(l.slot[s], static_cast<const char *>(src_host) + off, len);
            if (hipMemcpyAsync(static_cast<char *>(dst_dev) + off, l.slot[s], len, hipMemcpyHostToDevice,
                               l.stream) != hipSuccess ||
                hipEventRecord(l.done[s], l.stream) != hipSuccess) {
                err[t] = 1;
                return;
            }
        }
    };
    // a lane whose thread cannot be started runs on this thread afterwards
    std::vector<int> inline_lanes;
    for (int t = 0; t < kIoThreads; ++t) {
        try {
            th.emplace_back(lane_h2d, t);
        } catch (...) {
            inline_lanes.push_back(t);
        }
    }
    for (auto &x : th) x.join();
    for (int t : inline_lanes) lane_h2d(t);
    (void)hipEventDestroy(start);
    for (int t = 0; t < kIoThreads; ++t)
        if (err[t]) {
            set_error("host-to-device copy failed");
            return SID_ERR_HIP;
        }
    // work queued on st after this call sees the data: st waits for every lane
    for (auto &l : pool->lane) {
        hipEvent_t e = nullptr;
        SID_TRY_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        SID_TRY_HIP(hipEventRecord(e, l.stream));
        SID_TRY_HIP(hipStreamWaitEvent(st, e, 0));
        SID_TRY_HIP(hipEventDestroy(e));
    }
    return SID_OK;
}

int copy_d2h(void *dst_host, const void *src_dev, size_t bytes, hipStream_t st)
{
    if (bytes < kIoMin) {
        SID_TRY_HIP(hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost, st));
        SID_TRY_HIP(hipStreamSynchronize(st));
        return SID_OK;
    }
    std::lock_guard<std::mutex> lock(g_io_mutex);
    IoPool *pool = io_pool();
    if (!pool) {
        SID_TRY_HIP(hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost, st));
        SID_TRY_HIP(hipStreamSynchronize(st));
        return SID_OK;
    }
    // the source is complete once the work queued on st so far has finished
    hipEvent_t ready = nullptr;
    SID_TRY_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    SID_TRY_HIP(hipEventRecord(ready, st));
    int dev = pool->device;
    std::vector<int> err(kIoThreads, 0);
    std::vector<std::thread> th;
    auto lane_d2h = [&](int t) {
        if (hipSetDevice(dev) != hipSuccess) { err[t] = 1; return; }
        IoLane &l = pool->lane[t];
        if (hipStreamWaitEvent(l.stream, ready, 0) != hipSuccess) { err[t] = 1; return; }
        size_t b0, b1;
        lane_range(bytes, t, &b0, &b1);
        // piece k goes device -> slot (k & 1); piece k + 1's DMA is queued before piece k is copied out
        auto issue = [&](size_t off, int s) {
            const size_t len = std::min(kIoChunk, b1 - off);
            return hipMemcpyAsync(l.slot[s], static_cast<const char *>(src_dev) + off, len,
                                  hipMemcpyDeviceToHost, l.stream) == hipSuccess &&
                   hipEventRecord(l.done[s], l.stream) == hipSuccess;
        };
        if (b0 < b1 && !issue(b0, 0)) { err[t] = 1; return; }
        int k = 0;
        for (size_t off = b0; off < b1; off += kIoChunk, ++k) {
            const int s = k & 1;
            const size_t len = std::min(kIoChunk, b1 - off);
            if (off + kIoChunk < b1 && !issue(off + kIoChunk, s ^ 1)) { err[t] = 1; return; }
            if (hipEventSynchronize(l.done[s]) != hipSuccess) { err[t] = 1; return; }
            std::memcpy(static_cast<char *>(dst_host) + off, l.slot[s], len);
        }
    };
    std::vector<int> inline_lanes;
    for (int t = 0; t < kIoThreads; ++t) {
        try {
            th.emplace_back(lane_d2h, t);
        } catch (...) {
            inline_lanes.push_back(t);
        }
    }
    for (auto &x : th) x.join();
    for (int t : inline_lanes) lane_d2h(t);
    (void)hipEventDestroy(ready);
    for (int t = 0; t < kIoThreads; ++t)
        if (err[t]) {
            set_error("device-to-host copy failed");
            return SID_ERR_HIP;
        }
    return SID_OK;
}

}  // namespace sid
