// HIP runtime plumbing for libquest_mi355x: one device, one non-blocking stream
// (default priority), loud failure on any HIP error, event timing.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "qmx_device.h"

namespace qmx {

namespace {
std::once_flag g_once;
hipStream_t g_stream = nullptr;
double *g_pinned = nullptr;
hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
hipDeviceProp_t g_prop;
void (*g_pendingWork)() = nullptr;

// Staging ring for small per-launch inputs (gate and block matrices): pinned
// host slots and device slots of kStageSlotBytes, one event per slot. A slot's
// host side is refilled once the copy that last used it has completed; its
// device side is rewritten by a copy queued on the library stream, so after
// every kernel that read it. Allocated once, with the stream.
constexpr int kStageSlots = 256;
struct StageRing {
    char *dev = nullptr, *host = nullptr;
    hipEvent_t ev[kStageSlots];
    bool used[kStageSlots] = {};
    int next = 0;
} g_ring;

// inputs larger than a slot: one pinned / device pair, grown on demand
struct LargeStage {
    char *dev = nullptr, *host = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
    bool used = false;
} g_large;

// Deferred work (a composed density superoperator) is enqueued before anything
// else reaches the stream: launches, copies, synchronisation, timer events.
hipStream_t flushedStream() {
    if (g_pendingWork) {
        void (*fn)() = g_pendingWork;
        g_pendingWork = nullptr;
        fn();
    }
    return g_stream;
}
}  // namespace

void setPendingWork(void (*fn)()) { g_pendingWork = fn; }

[[noreturn]] void hipFatal(hipError_t err, const char *what, const char *file, int line) {
    std::fprintf(stderr, "libquest_mi355x: HIP error '%s' (%d) from %s at %s:%d\n",
                 hipGetErrorString(err), (int)err, what, file, line);
    std::fflush(stderr);
    std::abort();
}

void runtimeInit() {
    std::call_once(g_once, [] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n < 1) {
            std::fprintf(stderr,
                         "libquest_mi355x: no HIP device available (%s); this library has no CPU "
                         "fallback\n",
                         hipGetErrorString(e));
            std::fflush(stderr);
            std::abort();
        }
        QMX_HIP_CHECK(hipSetDevice(0));
        QMX_HIP_CHECK(hipGetDeviceProperties(&g_prop, 0));
        QMX_HIP_CHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
        QMX_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&g_pinned), 64 * sizeof(double),
                                    hipHostMallocDefault));
        QMX_HIP_CHECK(hipEventCreate(&g_ev0));
        QMX_HIP_CHECK(hipEventCreate(&g_ev1));
        QMX_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&g_ring.dev), kStageSlots * kStageSlotBytes));
        QMX_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&g_ring.host), kStageSlots * kStageSlotBytes,
                                    hipHostMallocDefault));
        for (int i = 0; i < kStageSlots; ++i)
            QMX_HIP_CHECK(hipEventCreateWithFlags(&g_ring.ev[i], hipEventDisableTiming));
        QMX_HIP_CHECK(hipEventCreateWithFlags(&g_large.ev, hipEventDisableTiming));
    });
}

Staged stageBegin(size_t bytes) {
    runtimeInit();
    if (bytes > kStageSlotBytes) {
        std::fprintf(stderr, "libquest_mi355x: %zu B exceed a staging slot (%zu B)\n", bytes, kStageSlotBytes);
        std::abort();
    }
    const int slot = g_ring.next;
    g_ring.next = (slot + 1) % kStageSlots;
    if (g_ring.used[slot]) QMX_HIP_CHECK(hipEventSynchronize(g_ring.ev[slot]));
    return Staged{g_ring.host + slot * kStageSlotBytes, g_ring.dev + slot * kStageSlotBytes, slot};
}

void stageCommit(const Staged &s, size_t bytes, void *stream) {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    QMX_HIP_CHECK(hipMemcpyAsync(s.dev, s.host, bytes, hipMemcpyHostToDevice, st));
    QMX_HIP_CHECK(hipEventRecord(g_ring.ev[s.slot], st));
    g_ring.used[s.slot] = true;
}

const void *stageLarge(const void *src, size_t bytes, void *stream) {
    runtimeInit();
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (bytes > g_large.bytes) {
        QMX_HIP_CHECK(hipStreamSynchronize(st));  // the old buffers may still be in use
        if (g_large.dev) QMX_HIP_CHECK(hipFree(g_large.dev));
        if (g_large.host) QMX_HIP_CHECK(hipHostFree(g_large.host));
        QMX_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&g_large.dev), bytes));
        QMX_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&g_large.host), bytes, hipHostMallocDefault));
        g_large.bytes = bytes;
        g_large.used = false;
    }
    if (g_large.used) QMX_HIP_CHECK(hipEventSynchronize(g_large.ev));
    std::memcpy(g_large.host, src, bytes);
    QMX_HIP_CHECK(hipMemcpyAsync(g_large.dev, g_large.host, bytes, hipMemcpyHostToDevice, st));
    QMX_HIP_CHECK(hipEventRecord(g_large.ev, st));
    g_large.used = true;
    return g_large.dev;
}

void *stream() { return flushedStream(); }

void deviceSync() {
    runtimeInit();
    QMX_HIP_CHECK(hipStreamSynchronize(flushedStream()));
}

void *devAlloc(size_t bytes) {
    runtimeInit();
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        std::fprintf(stderr, "libquest_mi355x: hipMalloc of %zu bytes failed: %s\n", bytes,
                     hipGetErrorString(e));
        std::fflush(stderr);
        std::abort();
    }
    return p;
}

void devFree(void *p) {
    if (!p) return;
    QMX_HIP_CHECK(hipStreamSynchronize(flushedStream()));
    QMX_HIP_CHECK(hipFree(p));
}

void hostToDevice(void *dst, const void *src, size_t bytes) {
    QMX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, flushedStream()));
    QMX_HIP_CHECK(hipStreamSynchronize(g_stream));
}

void deviceToHost(void *dst, const void *src, size_t bytes) {
    QMX_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, flushedStream()));
    QMX_HIP_CHECK(hipStreamSynchronize(g_stream));
}

double *pinnedScratch() {
    runtimeInit();
    return g_pinned;
}

void checkLaunch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) hipFatal(e, what, __FILE__, __LINE__);
}

void timerBegin() {
    runtimeInit();
    QMX_HIP_CHECK(hipEventRecord(g_ev0, flushedStream()));
}

double timerEndMs() {
    QMX_HIP_CHECK(hipEventRecord(g_ev1, flushedStream()));
    QMX_HIP_CHECK(hipEventSynchronize(g_ev1));
    float ms = 0.f;
    QMX_HIP_CHECK(hipEventElapsedTime(&ms, g_ev0, g_ev1));
    return (double)ms;
}

bool allowDynamicLds(const void *kernel, size_t bytes, const char *caller) {
    runtimeInit();
    if (bytes + 1024 > g_prop.sharedMemPerBlock) {
        std::fprintf(stderr, "libquest_mi355x: %s: %zu B of LDS exceeds the device's %zu B\n", caller, bytes,
                     (size_t)g_prop.sharedMemPerBlock);
        return false;
    }
    QMX_HIP_CHECK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return true;
}

int numCUs() {
    runtimeInit();
    return g_prop.multiProcessorCount;
}

int deviceDescription(char *buf, int len) {
    runtimeInit();
    size_t freeB = 0, totalB = 0;
    QMX_HIP_CHECK(hipMemGetInfo(&freeB, &totalB));
    return std::snprintf(buf, (size_t)len,
                         "device=%s arch=%s CUs=%d clockMHz=%d memGiB=%.1f freeGiB=%.1f",
                         g_prop.name, g_prop.gcnArchName, g_prop.multiProcessorCount,
                         g_prop.clockRate / 1000, totalB / 1073741824.0, freeB / 1073741824.0);
}

}  // namespace qmx

// Event pairs for timing individual launches on the library stream (used by
// the benchmark to time one kernel class inside a longer timed region).
namespace {
constexpr int kMaxSlots = 8192;
hipEvent_t g_slotEv[kMaxSlots][2];
bool g_slotMade[kMaxSlots];

hipEvent_t slotEvent(int slot, int which) {
    if (slot < 0 || slot >= kMaxSlots) {
        std::fprintf(stderr, "libquest_mi355x: timer slot %d out of range [0, %d)\n", slot, kMaxSlots);
        std::abort();
    }
    if (!g_slotMade[slot]) {
        QMX_HIP_CHECK(hipEventCreate(&g_slotEv[slot][0]));
        QMX_HIP_CHECK(hipEventCreate(&g_slotEv[slot][1]));
        g_slotMade[slot] = true;
    }
    return g_slotEv[slot][which];
}
}  // namespace

extern "C" void qmxTimerSlotBegin(int slot) {
    qmx::runtimeInit();
    QMX_HIP_CHECK(hipEventRecord(slotEvent(slot, 0), qmx::hstream()));
}

extern "C" void qmxTimerSlotEnd(int slot) { QMX_HIP_CHECK(hipEventRecord(slotEvent(slot, 1), qmx::hstream())); }

extern "C" double qmxTimerSlotMs(int slot) {
    QMX_HIP_CHECK(hipEventSynchronize(slotEvent(slot, 1)));
    float ms = 0.f;
    QMX_HIP_CHECK(hipEventElapsedTime(&ms, slotEvent(slot, 0), slotEvent(slot, 1)));
    return (double)ms;
}

// Register buffers of up to 1 MiB (registers of up to 16 FP64 / 17 FP32 qubits)
// come from per-size free lists: QuEST's tests and QuantumLab's per-test
// registers create and destroy such registers by the hundred thousand, and a
// hipFree synchronises the device each time. A buffer returned to a list may
// still be read or written by work queued on the library stream; every later
// use of it is queued on the same stream, so it runs after that work.
namespace {
constexpr int kPoolClasses = 17;          // 16 B .. 1 MiB, powers of two
constexpr size_t kPoolMaxPerClass = 256;  // buffers kept per size
std::vector<void *> g_pool[kPoolClasses];
std::mutex g_poolMu;

int poolClass(size_t bytes) {
    int c = 4;
    while (((size_t)1 << c) < bytes) ++c;
    return c - 4;
}
}  // namespace

void *qmxTryDevAlloc(size_t bytes);

namespace qmx {
void *poolAlloc(size_t bytes) {
    if (bytes <= ((size_t)1 << (kPoolClasses + 3))) {
        const int c = poolClass(bytes);
        {
            std::lock_guard<std::mutex> lk(g_poolMu);
            if (!g_pool[c].empty()) {
                void *p = g_pool[c].back();
                g_pool[c].pop_back();
                return p;
            }
        }
        return qmxTryDevAlloc((size_t)16 << c);
    }
    return qmxTryDevAlloc(bytes);
}

void poolFree(void *p, size_t bytes) {
    if (!p) return;
    if (bytes <= ((size_t)1 << (kPoolClasses + 3))) {
        const int c = poolClass(bytes);
        std::lock_guard<std::mutex> lk(g_poolMu);
        if (g_pool[c].size() < kPoolMaxPerClass) {
            g_pool[c].push_back(p);
            return;
        }
    }
    devFree(p);
}
}  // namespace qmx

// allocation that reports failure instead of aborting (register creation
// turns it into QuEST's "Could not allocate memory for Qureg on GPU" error)
void *qmxTryDevAlloc(size_t bytes) {
    qmx::runtimeInit();
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}
