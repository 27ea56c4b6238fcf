// HIP runtime plumbing for libquest_mi355x: one device, one non-blocking stream
// (default priority), loud failure on any HIP error, event timing.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "qmx_device.h"

namespace qmx {

namespace {
std::once_flag g_once;
hipStream_t g_stream = nullptr;
double *g_pinned = nullptr;
hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
hipDeviceProp_t g_prop;
void (*g_pendingWork)() = nullptr;

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
    });
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
