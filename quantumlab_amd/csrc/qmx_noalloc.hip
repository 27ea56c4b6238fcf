// An allocation-free form of calcProbOfAllOutcomes, for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// hosts that forbid hipMalloc). No counterpart in the reference.
//
// The histogram kernels (qmx_reduce.hip) need device scratch for per-chunk
// partials and the results; calcProbOfAllOutcomes keeps a growable buffer for
// that. Here the scratch is a fixed file-scope device array of
// kNoAllocScratchDoubles doubles (8 MiB, held by each GPU the library runs on
// until the process ends). Its device address comes from hipGetSymbolAddress,
// once per device, with that device current; host code uses only that address,
// never the array's name. Every kernel writes every scratch entry it later
// reads, so nothing is cleared between calls. Calls that overlap on one GPU
// share the array and race.
#include <cstdio>

#include "qmx_device.h"
#include "quest_mi355x.h"
#include "quest_validation.h"

namespace qmx {
namespace {

constexpr long long kNoAllocScratchDoubles = 1ll << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

double *g_scratchAddr[kMaxDevices];  // device address of `scratch`, per device

// the current device's address of the scratch array (looked up on the first
// call on that device)
double *scratchAddress(const char *caller) {
    int dev = 0;
    QMX_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        std::fprintf(stderr, "libquest_mi355x: %s: device %d beyond the %d this library tracks\n", caller, dev,
                     kMaxDevices);
        std::abort();
    }
    if (!g_scratchAddr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) hipFatal(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        g_scratchAddr[dev] = static_cast<double *>(p);
    }
    return g_scratchAddr[dev];
}

__global__ void __launch_bounds__(kBlock) kToFloat(const double *__restrict__ in, float *__restrict__ out,
                                                   long long n) {
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

}  // namespace
}  // namespace qmx

extern "C" void qmxCalcProbOfAllOutcomesNoAlloc(qreal *outcomeProbs, Qureg qureg, int *qubits, int numQubits) {
    using namespace qmx;
    try {
        qv::multiTargets(qureg, qubits, numQubits, __func__);
        const QuregDev *d = dev(qureg);
        const long long numOut = 1ll << numQubits;
        const long long floatDoubles = sizeof(qreal) == sizeof(float) ? (numOut + 1) / 2 : 0;
        const long long need = outcomeScratchDoubles(d, qubits, numQubits) + floatDoubles;
        if (need > kNoAllocScratchDoubles)
            qv::raise("Too many outcomes for the allocation-free histogram: its fixed device scratch holds "
                      "1048576 doubles (at most 19 qubits in all; fewer when the register is large and the "
                      "qubits are low). Use calcProbOfAllOutcomes.",
                      __func__);
        double *s = scratchAddress(__func__);
        const double *res = probOfAllOutcomesDevice(d, qubits, numQubits, s);
        const void *src = res;
        if (sizeof(qreal) == sizeof(float)) {
            float *f = reinterpret_cast<float *>(s + (need - floatDoubles));
            hipLaunchKernelGGL(kToFloat, dim3((unsigned)((numOut + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                               hstream(), res, f, numOut);
            checkLaunch("qmxCalcProbOfAllOutcomesNoAlloc");
            src = f;
        }
        QMX_HIP_CHECK(hipMemcpyAsync(outcomeProbs, src, (size_t)numOut * sizeof(qreal), hipMemcpyDeviceToHost,
                                     hstream()));
        QMX_HIP_CHECK(hipStreamSynchronize(hstream()));
    } catch (const qmx::InputError &) {
        return;
    }
}
