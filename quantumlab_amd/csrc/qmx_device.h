// Device-side helpers shared by the qmx_*.hip kernel files.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "qmx_internal.h"

namespace qmx {

// amplitude storage types: one 16 B (double) or 8 B (float) vector per amplitude
template <typename T> struct Amp;
template <> struct alignas(16) Amp<double> { double x, y; };
template <> struct alignas(8) Amp<float> { float x, y; };

constexpr int kBlock = 256;   // 4 wave64 per workgroup

// run a statement with T bound to the register's storage type
#define QMX_DISPATCH(prec, T, ...)                  \
    do {                                            \
        if ((prec) == FP64) {                       \
            using T = double;                       \
            __VA_ARGS__;                            \
        } else {                                    \
            using T = float;                        \
            __VA_ARGS__;                            \
        }                                           \
    } while (0)


#define QMX_HIP_CHECK(call)                                                            \
    do {                                                                               \
        hipError_t err_ = (call);                                                      \
        if (err_ != hipSuccess) ::qmx::hipFatal(err_, #call, __FILE__, __LINE__);      \
    } while (0)

[[noreturn]] void hipFatal(hipError_t err, const char *what, const char *file, int line);

inline hipStream_t hstream() { return reinterpret_cast<hipStream_t>(stream()); }

// 16 B (double) / 8 B (float) vector view of one amplitude, for single-instruction
// loads and stores, optionally nontemporal. Nontemporal streaming raised the
// 2x2 pair sweep at 32 qubits from 5.5 to 6.2-6.6 TB/s on MI355X when every
// wave's contiguous runs are >= 128 B; shorter runs (pair distance < 8
// amplitudes) are faster with ordinary accesses.
template <typename T> struct VecOf;
template <> struct VecOf<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct VecOf<float> { typedef float type __attribute__((ext_vector_type(2))); };

template <typename T, bool NT>
__device__ __forceinline__ Amp<T> ldA(const Amp<T> *p) {
    typedef typename VecOf<T>::type V;
    V v = NT ? __builtin_nontemporal_load(reinterpret_cast<const V *>(p)) : *reinterpret_cast<const V *>(p);
    return Amp<T>{v.x, v.y};
}

template <typename T, bool NT>
__device__ __forceinline__ void stA(Amp<T> *p, const Amp<T> &a) {
    typedef typename VecOf<T>::type V;
    V v;
    v.x = a.x;
    v.y = a.y;
    if (NT) __builtin_nontemporal_store(v, reinterpret_cast<V *>(p));
    else *reinterpret_cast<V *>(p) = v;
}

// XCD-aware workgroup order. MI355X dispatches consecutive workgroups
// round-robin over its 8 XCDs; remapping so that XCD x (= blockIdx % 8) owns
// the x-th contiguous eighth of the work lets each XCD stream one region.
// Measured +5-7% on streaming sweeps at 32 qubits.
__device__ __forceinline__ long long xcdBlock() {
    const long long b = blockIdx.x, nb = gridDim.x;
    if (nb & 7) return b;
    return (b & 7) * (nb >> 3) + (b >> 3);
}

// workgroup index in a grid from gridFor / gridOf
__device__ __forceinline__ long long flatBlock() { return (long long)blockIdx.y * gridDim.x + xcdBlock(); }
__device__ __forceinline__ long long flatBlockLinear() { return (long long)blockIdx.y * gridDim.x + blockIdx.x; }

// nontemporal accesses pay off once a wave's contiguous runs reach 128 B
inline bool useNT(int lowestFixedBit, int prec) {
    return ((1ll << lowestFixedBit) * (prec == FP64 ? 16 : 8)) >= 128;
}

// insert a zero bit at position p of x
__device__ __forceinline__ uint64_t insertZero(uint64_t x, int p) {
    uint64_t lo = x & ((1ull << p) - 1ull);
    return ((x >> p) << (p + 1)) | lo;
}

// insert zero bits at the ascending positions of a BitList
__device__ __forceinline__ uint64_t insertZeros(uint64_t x, const BitList &b) {
    for (int j = 0; j < b.n; ++j) x = insertZero(x, b.pos[j]);
    return x;
}

// parity of popcount
__device__ __forceinline__ int parity64(uint64_t x) { return __popcll(x) & 1; }

// Neumaier-compensated FP64 accumulator
struct Acc {
    double s, c;
    __device__ __forceinline__ Acc() : s(0.0), c(0.0) {}
    __device__ __forceinline__ void add(double v) {
        double t = s + v;
        if (fabs(s) >= fabs(v)) c += (s - t) + v;
        else c += (v - t) + s;
        s = t;
    }
    __device__ __forceinline__ void merge(double os, double oc) {
        add(os);
        c += oc;
    }
    __device__ __forceinline__ double value() const { return s + c; }
};

// Reduce NV compensated accumulators across the workgroup (kBlock threads).
// Thread 0 returns the block totals in out[]; deterministic order.
template <int NV>
__device__ __forceinline__ void blockReduce(Acc (&acc)[NV], double (&out)[NV]) {
    // wave64 butterfly on (s, c) pairs
#pragma unroll
    for (int v = 0; v < NV; ++v) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            double os = __shfl_xor(acc[v].s, off, 64);
            double oc = __shfl_xor(acc[v].c, off, 64);
            acc[v].merge(os, oc);
        }
    }
    __shared__ double lds[kBlock / 64][NV][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) { lds[wave][v][0] = acc[v].s; lds[wave][v][1] = acc[v].c; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            Acc a;
            for (int w = 0; w < kBlock / 64; ++w) a.merge(lds[w][v][0], lds[w][v][1]);
            out[v] = a.value();
        }
    }
}

// A 1-D index space as a launch grid. A dispatch counts work-items in 32 bits
// per dimension, so a 2^33-amplitude register cannot run one item per thread
// on a 1-D grid: rows of at most 2^23 workgroups (2^31 work-items of kBlock),
// as many rows as needed. Kernels take their workgroup index from
// flatBlock() (XCD-aware within a row; the row width is a multiple of 8 when
// there is more than one row) or flatBlockLinear().
inline dim3 gridOf(long long blocks) {
    constexpr long long kRow = 1ll << 23;
    if (blocks < 1) blocks = 1;
    if (blocks <= kRow) return dim3((unsigned)blocks);
    return dim3((unsigned)kRow, (unsigned)((blocks + kRow - 1) / kRow));
}

inline dim3 gridFor(long long work, int perThread) {
    const long long per = (long long)kBlock * perThread;
    return gridOf((work + per - 1) / per);
}

}  // namespace qmx
