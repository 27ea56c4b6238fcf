// Measurement collapse and density-matrix channels.
//
// A density matrix of n qubits is a 2n-qubit vector; element (r, c) sits at
// r + c*2^n, so "row bit q" is bit q and "column bit q" is bit q+n. A
// one-qubit channel on q touches quads {00, 01, 10, 11} of (row bit, column
// bit). Each channel is a single pass over its quads; the reference GPU
// backend ran a separate dephasing pass before depolarising / damping
// (QuEST_gpu.cu:3027-3078).
#include "qmx_device.h"

namespace qmx {

namespace {

constexpr int kPer = 4;

// statevec collapse (QuEST_cpu.c:3750-3800): keep half × 1/sqrt(p), zero the other
template <typename T>
__global__ void __launch_bounds__(kBlock) kCollapse(Amp<T> *__restrict__ a, long long work, int qubit,
                                                    int outcome, double renorm) {
    const uint64_t bit = 1ull << qubit;
    long long w = (long long)blockIdx.x * kBlock * kPer + threadIdx.x;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        long long wj = w + (long long)j * kBlock;
        if (wj < work) {
            uint64_t i0 = insertZero((uint64_t)wj, qubit);
            uint64_t keep = outcome ? (i0 | bit) : i0, kill = outcome ? i0 : (i0 | bit);
            Amp<T> v = a[keep];
            a[keep] = Amp<T>{(T)(v.x * renorm), (T)(v.y * renorm)};
            a[kill] = Amp<T>{(T)0, (T)0};
        }
    }
}

// density collapse (QuEST_cpu.c:798-877): keep |o><o| block / p, zero the rest
template <typename T>
__global__ void __launch_bounds__(kBlock) kDensCollapse(Amp<T> *__restrict__ a, long long n, int rowBit,
                                                        int colBit, int outcome, double prob) {
    long long i = (long long)blockIdx.x * kBlock * kPer + threadIdx.x;
#pragma unroll
    for (int j = 0; j < kPer; ++j, i += kBlock) {
        if (i < n) {
            int r = (int)((i >> rowBit) & 1), c = (int)((i >> colBit) & 1);
            if (r == outcome && c == outcome) {
                Amp<T> v = a[i];
                a[i] = Amp<T>{(T)(v.x / prob), (T)(v.y / prob)};
            } else {
                a[i] = Amp<T>{(T)0, (T)0};
            }
        }
    }
}

// MODE 0 dephase, 1 depolarise, 2 damp (QuEST_cpu.c:61-95, 138-235)
template <typename T, int MODE>
__global__ void __launch_bounds__(kBlock) kChannel1(Amp<T> *__restrict__ a, long long work, BitList fixed,
                                                    uint64_t rowM, uint64_t colM, double p1, double p2) {
    long long w = (long long)blockIdx.x * kBlock * kPer + threadIdx.x;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        long long wj = w + (long long)j * kBlock;
        if (wj >= work) continue;
        const uint64_t b00 = insertZeros((uint64_t)wj, fixed);
        const uint64_t b01 = b00 | rowM, b10 = b00 | colM, b11 = b00 | rowM | colM;
        Amp<T> x01 = a[b01], x10 = a[b10];
        const double off = p1;  // factor on the off-diagonal pair
        a[b01] = Amp<T>{(T)(off * x01.x), (T)(off * x01.y)};
        a[b10] = Amp<T>{(T)(off * x10.x), (T)(off * x10.y)};
        if (MODE == 1) {
            // depolarise: d = p2, retain = 1-d; diag -> retain*x + d*(x00+x11)/2
            Amp<T> x00 = a[b00], x11 = a[b11];
            double ar = ((double)x00.x + x11.x) / 2, ai = ((double)x00.y + x11.y) / 2;
            double keep = 1 - p2;
            a[b00] = Amp<T>{(T)(keep * x00.x + p2 * ar), (T)(keep * x00.y + p2 * ai)};
            a[b11] = Amp<T>{(T)(keep * x11.x + p2 * ar), (T)(keep * x11.y + p2 * ai)};
        } else if (MODE == 2) {
            // damp: g = p2; |0><0| += g |1><1|; |1><1| *= (1-g)
            Amp<T> x00 = a[b00], x11 = a[b11];
            a[b00] = Amp<T>{(T)(x00.x + p2 * x11.x), (T)(x00.y + p2 * x11.y)};
            a[b11] = Amp<T>{(T)((1 - p2) * x11.x), (T)((1 - p2) * x11.y)};
        }
    }
}

// two-qubit dephasing (QuEST_cpu.c:97-136): scale where either qubit's row and
// column bits differ
template <typename T>
__global__ void __launch_bounds__(kBlock) kDephase2(Amp<T> *__restrict__ a, long long n, int q1, int q2,
                                                    int shift, double retain) {
    long long i = (long long)blockIdx.x * kBlock * kPer + threadIdx.x;
#pragma unroll
    for (int j = 0; j < kPer; ++j, i += kBlock) {
        if (i < n) {
            bool m1 = ((i >> q1) & 1) != ((i >> (q1 + shift)) & 1);
            bool m2 = ((i >> q2) & 1) != ((i >> (q2 + shift)) & 1);
            if (m1 || m2) {
                Amp<T> v = a[i];
                a[i] = Amp<T>{(T)(retain * v.x), (T)(retain * v.y)};
            }
        }
    }
}

BitList bits2(int p, int q) {
    BitList b;
    b.n = 2;
    b.pos[0] = p < q ? p : q;
    b.pos[1] = p < q ? q : p;
    return b;
}

template <int MODE>
void channel1(QuregDev *q, int qubit, double p1, double p2) {
    const int shift = q->numQubitsRep;
    const long long work = q->numAmps >> 2;
    BitList fixed = bits2(qubit, qubit + shift);
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL((kChannel1<T, MODE>), dim3(gridFor(work, kPer)), dim3(kBlock), 0, hstream(),
                           (Amp<T> *)q->amps, work, fixed, 1ull << qubit, 1ull << (qubit + shift), p1, p2);
    });
    checkLaunch("density channel");
}

}  // namespace

void collapse(QuregDev *q, int qubit, int outcome, double prob) {
    const long long work = q->numAmps >> 1;
    const double renorm = 1.0 / sqrt(prob);
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kCollapse<T>, dim3(gridFor(work, kPer)), dim3(kBlock), 0, hstream(),
                           (Amp<T> *)q->amps, work, qubit, outcome, renorm);
    });
    checkLaunch("collapse");
}

void densCollapse(QuregDev *q, int qubit, int outcome, double prob) {
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kDensCollapse<T>, dim3(gridFor(q->numAmps, kPer)), dim3(kBlock), 0, hstream(),
                           (Amp<T> *)q->amps, q->numAmps, qubit, qubit + q->numQubitsRep, outcome, prob);
    });
    checkLaunch("densCollapse");
}

void densDephase(QuregDev *q, int qubit, double retain) { channel1<0>(q, qubit, retain, 0.0); }

void densDepolarise(QuregDev *q, int qubit, double depolLevel) {
    channel1<1>(q, qubit, 1.0 - depolLevel, depolLevel);
}

void densDamp(QuregDev *q, int qubit, double damping) {
    channel1<2>(q, qubit, sqrt(1.0 - damping), damping);
}

void densTwoQubitDephase(QuregDev *q, int q1, int q2, double retain) {
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kDephase2<T>, dim3(gridFor(q->numAmps, kPer)), dim3(kBlock), 0, hstream(),
                           (Amp<T> *)q->amps, q->numAmps, q1, q2, q->numQubitsRep, retain);
    });
    checkLaunch("densTwoQubitDephase");
}

}  // namespace qmx
