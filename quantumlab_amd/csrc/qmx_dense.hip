// A dense 2^B x 2^B complex matrix on B arbitrary bits of a register (B = 2..5),
// one group of 2^B amplitudes per thread held in registers, the matrix read
// through scalar loads (the same entry for every lane of the wave).
//
// Used for the density registers' composed superoperators (B = 2 or 4: the row
// and column bits of one or two qubits) and for uncontrolled 2- to 5-target
// unitaries on state vectors (multiQubitUnitary, twoQubitUnitary; QuEST_cpu.c:
// 1895-2007 gathers the same groups), where it replaces an LDS tile kernel
// that read a matrix row and the tile from LDS for every multiply-add.
//
// Matrices reach the device through a pinned ring of slots: a slot is
// rewritten only after the event recorded behind its last copy has fired, and
// copies and launches share the library stream, so a device slot is reused
// only after the kernel that read it has finished.
#include <algorithm>

#include "qmx_device.h"

namespace qmx {

namespace {

constexpr int kMaxB = 5;
constexpr int kMaxEntries = 1 << (2 * kMaxB);   // 1024 complex entries
constexpr int kRing = 256;
constexpr size_t kSlotBytes = kMaxEntries * sizeof(Amp<double>);

struct DenseSpec {
    int fix[kMaxB];                // ascending bit positions
    uint64_t off[1 << kMaxB];      // address offset of group entry e (bit j of e -> bit bits[j])
};

template <typename T, int B, bool NT>
__global__ void __launch_bounds__(kBlock) kDense(Amp<T> *__restrict__ a, long long groups, DenseSpec sp,
                                                 const Amp<T> *__restrict__ M) {
    constexpr int E = 1 << B;
    const long long g = xcdBlock() * kBlock + threadIdx.x;
    if (g >= groups) return;
    uint64_t base = (uint64_t)g;
#pragma unroll
    for (int j = 0; j < B; ++j) base = insertZero(base, sp.fix[j]);
    T xr[E], xi[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const Amp<T> v = ldA<T, NT>(a + (base | sp.off[e]));
        xr[e] = v.x;
        xi[e] = v.y;
    }
#pragma unroll
    for (int r = 0; r < E; ++r) {
        T yr = 0, yi = 0;
#pragma unroll
        for (int c = 0; c < E; ++c) {
            const Amp<T> m = M[r * E + c];
            yr = fma(m.x, xr[c], fma(-m.y, xi[c], yr));
            yi = fma(m.x, xi[c], fma(m.y, xr[c], yi));
        }
        stA<T, NT>(a + (base | sp.off[r]), Amp<T>{yr, yi});
    }
}

// Short runs (a target among the lowest bits): a workgroup of one wave takes
// 64 consecutive groups. Their 64 * 2^B amplitudes are the addresses base +
// deposit(j, V) with V = the lowest 6 free bits and the B target bits, sorted;
// V starts with bits 0-5, so load i of every lane covers 64 consecutive
// amplitudes (j = lane + 64 i). They go to LDS with one pad entry per 128 B,
// which keeps the strided gathers of the compute step free of bank conflicts;
// each lane then reads its own group (LDS index = its lane bits scattered to
// their places in V, or-ed with the entry's target bits), multiplies, writes
// it back, and the wave stores the block with the same coalesced pattern.
struct DenseLdsSpec {
    int fix[kMaxB];                // ascending target bit positions
    int lanePos[6];                // place in V of the lane's 6 bits
    unsigned eo[1 << kMaxB];       // place in V of entry e's target bits
    uint64_t hi[1 << kMaxB];       // address offset of load i: deposit(i, V[6..])
};

template <typename T, int B>
__global__ void __launch_bounds__(64) kDenseLds(Amp<T> *__restrict__ a, DenseLdsSpec sp,
                                                const Amp<T> *__restrict__ M) {
    constexpr int E = 1 << B;
    constexpr int PS = sizeof(T) == 8 ? 3 : 4;   // one pad entry per 128 B
    extern __shared__ unsigned char smemRaw[];
    Amp<T> *lds = reinterpret_cast<Amp<T> *>(smemRaw);
    const int l = threadIdx.x;
    uint64_t base = (uint64_t)xcdBlock() << 6;
#pragma unroll
    for (int j = 0; j < B; ++j) base = insertZero(base, sp.fix[j]);
    Amp<T> *src = a + base + l;
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int j = l + 64 * i;
        lds[j + (j >> PS)] = src[sp.hi[i]];
    }
    __syncthreads();
    unsigned sl = 0;
#pragma unroll
    for (int b = 0; b < 6; ++b) sl |= (unsigned)((l >> b) & 1) << sp.lanePos[b];
    T xr[E], xi[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const unsigned j = sl | sp.eo[e];
        const Amp<T> v = lds[j + (j >> PS)];
        xr[e] = v.x;
        xi[e] = v.y;
    }
#pragma unroll
    for (int r = 0; r < E; ++r) {
        T yr = 0, yi = 0;
#pragma unroll
        for (int c = 0; c < E; ++c) {
            const Amp<T> m = M[r * E + c];
            yr = fma(m.x, xr[c], fma(-m.y, xi[c], yr));
            yi = fma(m.x, xi[c], fma(m.y, xr[c], yi));
        }
        const unsigned j = sl | sp.eo[r];
        lds[j + (j >> PS)] = Amp<T>{yr, yi};
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int j = l + 64 * i;
        src[sp.hi[i]] = lds[j + (j >> PS)];
    }
}

struct MatrixRing {
    char *dev = nullptr, *host = nullptr;
    hipEvent_t ev[kRing];
    bool used[kRing] = {};
    int next = 0;
};
MatrixRing g_ring;

template <typename T>
const Amp<T> *stageMatrix(const Cplx *M, int entries, hipStream_t st) {
    if (!g_ring.dev) {
        g_ring.dev = static_cast<char *>(devAlloc(kRing * kSlotBytes));
        QMX_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&g_ring.host), kRing * kSlotBytes,
                                    hipHostMallocDefault));
        for (int i = 0; i < kRing; ++i)
            QMX_HIP_CHECK(hipEventCreateWithFlags(&g_ring.ev[i], hipEventDisableTiming));
    }
    const int slot = g_ring.next;
    g_ring.next = (slot + 1) % kRing;
    if (g_ring.used[slot]) QMX_HIP_CHECK(hipEventSynchronize(g_ring.ev[slot]));
    Amp<T> *h = reinterpret_cast<Amp<T> *>(g_ring.host + slot * kSlotBytes);
    for (int e = 0; e < entries; ++e) h[e] = Amp<T>{(T)M[e].re, (T)M[e].im};
    char *d = g_ring.dev + slot * kSlotBytes;
    QMX_HIP_CHECK(hipMemcpyAsync(d, h, entries * sizeof(Amp<T>), hipMemcpyHostToDevice, st));
    QMX_HIP_CHECK(hipEventRecord(g_ring.ev[slot], st));
    g_ring.used[slot] = true;
    return reinterpret_cast<const Amp<T> *>(d);
}

template <typename T, int B>
void launchDense(QuregDev *q, long long groups, const DenseSpec &sp, bool nt, const Cplx *M, hipStream_t st) {
    const Amp<T> *m = stageMatrix<T>(M, 1 << (2 * B), st);
    const unsigned grid = gridFor(groups, 1);
    if (nt)
        hipLaunchKernelGGL((kDense<T, B, true>), dim3(grid), dim3(kBlock), 0, st, (Amp<T> *)q->amps, groups, sp, m);
    else
        hipLaunchKernelGGL((kDense<T, B, false>), dim3(grid), dim3(kBlock), 0, st, (Amp<T> *)q->amps, groups, sp, m);
}

template <typename T, int B>
void launchDenseLds(QuregDev *q, long long groups, const DenseLdsSpec &sp, const Cplx *M, hipStream_t st) {
    const Amp<T> *m = stageMatrix<T>(M, 1 << (2 * B), st);
    constexpr int E = 1 << B;
    constexpr int PS = sizeof(T) == 8 ? 3 : 4;
    const size_t lds = (size_t)(64 * E + ((64 * E) >> PS)) * sizeof(Amp<T>);
    hipLaunchKernelGGL((kDenseLds<T, B>), dim3((unsigned)(groups >> 6)), dim3(64), lds, st, (Amp<T> *)q->amps, sp, m);
}

// the LDS form for targets among the bits whose runs are shorter than 128 B
void denseLds(QuregDev *q, const int *bits, int B, const int *sortedBits, const Cplx *M) {
    const int n = q->numQubits;
    DenseLdsSpec sp{};
    for (int j = 0; j < B; ++j) sp.fix[j] = sortedBits[j];
    // V: the lowest 6 free bits and the target bits, ascending
    uint64_t tmask = 0;
    for (int j = 0; j < B; ++j) tmask |= 1ull << bits[j];
    int V[6 + kMaxB], nv = 0, freeSeen = 0;
    for (int p = 0; p < n && nv < 6 + B; ++p) {
        if ((tmask >> p) & 1) {
            V[nv++] = p;
        } else if (freeSeen < 6) {
            sp.lanePos[freeSeen++] = nv;
            V[nv++] = p;
        }
    }
    auto placeOf = [&](int bit) {
        for (int i = 0; i < nv; ++i)
            if (V[i] == bit) return i;
        return -1;
    };
    for (int e = 0; e < (1 << B); ++e) {
        unsigned eo = 0;
        for (int j = 0; j < B; ++j)
            if ((e >> j) & 1) eo |= 1u << placeOf(bits[j]);
        sp.eo[e] = eo;
    }
    for (int i = 0; i < (1 << B); ++i) {
        uint64_t h = 0;
        for (int b = 0; b < B; ++b)
            if ((i >> b) & 1) h |= 1ull << V[6 + b];
        sp.hi[i] = h;
    }
    const long long groups = q->numAmps >> B;
    hipStream_t st = hstream();
    QMX_DISPATCH(q->prec, T, {
        switch (B) {
        case 2: launchDenseLds<T, 2>(q, groups, sp, M, st); break;
        case 3: launchDenseLds<T, 3>(q, groups, sp, M, st); break;
        case 4: launchDenseLds<T, 4>(q, groups, sp, M, st); break;
        default: launchDenseLds<T, 5>(q, groups, sp, M, st); break;
        }
    });
    checkLaunch("applyDense (LDS)");
}

}  // namespace

bool denseBitsSupported(int B) { return B >= 2 && B <= kMaxB; }

void applyDense(QuregDev *q, const int *bits, int B, const Cplx *M) {
    if (!denseBitsSupported(B)) {
        std::fprintf(stderr, "libquest_mi355x: applyDense supports 2 to %d bits (got %d)\n", kMaxB, B);
        std::abort();
    }
    DenseSpec sp{};
    for (int e = 0; e < (1 << B); ++e) {
        uint64_t off = 0;
        for (int j = 0; j < B; ++j)
            if ((e >> j) & 1) off |= 1ull << bits[j];
        sp.off[e] = off;
    }
    int pos[kMaxB];
    for (int j = 0; j < B; ++j) pos[j] = bits[j];
    std::sort(pos, pos + B);
    for (int j = 0; j < B; ++j) sp.fix[j] = pos[j];
    const long long groups = q->numAmps >> B;
    const bool nt = useNT(sp.fix[0], q->prec);
    if (!nt && q->numQubits >= B + 6) {
        denseLds(q, bits, B, sp.fix, M);
        return;
    }
    hipStream_t st = hstream();
    QMX_DISPATCH(q->prec, T, {
        switch (B) {
        case 2: launchDense<T, 2>(q, groups, sp, nt, M, st); break;
        case 3: launchDense<T, 3>(q, groups, sp, nt, M, st); break;
        case 4: launchDense<T, 4>(q, groups, sp, nt, M, st); break;
        default: launchDense<T, 5>(q, groups, sp, nt, M, st); break;
        }
    });
    checkLaunch("applyDense");
}

}  // namespace qmx
