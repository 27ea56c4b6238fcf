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
static_assert(kMaxEntries * sizeof(Amp<double>) <= kStageSlotBytes, "a dense matrix fits a staging slot");

struct DenseSpec {
    int fix[kMaxB];                // ascending bit positions
    uint64_t off[1 << kMaxB];      // address offset of group entry e (bit j of e -> bit bits[j])
};

template <typename T, int B, bool NT>
__global__ void __launch_bounds__(kBlock) kDense(Amp<T> *__restrict__ a, long long groups, DenseSpec sp,
                                                 const Amp<T> *__restrict__ M) {
    constexpr int E = 1 << B;
    const long long g = flatBlock() * kBlock + threadIdx.x;
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
    uint64_t base = (uint64_t)flatBlock() << 6;
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

// Two matrices on disjoint bit sets A (BA bits) and B (BB bits) in one sweep:
// they commute, and a workgroup can hold every A x B combination. The tile is
// 4096 entries — the A and B bits plus the lowest other bits. Each thread
// loads its 16 / 2^BA A-groups (a group: the 2^BA entries that differ only in
// A bits) straight from HBM, applies A and writes them to LDS; after a barrier
// it reads its B-groups from LDS, applies B and stores them straight back. A
// thread's 8 index bits take the lowest non-group tile bits, so a load or
// store instruction covers runs of the lowest free bits. One HBM read + write
// and one LDS round trip for two operations (the density registers' paired
// superoperator sweeps). Staging the whole tile through LDS on the way in and
// out, with coalesced tile-order accesses, measured 3 % slower (FP64, 14-qubit
// density) for two more LDS round trips.
constexpr int kPairTileBits = 12;

struct PairPhase {
    int thr[8];          // tile bit of thread-index bit j
    unsigned g[4];       // tile-index offset of the thread's group g
    unsigned e[16];      // tile-index offset of group entry e (bit i of e -> the matrix's bit i)
    uint64_t thrG[8], gG[4], eG[16];  // the same offsets deposited to global bit positions
};

struct DensePairSpec {
    BitList tile;        // the tile's global bit positions, ascending
    uint64_t lowOff[8];  // global offset of tile-index bit j < 8
    uint64_t hi[16];     // global offset of tile index 256 i
    PairPhase pa, pb;
};

// global index of a tile's entry 0: the tile index with zeros inserted at the
// tile's kPairTileBits positions. The count is a constant, so the loop unrolls
// and the positions stay in scalar registers (a loop over a run-time count
// re-reads each position from the kernel arguments).
__device__ __forceinline__ uint64_t pairTileBase(uint64_t tile, const BitList &b) {
#pragma unroll
    for (int j = 0; j < kPairTileBits; ++j) tile = insertZero(tile, b.pos[j]);
    return tile;
}

template <typename T, int BX, bool IN_GLOBAL>
__device__ __forceinline__ void pairPhase(Amp<T> *__restrict__ a, uint64_t tileBase, Amp<T> *lds, const PairPhase &ph,
                                          const Amp<T> *__restrict__ M) {
    constexpr int PS = sizeof(T) == 8 ? 3 : 4;
    constexpr int E = 1 << BX, G = 16 / E;
    const unsigned t = threadIdx.x;
    unsigned bt = 0;
    uint64_t btG = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        bt |= ((t >> j) & 1u) << ph.thr[j];
        if ((t >> j) & 1) btG |= ph.thrG[j];
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const unsigned gb = bt | ph.g[g];
        const uint64_t gbG = tileBase | btG | ph.gG[g];
        T xr[E], xi[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            Amp<T> v;
            if (IN_GLOBAL) {
                v = a[gbG | ph.eG[e]];
            } else {
                const unsigned j = gb | ph.e[e];
                v = lds[j + (j >> PS)];
            }
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
            if (IN_GLOBAL) {
                const unsigned j = gb | ph.e[r];
                lds[j + (j >> PS)] = Amp<T>{yr, yi};
            } else {
                a[gbG | ph.eG[r]] = Amp<T>{yr, yi};
            }
        }
    }
}

template <typename T, int BA, int BB>
__global__ void __launch_bounds__(kBlock) kDensePair(Amp<T> *__restrict__ a, DensePairSpec sp,
                                                     const Amp<T> *__restrict__ MA, const Amp<T> *__restrict__ MB) {
    extern __shared__ unsigned char smemRaw[];
    Amp<T> *lds = reinterpret_cast<Amp<T> *>(smemRaw);
    const uint64_t tileBase = insertZeros((uint64_t)xcdBlock(), sp.tile);
    pairPhase<T, BA, true>(a, tileBase, lds, sp.pa, MA);
    __syncthreads();
    pairPhase<T, BB, false>(a, tileBase, lds, sp.pb, MB);
}

// FP32 pairs of 2-qubit superoperators (4 + 4 bits) on the matrix cores. The
// VALU form spends 32 complex FMAs per entry, which at the non-packed FP32
// rate takes longer than the sweep's HBM traffic. Here a 16x16 superoperator S
// in real form R = [[Sr, -Si], [Si, Sr]] (32 x 32) multiplies 16 groups at a
// time as 2 row tiles x 8 K-steps of v_mfma_f32_16x16x4f32:
//   A (16 x 4): lane l holds R[16 rt + l%16][4 s + l/16], kept in registers;
//   B (4 x 16): lane l holds entry 4 s' + l/16 of group l%16 (real part for
//               K-step s', imaginary part for K-step s' + 4);
//   D (16 x 16): lane l holds entries 4 (l/16) + v of group l%16 (real parts
//               from row tile 0, imaginary from row tile 1).
// The tile, its two phases and their memory paths are kDensePair's.
template <typename T> struct MfmaAcc { typedef T type __attribute__((ext_vector_type(4))); };

__device__ __forceinline__ MfmaAcc<float>::type mfma16x16x4(float a, float b, MfmaAcc<float>::type c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ MfmaAcc<double>::type mfma16x16x4(double a, double b, MfmaAcc<double>::type c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// output row i of a lane's accumulator element v: FP32 4 (l/16) + v, FP64
// l/16 + 4 v (measured on gfx950 against a host product)
template <typename T> __device__ __forceinline__ int mfmaOutRow(int q, int v) {
    return sizeof(T) == 4 ? 4 * q + v : q + 4 * v;
}

template <typename T>
__device__ __forceinline__ void pairFragments(const Amp<T> *__restrict__ M, int col, int q, T (&f)[2][8]) {
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const Amp<T> m = M[col * 16 + ((4 * s + q) & 15)];
        f[0][s] = s < 4 ? m.x : -m.y;
        f[1][s] = s < 4 ? m.y : m.x;
    }
}

// element k of {v[0], v[1], v[2], v[3]} for a lane-dependent k in 0..3
template <typename V>
__device__ __forceinline__ V pick4(V v0, V v1, V v2, V v3, int k) {
    return k == 0 ? v0 : k == 1 ? v1 : k == 2 ? v2 : v3;
}

// per-lane offsets of one phase: entry s of the lane's input (4 s + q) and
// entry v of its output (4 q + v) of group col, in tile-index (LDS) or global
// form; a batch b adds the uniform group-id bits 4-7 (batchBits)
struct MfmaLayout {
    unsigned t[4];     // tile index of the lane's LDS entries
    uint64_t g[4];     // global offset of the lane's HBM entries
};

// LDS side from the entry set E0 (inputs: E0 = 4 s + q; outputs: mfmaOutRow),
// HBM side from the other
template <typename T>
__device__ __forceinline__ MfmaLayout mfmaLayout(const PairPhase &ph, int col, int q, bool ldsIsOutput) {
    unsigned lt = 0;
    uint64_t lg = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if ((col >> j) & 1) {
            lt |= 1u << ph.thr[j];
            lg |= ph.thrG[j];
        }
    MfmaLayout L;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const unsigned inT = pick4(ph.e[4 * s], ph.e[4 * s + 1], ph.e[4 * s + 2], ph.e[4 * s + 3], q);
        const unsigned outT = sizeof(T) == 4 ? pick4(ph.e[s], ph.e[4 + s], ph.e[8 + s], ph.e[12 + s], q) : inT;
        const uint64_t inG = pick4(ph.eG[4 * s], ph.eG[4 * s + 1], ph.eG[4 * s + 2], ph.eG[4 * s + 3], q);
        const uint64_t outG = sizeof(T) == 4 ? pick4(ph.eG[s], ph.eG[4 + s], ph.eG[8 + s], ph.eG[12 + s], q) : inG;
        L.t[s] = lt | (ldsIsOutput ? outT : inT);
        L.g[s] = lg | (ldsIsOutput ? inG : outG);
    }
    return L;
}

__device__ __forceinline__ void batchBits(const PairPhase &ph, int b, unsigned &bt, uint64_t &bg) {
    bt = 0;
    bg = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if ((b >> j) & 1) {
            bt |= 1u << ph.thr[4 + j];
            bg |= ph.thrG[4 + j];
        }
}

// 16 groups: OUT = R X on the matrix cores; x holds the lane's 4 input entries
template <typename T>
__device__ __forceinline__ void mfmaGroups(const T (&f)[2][8], const Amp<T> (&x)[4], typename MfmaAcc<T>::type &re,
                                           typename MfmaAcc<T>::type &im) {
    re = typename MfmaAcc<T>::type{0, 0, 0, 0};
    im = typename MfmaAcc<T>::type{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        re = mfma16x16x4(f[0][s], x[s].x, re);
        im = mfma16x16x4(f[1][s], x[s].x, im);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        re = mfma16x16x4(f[0][s + 4], x[s].y, re);
        im = mfma16x16x4(f[1][s + 4], x[s].y, im);
    }
}

constexpr int kMfmaBatches = 16 / (kBlock / 64);  // 16-group batches per wave and phase

// a persistent grid: each workgroup walks tiles w = blockIdx.x + k gridDim.x.
// The first superoperator runs on inputs already in registers; the next
// tile's inputs are loaded as soon as those registers are free, so the loads
// are in flight during the second superoperator and its HBM stores.
template <typename T>
__global__ void __launch_bounds__(kBlock) kDensePairMfma(Amp<T> *__restrict__ a, unsigned numTiles, DensePairSpec sp,
                                                         const Amp<T> *__restrict__ MA,
                                                         const Amp<T> *__restrict__ MB, unsigned perGroup) {
    constexpr int PS = sizeof(T) == 8 ? 3 : 4;
    extern __shared__ unsigned char smemRaw[];
    Amp<T> *lds = reinterpret_cast<Amp<T> *>(smemRaw);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, q = lane >> 4;
    T fa[2][8], fb[2][8];
    pairFragments(MA, col, q, fa);
    pairFragments(MB, col, q, fb);
    const MfmaLayout LA = mfmaLayout<T>(sp.pa, col, q, true), LB = mfmaLayout<T>(sp.pb, col, q, false);
    Amp<T> x[kMfmaBatches][4];
    auto loadInputs = [&](unsigned tile) {
        const Amp<T> *src = a + pairTileBase((uint64_t)tile, sp.tile);
#pragma unroll
        for (int i = 0; i < kMfmaBatches; ++i) {
            unsigned bt;
            uint64_t bg;
            batchBits(sp.pa, wave + i * (kBlock / 64), bt, bg);
#pragma unroll
            for (int s = 0; s < 4; ++s) x[i][s] = src[bg | LA.g[s]];
        }
    };
    // perGroup = 0: a persistent grid striding over the tiles; otherwise
    // perGroup consecutive tiles per workgroup (XCD order)
    unsigned w = perGroup ? (unsigned)xcdBlock() * perGroup : blockIdx.x;
    const unsigned wEnd = perGroup ? min(numTiles, w + perGroup) : numTiles;
    const unsigned stride = perGroup ? 1u : gridDim.x;
    if (w < wEnd) loadInputs(w);
    for (; w < wEnd; w += stride) {
        // first superoperator: registers -> LDS
#pragma unroll
        for (int i = 0; i < kMfmaBatches; ++i) {
            unsigned bt;
            uint64_t bg;
            batchBits(sp.pa, wave + i * (kBlock / 64), bt, bg);
            typename MfmaAcc<T>::type re, im;
            mfmaGroups(fa, x[i], re, im);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const unsigned j = bt | LA.t[v];
                lds[j + (j >> PS)] = Amp<T>{re[v], im[v]};
            }
        }
        if (w + stride < wEnd) loadInputs(w + stride);
        __syncthreads();
        // second superoperator: LDS -> HBM
        Amp<T> *dst = a + pairTileBase((uint64_t)w, sp.tile);
#pragma unroll
        for (int i = 0; i < kMfmaBatches; ++i) {
            unsigned bt;
            uint64_t bg;
            batchBits(sp.pb, wave + i * (kBlock / 64), bt, bg);
            Amp<T> y[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const unsigned j = bt | LB.t[s];
                y[s] = lds[j + (j >> PS)];
            }
            typename MfmaAcc<T>::type re, im;
            mfmaGroups(fb, y, re, im);
#pragma unroll
            for (int v = 0; v < 4; ++v) dst[bg | LB.g[v]] = Amp<T>{re[v], im[v]};
        }
        __syncthreads();  // the next tile's first phase overwrites the LDS copy
    }
}

// matrices reach the device through the runtime's staging ring (no
// allocation or synchronisation per launch)
template <typename T>
const Amp<T> *stageMatrix(const Cplx *M, int entries, hipStream_t st) {
    const size_t bytes = entries * sizeof(Amp<T>);
    const Staged s = stageBegin(bytes);
    Amp<T> *h = static_cast<Amp<T> *>(s.host);
    for (int e = 0; e < entries; ++e) h[e] = Amp<T>{(T)M[e].re, (T)M[e].im};
    stageCommit(s, bytes, st);
    return static_cast<const Amp<T> *>(s.dev);
}

template <typename T, int B>
void launchDense(QuregDev *q, long long groups, const DenseSpec &sp, bool nt, const Cplx *M, hipStream_t st) {
    const Amp<T> *m = stageMatrix<T>(M, 1 << (2 * B), st);
    const dim3 grid = gridFor(groups, 1);
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
    hipLaunchKernelGGL((kDenseLds<T, B>), gridOf(groups >> 6), dim3(64), lds, st, (Amp<T> *)q->amps, sp, m);
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

// tile-index layout of one phase: the matrix's bits sit at tile positions
// pos[0..BX-1]; a thread's 8 index bits and its groups take the other bits
PairPhase pairLayout(const int *pos, int BX) {
    PairPhase ph{};
    int other[kPairTileBits], no = 0;
    for (int b = 0; b < kPairTileBits; ++b)
        if (std::find(pos, pos + BX, b) == pos + BX) other[no++] = b;
    for (int j = 0; j < 8; ++j) ph.thr[j] = other[j];
    for (int g = 0; g < (16 >> BX); ++g) {
        unsigned o = 0;
        for (int j = 0; 8 + j < no; ++j)
            if ((g >> j) & 1) o |= 1u << other[8 + j];
        ph.g[g] = o;
    }
    for (int e = 0; e < (1 << BX); ++e) {
        unsigned o = 0;
        for (int i = 0; i < BX; ++i)
            if ((e >> i) & 1) o |= 1u << pos[i];
        ph.e[e] = o;
    }
    return ph;
}

void pairGlobal(PairPhase &ph, const int *tpos, int BX) {
    auto dep = [&](unsigned o) {
        uint64_t g = 0;
        for (int b = 0; b < kPairTileBits; ++b)
            if ((o >> b) & 1) g |= 1ull << tpos[b];
        return g;
    };
    for (int j = 0; j < 8; ++j) ph.thrG[j] = 1ull << tpos[ph.thr[j]];
    for (int g = 0; g < (16 >> BX); ++g) ph.gG[g] = dep(ph.g[g]);
    for (int e = 0; e < (1 << BX); ++e) ph.eG[e] = dep(ph.e[e]);
}

template <typename T, int BA, int BB>
void launchDensePair(QuregDev *q, const DensePairSpec &sp, const Cplx *MA, const Cplx *MB, hipStream_t st) {
    constexpr int PS = sizeof(T) == 8 ? 3 : 4;
    constexpr int entries = 1 << kPairTileBits;
    const size_t lds = (size_t)(entries + (entries >> PS)) * sizeof(Amp<T>);
    // 4 + 4 bits: the matrix-core form; 4 + 2 bits: the VALU form
    constexpr bool mfma = BA == 4 && BB == 4;
    const void *kern;
    if constexpr (mfma) kern = reinterpret_cast<const void *>(&kDensePairMfma<T>);
    else kern = reinterpret_cast<const void *>(&kDensePair<T, BA, BB>);
    static bool ldsAllowed = false;
    if (!ldsAllowed) {
        if (!allowDynamicLds(kern, lds, "applyDensePair")) std::abort();
        ldsAllowed = true;
    }
    // both matrices in one staging slot: A at 0, B at 256
    Cplx both[2 * 256];
    std::copy(MA, MA + (1 << (2 * BA)), both);
    std::copy(MB, MB + (1 << (2 * BB)), both + 256);
    const Amp<T> *m = stageMatrix<T>(both, 512, st);
    const unsigned grid = (unsigned)(q->numAmps >> kPairTileBits);
    // FP32: a persistent grid of 3 workgroups per CU (its occupancy); FP64:
    // 2 consecutive tiles per workgroup (at 14 qubits 0.36 s for the noisy
    // circuit against 0.39 s persistent, 2 workgroups per CU by its LDS; 4,
    // 8 and 16 tiles 0.37-0.38 s; FP32 gained nothing from it)
    if constexpr (mfma) {
        if constexpr (sizeof(T) == 8) {
            constexpr unsigned per = 2;
            hipLaunchKernelGGL(kDensePairMfma<T>, dim3(((grid + per - 1) / per + 7u) / 8u * 8u), dim3(kBlock), lds,
                               st, (Amp<T> *)q->amps, grid, sp, m, m + 256, per);
        } else {
            hipLaunchKernelGGL(kDensePairMfma<T>, dim3(std::min(grid, 3u * (unsigned)numCUs())), dim3(kBlock), lds,
                               st, (Amp<T> *)q->amps, grid, sp, m, m + 256, 0u);
        }
    } else
        hipLaunchKernelGGL((kDensePair<T, BA, BB>), dim3(grid), dim3(kBlock), lds, st, (Amp<T> *)q->amps, sp, m,
                           m + 256);
}

}  // namespace

bool densePairSupported(const QuregDev *q, int BA, int BB) {
    // 4 + 4 or 4 + 2 bits (two one-qubit superoperators are never pending
    // together: a lone one-qubit operation joins the pending one-qubit block)
    return ((BA == 4 && (BB == 2 || BB == 4)) || (BA == 2 && BB == 4)) && q->numQubits >= kPairTileBits;
}

void applyDensePair(QuregDev *q, const int *bitsA, int BA, const Cplx *MA, const int *bitsB, int BB,
                    const Cplx *MB) {
    if (!densePairSupported(q, BA, BB)) {
        std::fprintf(stderr, "libquest_mi355x: applyDensePair: unsupported shape (%d, %d bits on %d qubits)\n", BA,
                     BB, q->numQubits);
        std::abort();
    }
    if (BA < BB) {  // the matrices commute: the larger one first
        std::swap(bitsA, bitsB);
        std::swap(BA, BB);
        std::swap(MA, MB);
    }
    // tile bits: A, B and the lowest other bits
    uint64_t tm = 0;
    for (int i = 0; i < BA; ++i) tm |= 1ull << bitsA[i];
    for (int i = 0; i < BB; ++i) tm |= 1ull << bitsB[i];
    for (int b = 0; b < q->numQubits && __builtin_popcountll(tm) < kPairTileBits; ++b) tm |= 1ull << b;
    DensePairSpec sp{};
    int tpos[kPairTileBits];
    for (int b = 0, j = 0; b < q->numQubits; ++b)
        if ((tm >> b) & 1) {
            tpos[j] = b;
            sp.tile.pos[j++] = b;
        }
    sp.tile.n = kPairTileBits;
    for (int j = 0; j < 8; ++j) sp.lowOff[j] = 1ull << tpos[j];
    for (int i = 0; i < 16; ++i) {
        uint64_t o = 0;
        for (int k = 0; k < 4; ++k)
            if ((i >> k) & 1) o |= 1ull << tpos[8 + k];
        sp.hi[i] = o;
    }
    auto local = [&](int bit) { return (int)(std::find(tpos, tpos + kPairTileBits, bit) - tpos); };
    int pa[4], pb[4];
    for (int i = 0; i < BA; ++i) pa[i] = local(bitsA[i]);
    for (int i = 0; i < BB; ++i) pb[i] = local(bitsB[i]);
    sp.pa = pairLayout(pa, BA);
    sp.pb = pairLayout(pb, BB);
    pairGlobal(sp.pa, tpos, BA);
    pairGlobal(sp.pb, tpos, BB);
    hipStream_t st = hstream();
    QMX_DISPATCH(q->prec, T, {
        if (BB == 4) launchDensePair<T, 4, 4>(q, sp, MA, MB, st);
        else launchDensePair<T, 4, 2>(q, sp, MA, MB, st);
    });
    checkLaunch("applyDensePair");
}

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
