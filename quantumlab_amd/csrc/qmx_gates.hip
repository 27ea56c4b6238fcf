// Gate kernels: every gate is one streaming sweep over exactly the amplitudes
// it changes.
//
// Index scheme: a launch enumerates work items w in [0, N / 2^b), where b is
// the number of "fixed" bit positions (target + controls, or the tile bits of
// applyMatrixN). Each w is expanded to a state index by inserting zeros at
// those positions (ascending) and then OR-ing the control mask, so controlled
// gates touch only the 2^(n-c) amplitudes whose controls are 1 instead of
// launching N/2 threads and idling most of them (QuEST_gpu.cu:1239-1308).
// Consecutive lanes get consecutive w, so every load/store instruction of a
// wave64 covers contiguous memory for all bits below the lowest fixed bit.
#include <algorithm>
#include <cstring>
#include <vector>

#include "qmx_device.h"

namespace qmx {

namespace {

template <typename T> struct C2 { T re, im; };

template <typename T>
__device__ __forceinline__ Amp<T> cmul2(const C2<T> &u0, const Amp<T> &a, const C2<T> &u1,
                                        const Amp<T> &b) {
    // u0*a + u1*b
    T re = u0.re * a.x - u0.im * a.y + u1.re * b.x - u1.im * b.y;
    T im = u0.re * a.y + u0.im * a.x + u1.re * b.y + u1.im * b.x;
    return Amp<T>{re, im};
}

template <typename T> struct M2 { C2<T> m00, m01, m10, m11; };

// One work item per thread, consecutive lanes on consecutive items, XCD-aware
// workgroup order, nontemporal 16 B accesses when runs are >= 128 B (NT): the
// combination measured fastest for the FP64 sweep at 32 qubits.

// ---- dense 2x2 on one target (optionally controlled) ----------------------
// QuEST_cpu.c:1737-1794 (compactUnitary), :2020-2062 (unitary), :2251-2317
template <typename T, bool CTRL, bool NT>
__global__ void __launch_bounds__(kBlock) kMatrix2(Amp<T> *__restrict__ a, long long work, int target,
                                                   BitList fixed, uint64_t setMask, M2<T> u) {
    const long long w = flatBlock() * kBlock + threadIdx.x;
    if (w >= work) return;
    const uint64_t i0 = CTRL ? (insertZeros((uint64_t)w, fixed) | setMask) : insertZero((uint64_t)w, target);
    const uint64_t i1 = i0 | (1ull << target);
    const Amp<T> up = ldA<T, NT>(a + i0), lo = ldA<T, NT>(a + i1);
    stA<T, NT>(a + i0, cmul2(u.m00, up, u.m01, lo));
    stA<T, NT>(a + i1, cmul2(u.m10, up, u.m11, lo));
}

// Targets 0-5: the partner of amplitude i (i ^ 2^t) is held by lane l ^ 2^t
// of the same wave, so each thread loads and stores only its own amplitude —
// one contiguous nontemporal stream, like an in-place scale — and takes the
// partner through a cross-lane shuffle. With controls (CTRL), amplitudes whose
// control bits differ from setMask are written back unchanged: used where the
// pair form, touching only the controlled ones, would still move whole lines.
// Needs whole waves: n is a multiple of the workgroup (at least 8 qubits).
template <typename T, bool CTRL>
__global__ void __launch_bounds__(kBlock) kMatrix2Lane(Amp<T> *__restrict__ a, long long n, int target,
                                                       uint64_t ctrlMask, uint64_t setMask, M2<T> u) {
    const long long i = flatBlock() * kBlock + threadIdx.x;
    if (i >= n) return;
    const Amp<T> v = ldA<T, true>(a + i);
    const int m = 1 << target;
    const Amp<T> p{__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64)};
    // bit t of i clear: v is the pair's upper amplitude, p its lower one
    Amp<T> o = ((i >> target) & 1) ? cmul2(u.m10, p, u.m11, v) : cmul2(u.m00, v, u.m01, p);
    if (CTRL && ((uint64_t)i & ctrlMask) != setMask) o = v;
    stA<T, true>(a + i, o);
}

// Targets 6-10, no controls: a workgroup owns 2048 consecutive amplitudes, 8
// per thread, loaded and stored as contiguous nontemporal streams; each
// thread reads its amplitudes' partners from LDS after one barrier. Needs
// registers of at least 11 qubits (whole spans).
constexpr int kSpanPer = 8, kSpan = kBlock * kSpanPer;

template <typename T>
__global__ void __launch_bounds__(kBlock) kMatrix2Span(Amp<T> *__restrict__ a, int target, M2<T> u) {
    __shared__ Amp<T> s[kSpan];
    Amp<T> *src = a + flatBlock() * kSpan;
    Amp<T> v[kSpanPer];
#pragma unroll
    for (int k = 0; k < kSpanPer; ++k) {
        v[k] = ldA<T, true>(src + threadIdx.x + k * kBlock);
        s[threadIdx.x + k * kBlock] = v[k];
    }
    __syncthreads();
    const int m = 1 << target;
#pragma unroll
    for (int k = 0; k < kSpanPer; ++k) {
        const int j = threadIdx.x + k * kBlock;
        const Amp<T> p = s[j ^ m];
        const Amp<T> o = ((j >> target) & 1) ? cmul2(u.m10, p, u.m11, v[k]) : cmul2(u.m00, v[k], u.m01, p);
        stA<T, true>(src + j, o);
    }
}

// ---- uniformly controlled 2x2 ------------------------------------------------
// Matrix k of a table acts on the target wherever the controls read k (bit b
// of k = qubit ctrl[b]). QuantumLab builds the rotation forms as 2^c rotations
// and CNOTs (FRy / FRz / FRyz, FRGate.cu:181-460) or X flips and
// multi-controlled phases (FRp, FRGate.cu:460-515), one sweep each; here it is
// one sweep in all. The table (2^c x 4 complex) is read through the caches:
// neighbouring pairs share k unless a control is among the lowest bits.
constexpr int kMaxUniformControls = 16;
struct CtrlBits {
    int n;
    int pos[kMaxUniformControls];
};

template <typename T, bool NT>
__global__ void __launch_bounds__(kBlock) kUniform2(Amp<T> *__restrict__ a, long long work, int target,
                                                    CtrlBits ctrl, const M2<T> *__restrict__ mats) {
    const long long w = flatBlock() * kBlock + threadIdx.x;
    if (w >= work) return;
    const uint64_t i0 = insertZero((uint64_t)w, target), i1 = i0 | (1ull << target);
    int k = 0;
    for (int b = 0; b < ctrl.n; ++b) k |= (int)((i0 >> ctrl.pos[b]) & 1) << b;
    const M2<T> u = mats[k];
    const Amp<T> up = ldA<T, NT>(a + i0), lo = ldA<T, NT>(a + i1);
    stA<T, NT>(a + i0, cmul2(u.m00, up, u.m01, lo));
    stA<T, NT>(a + i1, cmul2(u.m10, up, u.m11, lo));
}

// ---- permutations ----------------------------------------------------------
// multiControlledMultiQubitNot (QuEST_cpu.c:2761-2810), controlledNot
// (QuEST_cpu.c:2662-2710): swap i <-> i ^ targMask where the controls are 1
template <typename T, bool NT>
__global__ void __launch_bounds__(kBlock) kNot(Amp<T> *__restrict__ a, long long work, BitList fixed,
                                               uint64_t setMask, uint64_t targMask) {
    const long long w = flatBlock() * kBlock + threadIdx.x;
    if (w >= work) return;
    const uint64_t i = insertZeros((uint64_t)w, fixed) | setMask, k = i ^ targMask;
    const Amp<T> x = ldA<T, NT>(a + i), y = ldA<T, NT>(a + k);
    stA<T, NT>(a + i, y);
    stA<T, NT>(a + k, x);
}

// Target bits all below 6: each thread streams its own amplitude and, where
// the controls are 1, takes its partner's (lane l ^ targMask of the same wave)
// through a cross-lane shuffle. It reads and writes every amplitude once,
// contiguously; the pair form touches only the controlled ones but, with a
// control or target among the lowest bits, still moves whole cache lines,
// partly written (a CNOT on qubits 0 and 1 at 32 qubits FP64: 108 GiB in
// 24.0 ms).
template <typename T>
__global__ void __launch_bounds__(kBlock) kNotLane(Amp<T> *__restrict__ a, long long n, uint64_t ctrlMask,
                                                   int targMask) {
    const long long i = flatBlock() * kBlock + threadIdx.x;
    if (i >= n) return;
    const Amp<T> v = ldA<T, true>(a + i);
    const Amp<T> p{__shfl_xor(v.x, targMask, 64), __shfl_xor(v.y, targMask, 64)};
    stA<T, true>(a + i, (((uint64_t)i & ctrlMask) == ctrlMask) ? p : v);
}

// swapGate (QuEST_cpu.c:3905-3950): |..0..1..> <-> |..1..0..>
template <typename T, bool NT>
__global__ void __launch_bounds__(kBlock) kSwap(Amp<T> *__restrict__ a, long long work, BitList fixed,
                                                uint64_t m1, uint64_t m2) {
    const long long w = flatBlock() * kBlock + threadIdx.x;
    if (w >= work) return;
    const uint64_t base = insertZeros((uint64_t)w, fixed);
    const uint64_t i = base | m2, k = base | m1;
    const Amp<T> x = ldA<T, NT>(a + i), y = ldA<T, NT>(a + k);
    stA<T, NT>(a + i, y);
    stA<T, NT>(a + k, x);
}

// swapGate with both qubits below 6 and one among the lowest bits: every
// amplitude streamed, exchanged with lane l ^ (m1 | m2) where the two bits
// differ (the pair form would move whole lines of which it changes half)
template <typename T>
__global__ void __launch_bounds__(kBlock) kSwapLane(Amp<T> *__restrict__ a, long long n, int m1, int m2) {
    const long long i = flatBlock() * kBlock + threadIdx.x;
    if (i >= n) return;
    const Amp<T> v = ldA<T, true>(a + i);
    const Amp<T> p{__shfl_xor(v.x, m1 | m2, 64), __shfl_xor(v.y, m1 | m2, 64)};
    const bool differ = (((int)i & m1) != 0) != (((int)i & m2) != 0);
    stA<T, true>(a + i, differ ? p : v);
}

// ---- diagonal gates --------------------------------------------------------
// phaseShiftByTerm / multiControlledPhaseShift / phase flips: amp *= term on
// the 2^(n-m) amplitudes whose mask bits are all 1
template <typename T, bool NT>
__global__ void __launch_bounds__(kBlock) kPhaseAllOnes(Amp<T> *__restrict__ a, long long work,
                                                        BitList fixed, uint64_t mask, C2<T> t) {
    const long long w = flatBlock() * kBlock + threadIdx.x;
    if (w >= work) return;
    const uint64_t i = insertZeros((uint64_t)w, fixed) | mask;
    const Amp<T> v = ldA<T, NT>(a + i);
    stA<T, NT>(a + i, Amp<T>{t.re * v.x - t.im * v.y, t.re * v.y + t.im * v.x});
}

// The same with every mask bit among the lowest bits: each line then holds
// amplitudes inside and outside the mask, so the direct form moves every line;
// the in-place stream multiplies the matching amplitudes and writes the others
// back unchanged.
template <typename T>
__global__ void __launch_bounds__(kBlock) kPhaseLane(Amp<T> *__restrict__ a, long long n, uint64_t mask, C2<T> t) {
    const long long i = flatBlock() * kBlock + threadIdx.x;
    if (i >= n) return;
    const Amp<T> v = ldA<T, true>(a + i);
    const bool hit = ((uint64_t)i & mask) == mask;
    stA<T, true>(a + i, hit ? Amp<T>{t.re * v.x - t.im * v.y, t.re * v.y + t.im * v.x} : v);
}

// multiRotateZ (QuEST_cpu.c:3299-3339): amp *= cos(θ/2) - i·f·sin(θ/2),
// f = +1 / -1 for even / odd parity of (i & targMask); only where ctrls are 1
template <typename T, bool NT>
__global__ void __launch_bounds__(kBlock) kRotZ(Amp<T> *__restrict__ a, long long work, BitList fixed,
                                                uint64_t setMask, uint64_t targMask, T c, T s) {
    const long long w = flatBlock() * kBlock + threadIdx.x;
    if (w >= work) return;
    const uint64_t i = insertZeros((uint64_t)w, fixed) | setMask;
    const T f = parity64(i & targMask) ? -s : s;
    const Amp<T> v = ldA<T, NT>(a + i);
    stA<T, NT>(a + i, Amp<T>{c * v.x + f * v.y, c * v.y - f * v.x});
}

// ---- Pauli-string rotation in one sweep -------------------------------------
// exp(-iθ/2 P) with P = i^nY X^x Z^z (Y = iXZ): out = cos·ψ - i·sin·(Pψ),
// (Pψ)_j = i^nY (-1)^{|(j^x)&z|} ψ_{j^x}. Replaces QuEST's basis change +
// multiRotateZ + undo (QuEST_common.c:411-448), i.e. 1 sweep instead of 1+2w.
template <typename T, bool NT>
__global__ void __launch_bounds__(kBlock) kPauliRot(Amp<T> *__restrict__ a, long long work,
                                                    BitList fixed, uint64_t setMask, uint64_t xMask,
                                                    uint64_t zMask, C2<T> ph, T c, T s) {
    const long long w = flatBlock() * kBlock + threadIdx.x;
    if (w >= work) return;
    // coefficient on the partner: -i·s·ph (times ±1 from the Z parity)
    const T kre = s * ph.im, kim = -s * ph.re;
    const uint64_t i = insertZeros((uint64_t)w, fixed) | setMask;
    const uint64_t p = i ^ xMask;
    const Amp<T> vi = ldA<T, NT>(a + i), vp = ldA<T, NT>(a + p);
    // out_i uses the sign of (p & z), out_p the sign of (i & z)
    const T si = parity64(p & zMask) ? (T)-1 : (T)1;
    const T sp = parity64(i & zMask) ? (T)-1 : (T)1;
    stA<T, NT>(a + i, Amp<T>{c * vi.x + si * (kre * vp.x - kim * vp.y), c * vi.y + si * (kre * vp.y + kim * vp.x)});
    stA<T, NT>(a + p, Amp<T>{c * vp.x + sp * (kre * vi.x - kim * vi.y), c * vp.y + sp * (kre * vi.y + kim * vi.x)});
}

// The same rotation with every X/Y qubit below 6: each thread streams its own
// amplitude and takes its partner (i ^ xMask) from lane l ^ xMask; with
// controls among the lowest bits (CTRL), uncontrolled amplitudes are written
// back unchanged. Output i uses the sign of parity((i ^ xMask) & zMask), as
// the pair form's out_i does.
template <typename T, bool CTRL>
__global__ void __launch_bounds__(kBlock) kPauliRotLane(Amp<T> *__restrict__ a, long long n, uint64_t ctrlMask,
                                                        int xMask, uint64_t zMask, C2<T> ph, T c, T s) {
    const long long i = flatBlock() * kBlock + threadIdx.x;
    if (i >= n) return;
    const T kre = s * ph.im, kim = -s * ph.re;
    const Amp<T> v = ldA<T, true>(a + i);
    const Amp<T> p{__shfl_xor(v.x, xMask, 64), __shfl_xor(v.y, xMask, 64)};
    const T si = parity64(((uint64_t)i ^ (uint64_t)xMask) & zMask) ? (T)-1 : (T)1;
    Amp<T> o{c * v.x + si * (kre * p.x - kim * p.y), c * v.y + si * (kre * p.y + kim * p.x)};
    if (CTRL && ((uint64_t)i & ctrlMask) != ctrlMask) o = v;
    stA<T, true>(a + i, o);
}

// ---- dense 2^k x 2^k on k targets: LDS-tiled --------------------------------
// QuEST_cpu.c:1895-2007 semantics (targs[0] = least significant bit of the
// matrix row/column index). A workgroup owns a tile of 2^(k+L) amplitudes:
// all target combinations times the L lowest non-target bits, with the tile's
// bit set S sorted so that loading tile index t from base | deposit(t, S)
// walks memory in increasing order (contiguous runs for the low bits).
// Matrix and tile live in LDS; no per-call global scratch (the reference
// allocated 2^k·grid·24 B of global memory per call, QuEST_gpu.cu:1065-1073).
struct TileSpec {
    int k;            // targets
    int tileBits;     // k + L
    int sPos[32];     // global bit position of each tile bit (ascending)
    int tLocal[16];   // tile-local position of targs[j]
    int gLocal[16];   // tile-local position of group bit j
    BitList outer;    // bits not in the tile and not controls (for block base), ascending (as fixed list)
    uint64_t ctrlMask;
};

template <typename T, int MAXK>
__global__ void __launch_bounds__(kBlock) kMatrixNTile(Amp<T> *__restrict__ a, long long numTiles,
                                                       TileSpec ts, const C2<T> *__restrict__ gU,
                                                       int matInLds) {
    extern __shared__ unsigned char smem[];
    const int k = ts.k, dim = 1 << k;
    const int tileAmps = 1 << ts.tileBits;
    const int L = ts.tileBits - k, G = 1 << L;
    Amp<T> *tile = reinterpret_cast<Amp<T> *>(smem);
    C2<T> *mat = reinterpret_cast<C2<T> *>(smem + sizeof(Amp<T>) * tileAmps);
    int *cOff = reinterpret_cast<int *>(smem + sizeof(Amp<T>) * tileAmps +
                                        (matInLds ? sizeof(C2<T>) * dim * dim : 0));

    if (matInLds)
        for (int e = threadIdx.x; e < dim * dim; e += kBlock) mat[e] = gU[e];
    for (int c = threadIdx.x; c < dim; c += kBlock) {
        int off = 0;
        for (int j = 0; j < k; ++j)
            if ((c >> j) & 1) off |= 1 << ts.tLocal[j];
        cOff[c] = off;
    }
    const C2<T> *U = matInLds ? mat : gU;

    for (long long tileId = blockIdx.x; tileId < numTiles; tileId += gridDim.x) {
        const uint64_t base = insertZeros((uint64_t)tileId, ts.outer) | ts.ctrlMask;
        __syncthreads();
        for (int t = threadIdx.x; t < tileAmps; t += kBlock) {
            uint64_t idx = base;
            for (int b = 0; b < ts.tileBits; ++b)
                if ((t >> b) & 1) idx |= 1ull << ts.sPos[b];
            tile[t] = a[idx];
        }
        __syncthreads();
        // each thread computes outputs o = tid + m*kBlock; r = o >> L, g = o & (G-1)
        constexpr int kMaxOut = 4;  // tiles hold at most 4 * kBlock amplitudes
        Amp<T> res[kMaxOut];
        int nOut = 0;
        for (int o = threadIdx.x; o < tileAmps && nOut < kMaxOut; o += kBlock, ++nOut) {
            int r = o >> L, g = o & (G - 1);
            int gOff = 0;
            for (int j = 0; j < L; ++j)
                if ((g >> j) & 1) gOff |= 1 << ts.gLocal[j];
            T re = 0, im = 0;
            const C2<T> *row = U + (size_t)r * dim;
            for (int c = 0; c < dim; ++c) {
                C2<T> m = row[c];
                Amp<T> v = tile[gOff | cOff[c]];
                re += m.re * v.x - m.im * v.y;
                im += m.re * v.y + m.im * v.x;
            }
            res[nOut] = Amp<T>{re, im};
        }
        __syncthreads();
        nOut = 0;
        for (int o = threadIdx.x; o < tileAmps && nOut < kMaxOut; o += kBlock, ++nOut) {
            int r = o >> L, g = o & (G - 1);
            int gOff = 0;
            for (int j = 0; j < L; ++j)
                if ((g >> j) & 1) gOff |= 1 << ts.gLocal[j];
            tile[gOff | cOff[r]] = res[nOut];
        }
        __syncthreads();
        for (int t = threadIdx.x; t < tileAmps; t += kBlock) {
            uint64_t idx = base;
            for (int b = 0; b < ts.tileBits; ++b)
                if ((t >> b) & 1) idx |= 1ull << ts.sPos[b];
            a[idx] = tile[t];
        }
    }
}

BitList makeBits(uint64_t mask) {
    BitList b;
    b.n = 0;
    for (int p = 0; p < 64; ++p)
        if ((mask >> p) & 1ull) b.pos[b.n++] = p;
    return b;
}

template <typename T> C2<T> c2(const Cplx &z) { return C2<T>{(T)z.re, (T)z.im}; }

int lowestBit(uint64_t m) { return m ? __builtin_ctzll(m) : 63; }

// Every control among the lowest bits (below the nontemporal threshold): each
// cache line then holds every control combination, so a pair form that
// touches only the controlled amplitudes still moves every line. With a
// control above it, the pair form moves only the lines that control selects
// and beats an in-place stream of the whole register.
bool controlsAllLow(uint64_t ctrlMask, int prec) {
    return ctrlMask && !useNT(63 - __builtin_clzll(ctrlMask), prec);
}

#define QMX_LAUNCH_NT(nt, work, KERN, ...)                                                            \
    do {                                                                                             \
        if (nt)                                                                                      \
            hipLaunchKernelGGL((KERN<T, true>), dim3(gridFor(work, 1)), dim3(kBlock), 0, hstream(),   \
                               __VA_ARGS__);                                                         \
        else                                                                                         \
            hipLaunchKernelGGL((KERN<T, false>), dim3(gridFor(work, 1)), dim3(kBlock), 0, hstream(),  \
                               __VA_ARGS__);                                                         \
    } while (0)

}  // namespace

void applyMatrix2(QuregDev *q, int target, uint64_t ctrlMask, uint64_t flipMask, const Cplx u[4]) {
    const uint64_t tbit = 1ull << target;
    const int c = __builtin_popcountll(ctrlMask);
    const long long work = q->numAmps >> (1 + c);
    const uint64_t setMask = ctrlMask & ~flipMask;
    BitList fixed = makeBits(ctrlMask | tbit);
    const bool nt = useNT(lowestBit(ctrlMask | tbit), q->prec);
    QMX_DISPATCH(q->prec, T, {
        M2<T> m{c2<T>(u[0]), c2<T>(u[1]), c2<T>(u[2]), c2<T>(u[3])};
        Amp<T> *a = (Amp<T> *)q->amps;
        const dim3 g(gridFor(work, 1));
        if (!ctrlMask && target < 6 && q->numQubits >= 8) {
            hipLaunchKernelGGL((kMatrix2Lane<T, false>), gridFor(q->numAmps, 1), dim3(kBlock), 0, hstream(), a,
                               q->numAmps, target, 0ull, 0ull, m);
        } else if (controlsAllLow(ctrlMask, q->prec) && target < 6 && q->numQubits >= 8) {
            // all controls among the lowest bits: the pair form would move
            // every line while changing only part of each
            hipLaunchKernelGGL((kMatrix2Lane<T, true>), gridFor(q->numAmps, 1), dim3(kBlock), 0, hstream(), a,
                               q->numAmps, target, ctrlMask, setMask, m);
        } else if (!ctrlMask && target < 11 && q->numQubits >= 11) {
            hipLaunchKernelGGL(kMatrix2Span<T>, gridOf(q->numAmps / kSpan), dim3(kBlock), 0, hstream(), a, target, m);
        } else if (ctrlMask) {
            if (nt) hipLaunchKernelGGL((kMatrix2<T, true, true>), g, dim3(kBlock), 0, hstream(), a, work, target, fixed, setMask, m);
            else hipLaunchKernelGGL((kMatrix2<T, true, false>), g, dim3(kBlock), 0, hstream(), a, work, target, fixed, setMask, m);
        } else {
            if (nt) hipLaunchKernelGGL((kMatrix2<T, false, true>), g, dim3(kBlock), 0, hstream(), a, work, target, fixed, setMask, m);
            else hipLaunchKernelGGL((kMatrix2<T, false, false>), g, dim3(kBlock), 0, hstream(), a, work, target, fixed, setMask, m);
        }
    });
    checkLaunch("applyMatrix2");
}

void applyNot(QuregDev *q, uint64_t ctrlMask, uint64_t targMask) {
    const uint64_t lowT = targMask & (~targMask + 1ull);
    const int c = __builtin_popcountll(ctrlMask);
    const long long work = q->numAmps >> (1 + c);
    BitList fixed = makeBits(ctrlMask | lowT);
    const bool nt = useNT(lowestBit(ctrlMask | targMask), q->prec);
    QMX_DISPATCH(q->prec, T, {
        // the lane form moves every amplitude: it wins where the pair form also
        // moves every line (no controls, or all of them among the lowest bits)
        // and loses where a high control halves the traffic
        if (targMask < 64 && q->numQubits >= 8 && (!ctrlMask || controlsAllLow(ctrlMask, q->prec)))
            hipLaunchKernelGGL(kNotLane<T>, gridFor(q->numAmps, 1), dim3(kBlock), 0, hstream(), (Amp<T> *)q->amps,
                               q->numAmps, ctrlMask, (int)targMask);
        else
            QMX_LAUNCH_NT(nt, work, kNot, (Amp<T> *)q->amps, work, fixed, ctrlMask, targMask);
    });
    checkLaunch("applyNot");
}

void applySwap(QuregDev *q, int q1, int q2) {
    const uint64_t m1 = 1ull << q1, m2 = 1ull << q2;
    const long long work = q->numAmps >> 2;
    BitList fixed = makeBits(m1 | m2);
    const bool nt = useNT(lowestBit(m1 | m2), q->prec);
    QMX_DISPATCH(q->prec, T, {
        if ((m1 | m2) < 64 && !nt && q->numQubits >= 8)
            hipLaunchKernelGGL(kSwapLane<T>, gridFor(q->numAmps, 1), dim3(kBlock), 0, hstream(), (Amp<T> *)q->amps,
                               q->numAmps, (int)m1, (int)m2);
        else
            QMX_LAUNCH_NT(nt, work, kSwap, (Amp<T> *)q->amps, work, fixed, m1, m2);
    });
    checkLaunch("applySwap");
}

void applyPhaseOnAllOnes(QuregDev *q, uint64_t mask, Cplx term) {
    const int m = __builtin_popcountll(mask);
    const long long work = q->numAmps >> m;
    BitList fixed = makeBits(mask);
    const bool nt = useNT(lowestBit(mask), q->prec);
    QMX_DISPATCH(q->prec, T, {
        if (controlsAllLow(mask, q->prec) && q->numQubits >= 8)
            hipLaunchKernelGGL(kPhaseLane<T>, gridFor(q->numAmps, 1), dim3(kBlock), 0, hstream(), (Amp<T> *)q->amps,
                               q->numAmps, mask, c2<T>(term));
        else
            QMX_LAUNCH_NT(nt, work, kPhaseAllOnes, (Amp<T> *)q->amps, work, fixed, mask, c2<T>(term));
    });
    checkLaunch("applyPhaseOnAllOnes");
}

void applyMultiRotateZ(QuregDev *q, uint64_t ctrlMask, uint64_t targMask, double angle) {
    const int c = __builtin_popcountll(ctrlMask);
    const long long work = q->numAmps >> c;
    BitList fixed = makeBits(ctrlMask);
    const double cs = cos(angle / 2.0), sn = sin(angle / 2.0);
    const bool nt = useNT(lowestBit(ctrlMask), q->prec);
    QMX_DISPATCH(q->prec, T, {
        QMX_LAUNCH_NT(nt, work, kRotZ, (Amp<T> *)q->amps, work, fixed, ctrlMask, targMask, (T)cs, (T)sn);
    });
    checkLaunch("applyMultiRotateZ");
}

void applyPauliRotation(QuregDev *q, uint64_t ctrlMask, uint64_t xMask, uint64_t zMask, int numY,
                        double angle) {
    if (xMask == 0) {
        // diagonal: Z-string rotation (numY == 0 here since Y sets x)
        applyMultiRotateZ(q, ctrlMask, zMask, angle);
        return;
    }
    const uint64_t lowX = xMask & (~xMask + 1ull);
    const int c = __builtin_popcountll(ctrlMask);
    const long long work = q->numAmps >> (1 + c);
    BitList fixed = makeBits(ctrlMask | lowX);
    const double cs = cos(angle / 2.0), sn = sin(angle / 2.0);
    // i^numY
    static const double phr[4] = {1, 0, -1, 0}, phi[4] = {0, 1, 0, -1};
    Cplx ph{phr[numY & 3], phi[numY & 3]};
    const bool nt = useNT(lowestBit(ctrlMask | xMask), q->prec);
    // X/Y qubits below 6 and no control, or all of them among the lowest
    // bits: the in-place stream (the pair form would move every line anyway)
    const bool lane = xMask < 64 && q->numQubits >= 8 && (!ctrlMask || controlsAllLow(ctrlMask, q->prec));
    QMX_DISPATCH(q->prec, T, {
        if (lane && ctrlMask)
            hipLaunchKernelGGL((kPauliRotLane<T, true>), gridFor(q->numAmps, 1), dim3(kBlock), 0, hstream(),
                               (Amp<T> *)q->amps, q->numAmps, ctrlMask, (int)xMask, zMask, c2<T>(ph), (T)cs, (T)sn);
        else if (lane)
            hipLaunchKernelGGL((kPauliRotLane<T, false>), gridFor(q->numAmps, 1), dim3(kBlock), 0, hstream(),
                               (Amp<T> *)q->amps, q->numAmps, 0ull, (int)xMask, zMask, c2<T>(ph), (T)cs, (T)sn);
        else
            QMX_LAUNCH_NT(nt, work, kPauliRot, (Amp<T> *)q->amps, work, fixed, ctrlMask, xMask, zMask, c2<T>(ph),
                          (T)cs, (T)sn);
    });
    checkLaunch("applyPauliRotation");
}

void applyMatrixN(QuregDev *q, uint64_t ctrlMask, const int *targs, int k, const Cplx *u) {
    const int n = q->numQubits;
    if (q->prec == FP32 && k == 5 && ctrlMask == 0 && n >= 9) {
        // uncontrolled 5-qubit matrices on FP32 registers run on the MFMA kernel
        fused5ApplyOnce(q, targs, u);
        return;
    }
    if (q->prec == FP64 && k == 5 && ctrlMask == 0 && n >= 11) {
        // and on FP64 registers large enough for whole tiles, the FP64 MFMA tiles
        fused5ApplyOnce64(q, targs, u);
        return;
    }
    if (ctrlMask == 0 && denseBitsSupported(k) && n >= k) {
        // uncontrolled 2..5-target matrices: one group per thread in registers
        applyDense(q, targs, k, u);
        return;
    }
    const int c = __builtin_popcountll(ctrlMask);
    uint64_t tmask = 0;
    for (int j = 0; j < k; ++j) tmask |= 1ull << targs[j];
    // choose L low non-target, non-control bits so a tile has >= 256 amps (>= 1 per lane)
    int free = n - k - c;
    int L = std::max(0, std::min(free, std::max(0, 8 - k)));
    uint64_t gmask = 0;
    for (int p = 0, got = 0; p < n && got < L; ++p) {
        uint64_t b = 1ull << p;
        if (!(tmask & b) && !(ctrlMask & b)) { gmask |= b; ++got; }
    }
    TileSpec ts;
    ts.k = k;
    ts.tileBits = k + L;
    uint64_t smask = tmask | gmask;
    int nb = 0;
    for (int p = 0; p < 64; ++p)
        if ((smask >> p) & 1ull) ts.sPos[nb++] = p;
    auto localOf = [&](int pos) {
        for (int b = 0; b < nb; ++b)
            if (ts.sPos[b] == pos) return b;
        return -1;
    };
    for (int j = 0; j < k; ++j) ts.tLocal[j] = localOf(targs[j]);
    for (int j = 0, p = 0; p < 64 && j < L; ++p)
        if ((gmask >> p) & 1ull) ts.gLocal[j++] = localOf(p);
    ts.outer = makeBits(smask | ctrlMask);
    ts.ctrlMask = ctrlMask;
    const long long numTiles = q->numAmps >> (k + L + c);
    const int dim = 1 << k;
    const int tileAmps = 1 << (k + L);
    if (k > 10) {
        std::fprintf(stderr, "libquest_mi355x: applyMatrixN supports at most 10 targets (got %d)\n", k);
        std::abort();
    }

    QMX_DISPATCH(q->prec, T, {
        size_t matBytes = sizeof(C2<T>) * (size_t)dim * dim;
        std::vector<C2<T>> hm((size_t)dim * dim);
        for (size_t e = 0; e < hm.size(); ++e) hm[e] = c2<T>(u[e]);
        hipStream_t st = hstream();
        const void *dMat;
        if (matBytes <= kStageSlotBytes) {
            const Staged sg = stageBegin(matBytes);
            std::memcpy(sg.host, hm.data(), matBytes);
            stageCommit(sg, matBytes, st);
            dMat = sg.dev;
        } else {
            dMat = stageLarge(hm.data(), matBytes, st);
        }
        size_t tileBytes = sizeof(Amp<T>) * tileAmps;
        int matInLds = (tileBytes + matBytes + dim * sizeof(int) <= 64 * 1024) ? 1 : 0;
        size_t shm = tileBytes + (matInLds ? matBytes : 0) + dim * sizeof(int);
        long long grid = std::min<long long>(numTiles, 65536);
        hipLaunchKernelGGL((kMatrixNTile<T, 16>), dim3((unsigned)grid), dim3(kBlock), shm, st,
                           (Amp<T> *)q->amps, numTiles, ts, (const C2<T> *)dMat, matInLds);
    });
    checkLaunch("applyMatrixN");
}

// Uniformly controlled 2x2: u holds 4 entries (row-major) per control value.
void applyUniform2(QuregDev *q, const int *controls, int numControls, int target, const Cplx *u) {
    if (numControls > kMaxUniformControls) {
        std::fprintf(stderr, "libquest_mi355x: applyUniform2 supports at most %d controls (got %d)\n",
                     kMaxUniformControls, numControls);
        std::abort();
    }
    CtrlBits cb{};
    cb.n = numControls;
    for (int b = 0; b < numControls; ++b) cb.pos[b] = controls[b];
    const long long work = q->numAmps >> 1;
    const bool nt = useNT(target, q->prec);
    const size_t count = (size_t)1 << numControls;
    QMX_DISPATCH(q->prec, T, {
        const size_t bytes = count * sizeof(M2<T>);
        hipStream_t st = hstream();
        const void *dMat;
        if (bytes <= kStageSlotBytes) {
            const Staged sg = stageBegin(bytes);
            M2<T> *h = static_cast<M2<T> *>(sg.host);
            for (size_t k = 0; k < count; ++k)
                h[k] = M2<T>{c2<T>(u[4 * k]), c2<T>(u[4 * k + 1]), c2<T>(u[4 * k + 2]), c2<T>(u[4 * k + 3])};
            stageCommit(sg, bytes, st);
            dMat = sg.dev;
        } else {
            std::vector<M2<T>> h(count);
            for (size_t k = 0; k < count; ++k)
                h[k] = M2<T>{c2<T>(u[4 * k]), c2<T>(u[4 * k + 1]), c2<T>(u[4 * k + 2]), c2<T>(u[4 * k + 3])};
            dMat = stageLarge(h.data(), bytes, st);
        }
        if (nt)
            hipLaunchKernelGGL((kUniform2<T, true>), gridFor(work, 1), dim3(kBlock), 0, st, (Amp<T> *)q->amps, work,
                               target, cb, (const M2<T> *)dMat);
        else
            hipLaunchKernelGGL((kUniform2<T, false>), gridFor(work, 1), dim3(kBlock), 0, st, (Amp<T> *)q->amps,
                               work, target, cb, (const M2<T> *)dMat);
    });
    checkLaunch("applyUniform2");
}

}  // namespace qmx
