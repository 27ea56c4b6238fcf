// Runs of Pauli exponentials in one sweep (Trotter circuits).
//
// A Trotter repetition is an ordered list of rotations exp(-i theta/2 P_j),
// P = i^nY X^x Z^z. Consecutive rotations whose X/Y supports together span at
// most K qubits form a block: every amplitude pair a rotation of the block
// touches lies inside one coset of those K bits, so a workgroup can load a
// tile — the block's X/Y bits plus the lowest other bits, 2^tileBits
// amplitudes, contiguous runs in memory — into LDS, apply all of the block's
// rotations there in order, and write it back: one HBM read + write per block
// instead of per rotation. Z parities split as parity(base & z) (one per tile
// and rotation) xor parity(local & z_local), and a pair's partner parity is
// the pair's own xor parity(lx & z_local). The per-rotation arithmetic is the
// one-sweep kernels' (kPauliRot, kRotZ), term for term. Inside a tile the
// rotations run in passes: consecutive rotations whose X/Y masks coincide
// (e.g. XX and YY on one bond), plus any diagonal rotations between them, are
// applied to each thread's amplitude pairs in registers in one LDS round trip.
// Full tiles use the coset form instead (kPauliCoset): one LDS round trip per
// group of rotations whose X/Y masks span at most 4 (FP32) / 3 (FP64)
// dimensions, each thread holding one coset of that span in registers.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "qmx_device.h"

namespace qmx {

namespace {

constexpr int kMaxTileBits = 12;

struct PauliExpDev {
    uint64_t z;          // global z mask (parity of the tile base)
    unsigned lz;         // tile-local z mask
    unsigned pxz;        // parity(lx & lz) (or parity(half & lz) in a diagonal-only pass): partner vs own sign
    unsigned kpar;       // bit k: parity(D_k & lz), D_k = offset of a thread's k-th pair from its first
    int kind;            // 0 diagonal, 1 partner factor real, 2 partner factor imaginary; +3: tangent form
    double c;            // cos(theta/2)
    double kap;          // sin(theta/2) (diagonal), else the non-zero part of -i sin(theta/2) i^nY;
                         // divided by cos(theta/2) in the tangent form
};

// a pass: consecutive rotations of a block whose X/Y parts share one tile-local
// mask lx (diagonal rotations join any pass); lx == 0: diagonal rotations only
struct PassDev {
    unsigned lx;
    int first, count;
    int pad;
    unsigned D[8];       // tile-index offset of a thread's k-th pair from its first
};

constexpr int kMaxPer = 1 << (kMaxTileBits - 8);   // amplitudes per thread: tile / 256

struct PauliTile {
    int tb;                   // tile bits
    int pos[kMaxTileBits];    // global bit of tile bit j, ascending
    BitList fixed;            // the same positions (tile base = insert zeros there)
    uint64_t hi[kMaxPer];     // global offset of tile index 256 k (tile bits 8 and up)
};

// v * (-1)^(bit 0 of p), by flipping the sign bit
__device__ __forceinline__ float signFlip(float v, unsigned p) { return __uint_as_float(__float_as_uint(v) ^ (p << 31)); }
__device__ __forceinline__ double signFlip(double v, unsigned p) {
    return __longlong_as_double(__double_as_longlong(v) ^ ((unsigned long long)(p & 1u) << 63));
}

// PPT amplitude pairs per thread and pass (tiles of 2^(9 + log2 PPT) amplitudes):
// a thread keeps its pairs in registers while every rotation of the pass is
// applied, so a rotation's constants are read once per thread, from scalar
// loads, with the tile base's parity folded into its signs. PREFETCH: a
// persistent grid (4 workgroups per CU, the occupancy its 119 VGPRs allow)
// walks tiles w = blockIdx.x + k gridDim.x and loads the next tile into
// registers while the current one is in LDS: 14 % faster for FP32 at 24
// qubits, measured while this form served full tiles. FP64 runs one workgroup
// per tile (its prefetch buffer was placed in scratch memory, and the variant
// without it is faster). Full tiles now take the coset form, so this kernel
// serves registers smaller than a tile, which are a single tile.
template <typename T, int PPT, bool PREFETCH>
__global__ void __launch_bounds__(kBlock) kPauliBlock(Amp<T> *__restrict__ a, unsigned numTiles, PauliTile tile,
                                                      const PassDev *__restrict__ passes, int numPasses,
                                                      const PauliExpDev *__restrict__ ops, double scale) {
    constexpr int kPer = 2 * PPT;  // amplitudes a thread loads and stores per tile
    extern __shared__ unsigned char smem[];
    Amp<T> *st = reinterpret_cast<Amp<T> *>(smem);
    const int dim = 1 << tile.tb;
    const int half = dim >> 1;
    // tile index t = threadIdx.x + kBlock k: bits 0-7 come from the thread
    // (scattered to dt), bits 8 and up from k (gk[k], the same in every thread)
    const int lowBits = tile.tb < 8 ? tile.tb : 8;
    uint64_t dt = 0;
    for (int j = 0; j < lowBits; ++j)
        if ((threadIdx.x >> j) & 1) dt |= 1ull << tile.pos[j];
    const bool active = (int)threadIdx.x < dim;
    Amp<T> next[PREFETCH ? kPer : 1];
    unsigned w = blockIdx.x;
    if (PREFETCH) {
        Amp<T> *ns = a + (insertZeros((uint64_t)w, tile.fixed) | dt);
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if (active && (k << 8) < dim) next[k] = ns[tile.hi[k]];
    }
    for (; w < numTiles; w += PREFETCH ? gridDim.x : numTiles) {
    const uint64_t base = insertZeros((uint64_t)w, tile.fixed);
    Amp<T> *src = a + (base | dt);
#pragma unroll
    for (int k = 0; k < kPer; ++k)
        if (active && (k << 8) < dim) st[threadIdx.x + (k << 8)] = PREFETCH ? next[k] : src[tile.hi[k]];
    __syncthreads();
    if (PREFETCH && w + gridDim.x < numTiles) {
        Amp<T> *ns = a + (insertZeros((uint64_t)(w + gridDim.x), tile.fixed) | dt);
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if (active && (k << 8) < dim) next[k] = ns[tile.hi[k]];
    }
    for (int ps = 0; ps < numPasses; ++ps) {
        const PassDev pass = passes[ps];
        // pairs (i, i ^ lx) of this thread; lx == 0: amplitudes i and i + half.
        // Pair k's index is the first pair's with the uniform offset D[k] or-ed in.
        const unsigned lx = pass.lx;
        const unsigned pm = lx ? lx : (unsigned)half;
        const unsigned low = lx & (~lx + 1u);
        const int lb = lx ? __ffs(lx) - 1 : 0;
        const unsigned tx = threadIdx.x;
        const unsigned idx0 = lx ? (((tx >> lb) << (lb + 1)) | (tx & (low - 1u))) : tx;
        unsigned idx[PPT];
        bool ok[PPT];
        Amp<T> vi[PPT], vp[PPT];
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            ok[k] = (int)(tx + kBlock * k) < half;
            idx[k] = idx0 | pass.D[k];
            if (ok[k]) {
                vi[k] = st[idx[k]];
                vp[k] = st[idx[k] ^ pm];
            }
        }
        // Signs: a thread's pairs differ from its first pair by tile bits
        // that are the same in every thread (D_k), so a rotation's sign for
        // pair k is the first pair's parity xor a bit the host precomputed
        // (kpar); the partner's is that xor pxz. The partner factor
        // -i sin(theta/2) i^nY is real or imaginary, never both.
        for (int o = 0; o < pass.count; ++o) {
            const PauliExpDev e = ops[pass.first + o];
            const unsigned flip = (unsigned)__popcll(base & e.z);
            const T c = (T)e.c;
            const T kap = signFlip((T)e.kap, flip);
            const T a0 = signFlip(kap, (unsigned)__popc(idx0 & e.lz));
            const T b0 = signFlip(a0, e.pxz);
            switch (e.kind) {
            case 0:
#pragma unroll
                for (int k = 0; k < PPT; ++k) {
                    const T A = signFlip(a0, e.kpar >> k), B = signFlip(b0, e.kpar >> k);
                    const Amp<T> x = vi[k], y = vp[k];
                    vi[k] = Amp<T>{c * x.x + A * x.y, c * x.y - A * x.x};
                    vp[k] = Amp<T>{c * y.x + B * y.y, c * y.y - B * y.x};
                }
                break;
            case 1:
#pragma unroll
                for (int k = 0; k < PPT; ++k) {
                    const T A = signFlip(a0, e.kpar >> k), B = signFlip(b0, e.kpar >> k);
                    const Amp<T> x = vi[k], y = vp[k];
                    vi[k] = Amp<T>{c * x.x + B * y.x, c * x.y + B * y.y};
                    vp[k] = Amp<T>{c * y.x + A * x.x, c * y.y + A * x.y};
                }
                break;
            case 2:
#pragma unroll
                for (int k = 0; k < PPT; ++k) {
                    const T A = signFlip(a0, e.kpar >> k), B = signFlip(b0, e.kpar >> k);
                    const Amp<T> x = vi[k], y = vp[k];
                    vi[k] = Amp<T>{c * x.x - B * y.y, c * x.y + B * y.x};
                    vp[k] = Amp<T>{c * y.x - A * x.y, c * y.y + A * x.x};
                }
                break;
            // tangent form: the rotation divided by cos(theta/2), whose product
            // over the block multiplies the tile once, at the store
            case 3:
#pragma unroll
                for (int k = 0; k < PPT; ++k) {
                    const T A = signFlip(a0, e.kpar >> k), B = signFlip(b0, e.kpar >> k);
                    const Amp<T> x = vi[k], y = vp[k];
                    vi[k] = Amp<T>{x.x + A * x.y, x.y - A * x.x};
                    vp[k] = Amp<T>{y.x + B * y.y, y.y - B * y.x};
                }
                break;
            case 4:
#pragma unroll
                for (int k = 0; k < PPT; ++k) {
                    const T A = signFlip(a0, e.kpar >> k), B = signFlip(b0, e.kpar >> k);
                    const Amp<T> x = vi[k], y = vp[k];
                    vi[k] = Amp<T>{x.x + B * y.x, x.y + B * y.y};
                    vp[k] = Amp<T>{y.x + A * x.x, y.y + A * x.y};
                }
                break;
            default:
#pragma unroll
                for (int k = 0; k < PPT; ++k) {
                    const T A = signFlip(a0, e.kpar >> k), B = signFlip(b0, e.kpar >> k);
                    const Amp<T> x = vi[k], y = vp[k];
                    vi[k] = Amp<T>{x.x - B * y.y, x.y + B * y.x};
                    vp[k] = Amp<T>{y.x - A * x.y, y.y + A * x.x};
                }
                break;
            }
        }
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            if (ok[k]) {
                st[idx[k]] = vi[k];
                st[idx[k] ^ pm] = vp[k];
            }
        }
        __syncthreads();
    }
    const T sc = (T)scale;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (active && (k << 8) < dim) {
            const Amp<T> v = st[threadIdx.x + (k << 8)];
            src[tile.hi[k]] = Amp<T>{sc * v.x, sc * v.y};
        }
    }
    if (PREFETCH) __syncthreads();  // the next tile overwrites the LDS copy
    }
}

// ---- coset form: a thread holds a coset of the exchange masks' span ----------
// Full tiles only (2^(8 + R) amplitudes, R = 4 for FP32, 3 for FP64). A block's
// rotations split into groups of consecutive rotations whose tile-local X/Y
// masks span at most R dimensions over GF(2). For a group, the host puts a
// basis b_0..b_{R-1} of a span containing them in reduced echelon form (pivot
// bit piv_j set in b_j only) and thread t takes the coset r + span, where r =
// t with zeros inserted at the pivots: element v (bit j of v selects b_j) is
// tile index r ^ C[v]. Every pair (i, i ^ lx) of every rotation in the group is
// then a pair (v, v ^ m) of one thread's registers, m = lx's coordinates, so
// the group costs one LDS round trip where the pass form costs one per
// distinct mask. Signs: parity(index & lz) = parity(r & lz) xor parity(v & zv)
// with zv_j = parity(b_j & lz).
struct CosetExpDev {
    uint64_t z;          // global z mask (parity of the tile base)
    unsigned lz;         // tile-local z mask
    unsigned m;          // the X/Y mask's coordinates in the group's basis (0: diagonal)
    unsigned zv;         // bit j: parity(b_j & lz)
    int kind;            // as PauliExpDev
    double c, kap;
};

constexpr int kMaxCosetBits = 4;

struct CosetGroupDev {
    int first, count;             // the group's rotations (block-relative)
    int piv[kMaxCosetBits];       // pivot tile bits, ascending
    unsigned C[1 << kMaxCosetBits];  // tile offset of coset element v
};

template <typename T> constexpr int cosetBits() { return sizeof(T) == 8 ? 3 : 4; }

// A thread's coset, real and imaginary parts as vector values: they stay SSA
// values through the exchange-mask switch below (an array there is kept in
// scratch memory once the compiler merges the cases' stores).
template <typename T, int E> struct CosetRegs {
    typedef T V __attribute__((ext_vector_type(E)));
    V re, im;
};

// out_u = c x_u + kap s(u ^ M) F x_{u ^ M} for every u, F = 1 (KIND 1, 4) or i
// (KIND 2, 5), s(v) = (-1)^parity(v & zv) (the thread's own parity is in kap);
// KIND 4, 5: the tangent form (c = 1)
template <typename T, int E, int M, int KIND>
__device__ __forceinline__ void cosetPairs(CosetRegs<T, E> &x, T c, T kap, unsigned zv) {
    if constexpr (M < E) {
        constexpr int lowM = M & -M;
#pragma unroll
        for (int u = 0; u < E; ++u) {
            if (u & lowM) continue;
            const int p = u ^ M;
            const T Bu = signFlip(kap, (unsigned)__popc((unsigned)p & zv));  // for out_u: the partner's sign
            const T Ap = signFlip(kap, (unsigned)__popc((unsigned)u & zv));  // for out_p
            const T ur = x.re[u], ui = x.im[u], pr = x.re[p], pi = x.im[p];
            if constexpr (KIND == 1) {
                x.re[u] = c * ur + Bu * pr;
                x.im[u] = c * ui + Bu * pi;
                x.re[p] = c * pr + Ap * ur;
                x.im[p] = c * pi + Ap * ui;
            } else if constexpr (KIND == 2) {
                x.re[u] = c * ur - Bu * pi;
                x.im[u] = c * ui + Bu * pr;
                x.re[p] = c * pr - Ap * ui;
                x.im[p] = c * pi + Ap * ur;
            } else if constexpr (KIND == 4) {
                x.re[u] = ur + Bu * pr;
                x.im[u] = ui + Bu * pi;
                x.re[p] = pr + Ap * ur;
                x.im[p] = pi + Ap * ui;
            } else {
                x.re[u] = ur - Bu * pi;
                x.im[u] = ui + Bu * pr;
                x.re[p] = pr - Ap * ui;
                x.im[p] = pi + Ap * ur;
            }
        }
    }
}

template <typename T, int E, int KIND>
__device__ __forceinline__ void cosetExchange(CosetRegs<T, E> &x, unsigned m, T c, T kap, unsigned zv) {
    switch (m) {
    case 1: cosetPairs<T, E, 1, KIND>(x, c, kap, zv); break;
    case 2: cosetPairs<T, E, 2, KIND>(x, c, kap, zv); break;
    case 3: cosetPairs<T, E, 3, KIND>(x, c, kap, zv); break;
    case 4: cosetPairs<T, E, 4, KIND>(x, c, kap, zv); break;
    case 5: cosetPairs<T, E, 5, KIND>(x, c, kap, zv); break;
    case 6: cosetPairs<T, E, 6, KIND>(x, c, kap, zv); break;
    case 7: cosetPairs<T, E, 7, KIND>(x, c, kap, zv); break;
    case 8: cosetPairs<T, E, 8, KIND>(x, c, kap, zv); break;
    case 9: cosetPairs<T, E, 9, KIND>(x, c, kap, zv); break;
    case 10: cosetPairs<T, E, 10, KIND>(x, c, kap, zv); break;
    case 11: cosetPairs<T, E, 11, KIND>(x, c, kap, zv); break;
    case 12: cosetPairs<T, E, 12, KIND>(x, c, kap, zv); break;
    case 13: cosetPairs<T, E, 13, KIND>(x, c, kap, zv); break;
    case 14: cosetPairs<T, E, 14, KIND>(x, c, kap, zv); break;
    default: cosetPairs<T, E, 15, KIND>(x, c, kap, zv); break;
    }
}

// x_u *= c - i kap s(u) (TAN: c = 1)
template <typename T, int E, bool TAN>
__device__ __forceinline__ void cosetDiagonal(CosetRegs<T, E> &x, T c, T kap, unsigned zv) {
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const T A = signFlip(kap, (unsigned)__popc((unsigned)u & zv));
        const T vr = x.re[u], vi = x.im[u];
        x.re[u] = TAN ? vr + A * vi : c * vr + A * vi;
        x.im[u] = TAN ? vi - A * vr : c * vi - A * vr;
    }
}

// Tile staging and the final scaled store as in kPauliBlock, one tile per
// workgroup; the groups replace its passes. A persistent FP32 grid that
// prefetched the next tile into registers measured slower once the groups had
// cut the LDS round trips (24 qubits: 290 vs 284 ms; 27 qubits: 1470 vs 1298
// ms; 29 qubits: 6318 vs 5580 ms for the same solves, identical states).
template <typename T>
__global__ void __launch_bounds__(kBlock) kPauliCoset(Amp<T> *__restrict__ a, PauliTile tile,
                                                      const CosetGroupDev *__restrict__ groups, int numGroups,
                                                      const CosetExpDev *__restrict__ ops, double scale) {
    constexpr int R = cosetBits<T>();
    constexpr int E = 1 << R;   // amplitudes per thread; the tile is kBlock * E
    extern __shared__ unsigned char smem[];
    Amp<T> *st = reinterpret_cast<Amp<T> *>(smem);
    const unsigned tx = threadIdx.x;
    uint64_t dt = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        if ((tx >> j) & 1) dt |= 1ull << tile.pos[j];
    const uint64_t base = insertZeros((uint64_t)flatBlockLinear(), tile.fixed);
    Amp<T> *src = a + (base | dt);
#pragma unroll
    for (int k = 0; k < E; ++k) st[tx + (k << 8)] = src[tile.hi[k]];
    __syncthreads();
    for (int g = 0; g < numGroups; ++g) {
        const CosetGroupDev &grp = groups[g];
        unsigned r = tx;
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const int p = grp.piv[j];
            r = ((r >> p) << (p + 1)) | (r & ((1u << p) - 1u));
        }
        CosetRegs<T, E> x;
#pragma unroll
        for (int v = 0; v < E; ++v) {
            const Amp<T> t = st[r ^ grp.C[v]];
            x.re[v] = t.x;
            x.im[v] = t.y;
        }
        const int first = grp.first, count = grp.count;
        for (int o = 0; o < count; ++o) {
            const CosetExpDev e = ops[first + o];
            const unsigned p0 = (unsigned)__popcll(base & e.z) + (unsigned)__popc(r & e.lz);
            const T kap = signFlip((T)e.kap, p0);
            const T c = (T)e.c;
            switch (e.kind) {
            case 0: cosetDiagonal<T, E, false>(x, c, kap, e.zv); break;
            case 1: cosetExchange<T, E, 1>(x, e.m, c, kap, e.zv); break;
            case 2: cosetExchange<T, E, 2>(x, e.m, c, kap, e.zv); break;
            case 3: cosetDiagonal<T, E, true>(x, c, kap, e.zv); break;
            case 4: cosetExchange<T, E, 4>(x, e.m, c, kap, e.zv); break;
            default: cosetExchange<T, E, 5>(x, e.m, c, kap, e.zv); break;
            }
        }
#pragma unroll
        for (int v = 0; v < E; ++v) st[r ^ grp.C[v]] = Amp<T>{x.re[v], x.im[v]};
        __syncthreads();
    }
    const T sc = (T)scale;
#pragma unroll
    for (int k = 0; k < E; ++k) {
        const Amp<T> v = st[tx + (k << 8)];
        src[tile.hi[k]] = Amp<T>{sc * v.x, sc * v.y};
    }
}

constexpr int kMaxBlockOps = 1 << 16;  // rotations per block (bounds the host's records only)

struct Block {
    PauliTile tile;
    int firstPass, numPasses;   // coset form: its groups
    int firstOp, numOps;
    int single;        // >= 0: one rotation too wide for a tile, applied alone (index into seq)
    bool coset;        // coset form (full tiles) or pass form
    double scale;      // product of cos(theta/2) over the block's tangent-form rotations
};

// Split a block's rotations (tile-local X/Y masks lx, z masks lz) into coset
// groups of at most R exchange dimensions; fills the ops' m and zv.
void buildCosetGroups(const std::vector<unsigned> &lx, std::vector<CosetExpDev> &ops, int firstOp, int tb, int R,
                      std::vector<CosetGroupDev> &groups) {
    const int count = (int)lx.size();
    int start = 0;
    while (start < count) {
        // grow the group while the masks' span stays within R dimensions
        std::vector<unsigned> ech;  // echelon form by highest bit, for the membership test
        int end = start;
        for (; end < count; ++end) {
            unsigned v = lx[end];
            for (unsigned e : ech)
                if ((v ^ e) < v) v ^= e;
            if (v == 0) continue;
            if ((int)ech.size() == R) break;
            ech.push_back(v);
            std::sort(ech.begin(), ech.end(), [](unsigned p, unsigned q) { return p > q; });
        }
        // reduced echelon basis (pivot = highest bit), extended by unit
        // vectors at the highest free tile bits to exactly R vectors
        std::vector<unsigned> b;
        std::vector<int> piv;
        auto add = [&](unsigned v) {
            for (size_t j = 0; j < b.size(); ++j)
                if ((v >> piv[j]) & 1) v ^= b[j];
            const int p = 31 - __builtin_clz(v);
            for (size_t j = 0; j < b.size(); ++j)
                if ((b[j] >> p) & 1) b[j] ^= v;
            b.push_back(v);
            piv.push_back(p);
        };
        for (unsigned e : ech) add(e);
        for (int p = tb - 1; p >= 0 && (int)b.size() < R; --p)
            if (std::find(piv.begin(), piv.end(), p) == piv.end()) add(1u << p);
        // order the basis by pivot, ascending
        std::vector<int> ord(R);
        for (int j = 0; j < R; ++j) ord[j] = j;
        std::sort(ord.begin(), ord.end(), [&](int p, int q) { return piv[p] < piv[q]; });
        CosetGroupDev g{};
        g.first = start;
        g.count = end - start;
        unsigned bs[kMaxCosetBits];
        for (int j = 0; j < R; ++j) {
            g.piv[j] = piv[ord[j]];
            bs[j] = b[ord[j]];
        }
        for (int v = 0; v < (1 << R); ++v) {
            unsigned cv = 0;
            for (int j = 0; j < R; ++j)
                if ((v >> j) & 1) cv ^= bs[j];
            g.C[v] = cv;
        }
        for (int k = start; k < end; ++k) {
            CosetExpDev &d = ops[firstOp + k];
            unsigned m = 0, back = 0;
            for (int j = 0; j < R; ++j)
                if ((lx[k] >> g.piv[j]) & 1) {
                    m |= 1u << j;
                    back ^= bs[j];
                }
            if (back != lx[k]) {
                std::fprintf(stderr, "libquest_mi355x: Pauli coset basis does not span a rotation's mask\n");
                std::abort();
            }
            d.m = m;
            d.zv = 0;
            for (int j = 0; j < R; ++j) d.zv |= ((unsigned)__builtin_popcount(bs[j] & d.lz) & 1u) << j;
        }
        groups.push_back(g);
        start = end;
    }
}

// Reorders one repetition so that blocks carry more rotations. Building a
// block, a later rotation joins it when its X/Y bits still fit and it commutes
// with every rotation it moves past (exp(A) exp(B) = exp(B) exp(A) when
// [A, B] = 0; Pauli strings commute iff |x1 & z2| + |z1 & x2| is even), so
// the product is unchanged. Greedy over a bounded look-ahead; identity strings
// (no-ops, as in multiRotatePauli) are dropped. On the 24-qubit TFIM and
// Heisenberg chains with 32 random strings this cuts sweeps by 19-34 %.
std::vector<PauliExp> packCommuting(const PauliExp *seq, int count, int maxX) {
    constexpr size_t kWindow = 256, kMaxSkipped = 64;
    auto commutes = [](const PauliExp &a, const PauliExp &b) {
        return (__builtin_popcountll((a.x & b.z) ^ (a.z & b.x)) & 1) == 0;
    };
    std::vector<int> rem;
    for (int i = 0; i < count; ++i)
        if ((seq[i].x | seq[i].z) != 0) rem.push_back(i);
    std::vector<PauliExp> out;
    out.reserve(rem.size());
    std::vector<int> skipped, keep;
    while (!rem.empty()) {
        const PauliExp &f = seq[rem[0]];
        out.push_back(f);
        if (__builtin_popcountll(f.x) > maxX) {  // too wide for a tile: applied alone
            rem.erase(rem.begin());
            continue;
        }
        uint64_t xu = f.x;
        skipped.clear();
        keep.clear();
        size_t p = 1;
        for (; p < rem.size() && p < kWindow; ++p) {
            const PauliExp &e = seq[rem[p]];
            bool ok = __builtin_popcountll(xu | e.x) <= maxX;
            for (size_t k = 0; ok && k < skipped.size(); ++k) ok = commutes(e, seq[skipped[k]]);
            if (ok) {
                out.push_back(e);
                xu |= e.x;
            } else {
                skipped.push_back(rem[p]);
                keep.push_back(rem[p]);
                if (skipped.size() > kMaxSkipped) {
                    ++p;
                    break;
                }
            }
        }
        keep.insert(keep.end(), rem.begin() + p, rem.end());
        rem.swap(keep);
    }
    return out;
}

}  // namespace

// tile bits (32 KiB of LDS per tile) and the widest X/Y union of a block: the
// tile keeps the lowest other bits for contiguous runs, 4 in FP64 (256 B) and
// 3 in FP32 (64 B; 9 instead of 8 X/Y qubits per block made 24-qubit FP32
// Trotter solves 3.5 % faster, 74.2 -> 71.6 ms on 32 instances; 2 low bits,
// 10 X/Y qubits, 78 ms)
static int pauliTileBits(int n, int prec) { return std::min(n, prec == FP64 ? 11 : 12); }
static int pauliMaxX(int tb, int prec) { return std::max(1, tb - std::min(tb, prec == FP64 ? 4 : 3)); }

void applyPauliSequence(QuregDev *q, const PauliExp *seqIn, int countIn, int reps) {
    const int n = q->numQubits;
    const int tb = pauliTileBits(n, q->prec);
    const int maxX = pauliMaxX(tb, q->prec);
    const std::vector<PauliExp> packed = packCommuting(seqIn, countIn, maxX);
    const PauliExp *seq = packed.data();
    const int count = (int)packed.size();
    const int R = q->prec == FP64 ? cosetBits<double>() : cosetBits<float>();
    const bool coset = tb == 8 + R;                           // full tiles: the coset form
    const double scaleFloor = q->prec == FP64 ? std::ldexp(1.0, -500) : std::ldexp(1.0, -40);
    std::vector<PauliExpDev> ops;
    std::vector<PassDev> passes;
    std::vector<CosetExpDev> cops;
    std::vector<CosetGroupDev> groups;
    std::vector<Block> blocks;
    int i = 0;
    while (i < count) {
        if ((seq[i].x | seq[i].z) == 0) {  // identity string: a no-op, as in multiRotatePauli
            ++i;
            continue;
        }
        if (__builtin_popcountll(seq[i].x) > maxX) {
            blocks.push_back(Block{PauliTile{}, 0, 0, 0, 0, i, false, 1.0});
            ++i;
            continue;
        }
        uint64_t xu = 0;
        int j = i, nops = 0;
        while (j < count && nops < kMaxBlockOps && __builtin_popcountll(xu | seq[j].x) <= maxX) {
            xu |= seq[j].x;
            nops += (seq[j].x | seq[j].z) != 0;
            ++j;
        }
        // tile bits: the X/Y union plus the lowest other bits
        uint64_t tm = xu;
        for (int b = 0; b < n && __builtin_popcountll(tm) < tb; ++b) tm |= 1ull << b;
        Block blk{};
        blk.single = -1;
        blk.scale = 1.0;
        blk.tile.fixed.n = 0;
        for (int b = 0; b < n; ++b)
            if ((tm >> b) & 1) {
                blk.tile.pos[blk.tile.tb++] = b;
                blk.tile.fixed.pos[blk.tile.fixed.n++] = b;
            }
        for (int k = 0; k < kMaxPer; ++k) {
            uint64_t g = 0;
            for (int j = 8; j < blk.tile.tb; ++j)
                if ((k >> (j - 8)) & 1) g |= 1ull << blk.tile.pos[j];
            blk.tile.hi[k] = g;
        }
        blk.coset = coset;
        blk.firstPass = coset ? (int)groups.size() : (int)passes.size();
        blk.firstOp = coset ? (int)cops.size() : (int)ops.size();
        std::vector<unsigned> lxs;  // coset form: the rotations' tile-local X/Y masks
        int open = -1;  // pass accepting rotations
        for (int k = i; k < j; ++k) {
            const PauliExp &e = seq[k];
            if ((e.x | e.z) == 0) continue;
            PauliExpDev d{};
            unsigned lx = 0;
            d.z = e.z;
            for (int t = 0; t < blk.tile.tb; ++t) {
                if ((e.x >> blk.tile.pos[t]) & 1) lx |= 1u << t;
                if ((e.z >> blk.tile.pos[t]) & 1) d.lz |= 1u << t;
            }
            // partner factor -i s i^nY: nY = 0, 2 -> imaginary -s, +s; nY = 1, 3 -> real s, -s
            const double sn = std::sin(e.angle / 2.0);
            d.c = std::cos(e.angle / 2.0);
            if (e.x == 0) {
                d.kind = 0;
                d.kap = sn;
            } else if (e.ny & 1) {
                d.kind = 1;
                d.kap = (e.ny & 2) ? -sn : sn;
            } else {
                d.kind = 2;
                d.kap = (e.ny & 2) ? sn : -sn;
            }
            // |tan(theta/2)| <= sqrt(3): divide by cos, scale the tile by it once.
            // Amplitudes grow by up to 1/|cos| per such rotation until the final
            // scaled store, so the running product is kept above a floor (amplitudes
            // stay below 2^40 in FP32, 2^500 in FP64, and the scale a normal number);
            // rotations past it use the plain form.
            if (std::fabs(d.c) >= 0.5 && std::fabs(blk.scale * d.c) >= scaleFloor) {
                d.kind += 3;
                d.kap /= d.c;
                blk.scale *= d.c;
            }
            if (coset) {
                cops.push_back(CosetExpDev{d.z, d.lz, 0, 0, d.kind, d.c, d.kap});
                lxs.push_back(lx);
                continue;
            }
            const int rel = (int)ops.size() - blk.firstOp;  // passes index the block's own ops
            if (open < 0) {
                passes.push_back(PassDev{lx, rel, 0, 0, {}});
                open = (int)passes.size() - 1;
            } else if (lx != 0 && passes[open].lx != 0 && passes[open].lx != lx) {
                passes.push_back(PassDev{lx, rel, 0, 0, {}});
                open = (int)passes.size() - 1;
            } else if (lx != 0 && passes[open].lx == 0) {
                passes[open].lx = lx;  // diagonal rotations so far ride along with this exchange
            }
            ops.push_back(d);
            ++passes[open].count;
        }
        if (coset) {
            buildCosetGroups(lxs, cops, blk.firstOp, blk.tile.tb, R, groups);
            blk.numPasses = (int)groups.size() - blk.firstPass;
            blk.numOps = (int)cops.size() - blk.firstOp;
            if (blk.numPasses > 0) blocks.push_back(blk);
            i = j;
            continue;
        }
        blk.numPasses = (int)passes.size() - blk.firstPass;
        blk.numOps = (int)ops.size() - blk.firstOp;
        // signs with each pass's final exchange mask: partner vs own parity, and
        // the parity of the offset D_k of a thread's k-th pair (k < 8) from its
        // first: tile index u + 256 k with a zero inserted at lx's lowest bit
        for (int ps = blk.firstPass; ps < (int)passes.size(); ++ps) {
            const unsigned lx = passes[ps].lx;
            const unsigned half = 1u << (blk.tile.tb - 1);
            const int lb = lx ? __builtin_ctz(lx) : 0;
            for (unsigned k = 0; k < 8; ++k) {
                const unsigned u = 256u * k;
                passes[ps].D[k] = lx ? (((u >> lb) << (lb + 1)) | (u & ((1u << lb) - 1u))) : u;
            }
            for (int o = 0; o < passes[ps].count; ++o) {
                PauliExpDev &d = ops[blk.firstOp + passes[ps].first + o];
                d.pxz = (unsigned)__builtin_popcount((lx ? lx : half) & d.lz) & 1u;
                d.kpar = 0;
                for (unsigned k = 0; k < 8; ++k)
                    d.kpar |= ((unsigned)__builtin_popcount(passes[ps].D[k] & d.lz) & 1u) << k;
            }
        }
        if (blk.numPasses > 0) blocks.push_back(blk);
        i = j;
    }
    if (blocks.empty()) return;
    std::vector<void *> owned;
    auto upload = [&](const void *h, size_t bytes) -> void * {
        if (bytes == 0) return nullptr;
        void *d = devAlloc(bytes);
        hostToDevice(d, h, bytes);
        owned.push_back(d);
        return d;
    };
    const PauliExpDev *dOps = static_cast<const PauliExpDev *>(upload(ops.data(), ops.size() * sizeof(PauliExpDev)));
    const PassDev *dPasses = static_cast<const PassDev *>(upload(passes.data(), passes.size() * sizeof(PassDev)));
    const CosetExpDev *dCops =
        static_cast<const CosetExpDev *>(upload(cops.data(), cops.size() * sizeof(CosetExpDev)));
    const CosetGroupDev *dGroups =
        static_cast<const CosetGroupDev *>(upload(groups.data(), groups.size() * sizeof(CosetGroupDev)));
    const unsigned prefetchGrid = 4u * (unsigned)numCUs();
    for (int r = 0; r < reps; ++r) {
        for (const Block &b : blocks) {
            if (b.single >= 0) {
                const PauliExp &e = seq[b.single];
                applyPauliRotation(q, 0, e.x, e.z, e.ny, e.angle);
                continue;
            }
            const unsigned tiles = (unsigned)(q->numAmps >> b.tile.tb);
            const size_t lds = (q->prec == FP64 ? sizeof(Amp<double>) : sizeof(Amp<float>)) << b.tile.tb;
            if (b.coset) {
                if (q->prec == FP64)
                    hipLaunchKernelGGL(kPauliCoset<double>, dim3(tiles), dim3(kBlock), lds, hstream(),
                                       (Amp<double> *)q->amps, b.tile, dGroups + b.firstPass, b.numPasses,
                                       dCops + b.firstOp, b.scale);
                else
                    hipLaunchKernelGGL(kPauliCoset<float>, dim3(tiles), dim3(kBlock), lds, hstream(),
                                       (Amp<float> *)q->amps, b.tile, dGroups + b.firstPass, b.numPasses,
                                       dCops + b.firstOp, b.scale);
                continue;
            }
            // pairs per thread: 2^tb / 2 / kBlock at the full tile size (8 FP32, 4 FP64)
            if (q->prec == FP64)
                hipLaunchKernelGGL((kPauliBlock<double, 4, false>), dim3(tiles), dim3(kBlock), lds, hstream(),
                                   (Amp<double> *)q->amps, tiles, b.tile, dPasses + b.firstPass, b.numPasses,
                                   dOps + b.firstOp, b.scale);
            else
                hipLaunchKernelGGL((kPauliBlock<float, 8, true>), dim3(std::min(tiles, prefetchGrid)),
                                   dim3(kBlock), lds, hstream(), (Amp<float> *)q->amps, tiles, b.tile,
                                   dPasses + b.firstPass, b.numPasses, dOps + b.firstOp, b.scale);
        }
        checkLaunch("applyPauliSequence");
    }
    if (!owned.empty()) {
        deviceSync();
        for (void *d : owned) devFree(d);
    }
}

int pauliSequenceBlocks(int n, int prec, const PauliExp *seqIn, int countIn) {
    // the block count applyPauliSequence would launch per repetition
    const int tb = pauliTileBits(n, prec);
    const int maxX = pauliMaxX(tb, prec);
    const std::vector<PauliExp> packed = packCommuting(seqIn, countIn, maxX);
    const PauliExp *seq = packed.data();
    const int count = (int)packed.size();
    int blocks = 0, i = 0;
    while (i < count) {
        if ((seq[i].x | seq[i].z) == 0) { ++i; continue; }
        if (__builtin_popcountll(seq[i].x) > maxX) { ++blocks; ++i; continue; }
        uint64_t xu = 0;
        int nops = 0;
        while (i < count && nops < kMaxBlockOps && __builtin_popcountll(xu | seq[i].x) <= maxX) {
            nops += (seq[i].x | seq[i].z) != 0;
            xu |= seq[i++].x;
        }
        ++blocks;
    }
    return blocks;
}

}  // namespace qmx
