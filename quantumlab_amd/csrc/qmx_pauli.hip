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
// and rotation) xor parity(local & z_local). The per-rotation arithmetic is
// the one-sweep kernels' (kPauliRot, kRotZ), term for term. Inside a tile the
// rotations run in passes: consecutive rotations whose X/Y masks coincide
// (e.g. XX and YY on one bond), plus any diagonal rotations between them, are
// applied to each amplitude pair in registers in one LDS round trip.
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
    int diag;            // no X/Y factor
    double c, s;         // cos(theta/2), sin(theta/2)
    double kre, kim;     // -i sin(theta/2) i^nY
};

// a pass: consecutive rotations of a block whose X/Y parts share one tile-local
// mask lx (diagonal rotations join any pass); lx == 0: diagonal rotations only
struct PassDev {
    unsigned lx;
    int first, count;
    int pad;
};

struct PauliTile {
    int tb;                   // tile bits
    int pos[kMaxTileBits];    // global bit of tile bit j, ascending
    BitList fixed;            // the same positions (tile base = insert zeros there)
};

// a rotation as one tile uses it: working precision, and the parity of the
// tile base folded into the signs (s for diagonal rotations, k otherwise)
template <typename T>
struct OpT {
    unsigned lz;
    int diag;
    T c, s, kre, kim;
};

template <typename T>
__device__ __forceinline__ void rotatePair(Amp<T> &vi, Amp<T> &vp, const OpT<T> &e, unsigned i, unsigned p) {
    if (e.diag) {
        const T fi = (__popc(i & e.lz) & 1) ? -e.s : e.s;
        const T fp = (__popc(p & e.lz) & 1) ? -e.s : e.s;
        vi = Amp<T>{e.c * vi.x + fi * vi.y, e.c * vi.y - fi * vi.x};
        vp = Amp<T>{e.c * vp.x + fp * vp.y, e.c * vp.y - fp * vp.x};
        return;
    }
    const T si = (__popc(p & e.lz) & 1) ? (T)-1 : (T)1;
    const T sp = (__popc(i & e.lz) & 1) ? (T)-1 : (T)1;
    const Amp<T> a = vi, b = vp;
    vi = Amp<T>{e.c * a.x + si * (e.kre * b.x - e.kim * b.y), e.c * a.y + si * (e.kre * b.y + e.kim * b.x)};
    vp = Amp<T>{e.c * b.x + sp * (e.kre * a.x - e.kim * a.y), e.c * b.y + sp * (e.kre * a.y + e.kim * a.x)};
}

constexpr int kMaxBlockOps = 128;  // rotations staged per tile (LDS)

template <typename T>
__global__ void __launch_bounds__(kBlock) kPauliBlock(Amp<T> *__restrict__ a, PauliTile tile,
                                                      const PassDev *__restrict__ passes, int numPasses,
                                                      const PauliExpDev *__restrict__ ops, int numOps) {
    extern __shared__ unsigned char smem[];
    __shared__ OpT<T> sOps[kMaxBlockOps];
    Amp<T> *st = reinterpret_cast<Amp<T> *>(smem);
    const int dim = 1 << tile.tb;
    const uint64_t base = insertZeros((uint64_t)blockIdx.x, tile.fixed);
    for (int o = threadIdx.x; o < numOps; o += kBlock) {
        const PauliExpDev e = ops[o];
        const bool flip = (__popcll(base & e.z) & 1) != 0;
        OpT<T> t;
        t.lz = e.lz;
        t.diag = e.diag;
        t.c = (T)e.c;
        t.s = flip ? -(T)e.s : (T)e.s;
        t.kre = flip ? -(T)e.kre : (T)e.kre;
        t.kim = flip ? -(T)e.kim : (T)e.kim;
        sOps[o] = t;
    }
    // tile index t = threadIdx.x + kBlock k: the scatter of its bits splits the same way
    const int lowBits = tile.tb < 8 ? tile.tb : 8;
    uint64_t dt = base;
    for (int j = 0; j < lowBits; ++j)
        if ((threadIdx.x >> j) & 1) dt |= 1ull << tile.pos[j];
    const bool active = (int)threadIdx.x < dim;
    for (int k = 0; active && (k << 8) < dim; ++k) {
        uint64_t g = dt;
        for (int j = 8; j < tile.tb; ++j)
            if ((k >> (j - 8)) & 1) g |= 1ull << tile.pos[j];
        st[threadIdx.x + (k << 8)] = a[g];
    }
    __syncthreads();
    for (int ps = 0; ps < numPasses; ++ps) {
        const PassDev pass = passes[ps];
        if (pass.lx == 0) {
            // diagonal rotations: own amplitudes, no exchange
            for (int t = threadIdx.x; t < dim; t += kBlock) {
                Amp<T> v = st[t], dummy = v;
                for (int o = 0; o < pass.count; ++o) rotatePair<T>(v, dummy, sOps[pass.first + o], t, t);
                st[t] = v;
            }
            __syncthreads();
            continue;
        }
        const unsigned low = pass.lx & (~pass.lx + 1u);
        const int lb = __ffs(pass.lx) - 1;
        for (int w = threadIdx.x; w < (dim >> 1); w += kBlock) {
            const unsigned i = (((unsigned)w >> lb) << (lb + 1)) | ((unsigned)w & (low - 1u));
            const unsigned p = i ^ pass.lx;
            Amp<T> vi = st[i], vp = st[p];
            for (int o = 0; o < pass.count; ++o) rotatePair<T>(vi, vp, sOps[pass.first + o], i, p);
            st[i] = vi;
            st[p] = vp;
        }
        __syncthreads();
    }
    for (int k = 0; active && (k << 8) < dim; ++k) {
        uint64_t g = dt;
        for (int j = 8; j < tile.tb; ++j)
            if ((k >> (j - 8)) & 1) g |= 1ull << tile.pos[j];
        a[g] = st[threadIdx.x + (k << 8)];
    }
}

struct Block {
    PauliTile tile;
    int firstPass, numPasses;
    int firstOp, numOps;
    int single;        // >= 0: one rotation too wide for a tile, applied alone (index into seq)
};

}  // namespace

void applyPauliSequence(QuregDev *q, const PauliExp *seq, int count, int reps) {
    const int n = q->numQubits;
    const int tb = std::min(n, q->prec == FP64 ? 11 : 12);  // 32 KiB of LDS per tile
    const int maxX = std::max(1, tb - std::min(tb, 4));     // keep >= 4 low bits for contiguous runs
    std::vector<PauliExpDev> ops;
    std::vector<PassDev> passes;
    std::vector<Block> blocks;
    int i = 0;
    while (i < count) {
        if ((seq[i].x | seq[i].z) == 0) {  // identity string: a no-op, as in multiRotatePauli
            ++i;
            continue;
        }
        if (__builtin_popcountll(seq[i].x) > maxX) {
            blocks.push_back(Block{PauliTile{}, 0, 0, 0, 0, i});
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
        blk.tile.fixed.n = 0;
        for (int b = 0; b < n; ++b)
            if ((tm >> b) & 1) {
                blk.tile.pos[blk.tile.tb++] = b;
                blk.tile.fixed.pos[blk.tile.fixed.n++] = b;
            }
        blk.firstPass = (int)passes.size();
        blk.firstOp = (int)ops.size();
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
            d.diag = e.x == 0;
            d.c = std::cos(e.angle / 2.0);
            d.s = std::sin(e.angle / 2.0);
            static const double phr[4] = {1, 0, -1, 0}, phi[4] = {0, 1, 0, -1};
            d.kre = d.s * phi[e.ny & 3];
            d.kim = -d.s * phr[e.ny & 3];
            const int rel = (int)ops.size() - blk.firstOp;  // passes index the block's own ops
            if (open < 0) {
                passes.push_back(PassDev{lx, rel, 0, 0});
                open = (int)passes.size() - 1;
            } else if (lx != 0 && passes[open].lx != 0 && passes[open].lx != lx) {
                passes.push_back(PassDev{lx, rel, 0, 0});
                open = (int)passes.size() - 1;
            } else if (lx != 0 && passes[open].lx == 0) {
                passes[open].lx = lx;  // diagonal rotations so far ride along with this exchange
            }
            ops.push_back(d);
            ++passes[open].count;
        }
        blk.numPasses = (int)passes.size() - blk.firstPass;
        blk.numOps = (int)ops.size() - blk.firstOp;
        if (blk.numPasses > 0) blocks.push_back(blk);
        i = j;
    }
    if (blocks.empty()) return;
    PauliExpDev *dOps = nullptr;
    PassDev *dPasses = nullptr;
    if (!ops.empty()) {
        dOps = static_cast<PauliExpDev *>(devAlloc(ops.size() * sizeof(PauliExpDev)));
        hostToDevice(dOps, ops.data(), ops.size() * sizeof(PauliExpDev));
        dPasses = static_cast<PassDev *>(devAlloc(passes.size() * sizeof(PassDev)));
        hostToDevice(dPasses, passes.data(), passes.size() * sizeof(PassDev));
    }
    for (int r = 0; r < reps; ++r) {
        for (const Block &b : blocks) {
            if (b.single >= 0) {
                const PauliExp &e = seq[b.single];
                applyPauliRotation(q, 0, e.x, e.z, e.ny, e.angle);
                continue;
            }
            const unsigned tiles = (unsigned)(q->numAmps >> b.tile.tb);
            QMX_DISPATCH(q->prec, T, {
                hipLaunchKernelGGL(kPauliBlock<T>, dim3(tiles), dim3(kBlock), (size_t)sizeof(Amp<T>) << b.tile.tb,
                                   hstream(), (Amp<T> *)q->amps, b.tile, dPasses + b.firstPass, b.numPasses,
                                   dOps + b.firstOp, b.numOps);
            });
        }
        checkLaunch("applyPauliSequence");
    }
    if (dOps) {
        deviceSync();
        devFree(dOps);
        devFree(dPasses);
    }
}

int pauliSequenceBlocks(int n, int prec, const PauliExp *seq, int count) {
    // the block count applyPauliSequence would launch per repetition
    const int tb = std::min(n, prec == FP64 ? 11 : 12);
    const int maxX = std::max(1, tb - std::min(tb, 4));
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
