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
    unsigned pxz;        // parity(lx & lz) of the pass's exchange mask: sign of the partner
    int diag;            // no X/Y factor
    int pad;
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

// PPT amplitude pairs per thread and pass (tiles of 2^(9 + log2 PPT) amplitudes):
// a thread keeps its pairs in registers while every rotation of the pass is
// applied, so a rotation's constants are read once per thread, from scalar
// loads, with the tile base's parity folded into its signs. A workgroup walks
// tiles w = blockIdx.x + k gridDim.x, issuing the global loads of its next
// tile before the passes of the current one.
template <typename T, int PPT>
__global__ void __launch_bounds__(kBlock) kPauliBlock(Amp<T> *__restrict__ a, unsigned numTiles, PauliTile tile,
                                                      const PassDev *__restrict__ passes, int numPasses,
                                                      const PauliExpDev *__restrict__ ops) {
    constexpr int kPer = 2 * PPT;  // amplitudes a thread loads and stores per tile
    extern __shared__ unsigned char smem[];
    Amp<T> *st = reinterpret_cast<Amp<T> *>(smem);
    const int dim = 1 << tile.tb;
    const int half = dim >> 1;
    // tile index t = threadIdx.x + kBlock k: the scatter of its bits splits the same way
    const int lowBits = tile.tb < 8 ? tile.tb : 8;
    uint64_t dt = 0;
    for (int j = 0; j < lowBits; ++j)
        if ((threadIdx.x >> j) & 1) dt |= 1ull << tile.pos[j];
    uint64_t dk[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        uint64_t g = dt;
        for (int j = 8; j < tile.tb; ++j)
            if ((k >> (j - 8)) & 1) g |= 1ull << tile.pos[j];
        dk[k] = g;
    }
    const bool active = (int)threadIdx.x < dim;
    Amp<T> next[kPer];
    unsigned w = blockIdx.x;
    if (w < numTiles) {
        const uint64_t nb = insertZeros((uint64_t)w, tile.fixed);
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if (active && (k << 8) < dim) next[k] = a[nb | dk[k]];
    }
    for (; w < numTiles; w += gridDim.x) {
        const uint64_t base = insertZeros((uint64_t)w, tile.fixed);
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if (active && (k << 8) < dim) st[threadIdx.x + (k << 8)] = next[k];
        __syncthreads();
        const unsigned wn = w + gridDim.x;
        if (wn < numTiles) {
            const uint64_t nb = insertZeros((uint64_t)wn, tile.fixed);
#pragma unroll
            for (int k = 0; k < kPer; ++k)
                if (active && (k << 8) < dim) next[k] = a[nb | dk[k]];
        }
        for (int ps = 0; ps < numPasses; ++ps) {
            const PassDev pass = passes[ps];
            // pairs (i, i ^ lx) of this thread; lx == 0: amplitudes i and i + half
            const unsigned lx = pass.lx;
            const unsigned low = lx & (~lx + 1u);
            const int lb = lx ? __ffs(lx) - 1 : 0;
            unsigned idx[PPT];
            bool ok[PPT];
            Amp<T> vi[PPT], vp[PPT];
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                const unsigned u = threadIdx.x + kBlock * k;
                ok[k] = (int)u < half;
                idx[k] = lx ? (((u >> lb) << (lb + 1)) | (u & (low - 1u))) : u;
                if (ok[k]) {
                    vi[k] = st[idx[k]];
                    vp[k] = st[lx ? (idx[k] ^ lx) : (idx[k] + half)];
                }
            }
            for (int o = 0; o < pass.count; ++o) {
                const PauliExpDev e = ops[pass.first + o];
                const bool flip = (__popcll(base & e.z) & 1) != 0;
                const T c = (T)e.c;
                if (e.diag) {
                    const T s = flip ? -(T)e.s : (T)e.s;
#pragma unroll
                    for (int k = 0; k < PPT; ++k) {
                        const unsigned pi = (unsigned)__popc(idx[k] & e.lz) & 1u;
                        const unsigned pp = lx ? (pi ^ e.pxz) : ((unsigned)__popc((idx[k] + half) & e.lz) & 1u);
                        const T fi = pi ? -s : s, fp = pp ? -s : s;
                        const Amp<T> x = vi[k], y = vp[k];
                        vi[k] = Amp<T>{c * x.x + fi * x.y, c * x.y - fi * x.x};
                        vp[k] = Amp<T>{c * y.x + fp * y.y, c * y.y - fp * y.x};
                    }
                } else {
                    const T kre = flip ? -(T)e.kre : (T)e.kre, kim = flip ? -(T)e.kim : (T)e.kim;
#pragma unroll
                    for (int k = 0; k < PPT; ++k) {
                        // the sign on i's partner term is p's parity, and vice versa
                        const unsigned pi = (unsigned)__popc(idx[k] & e.lz) & 1u;
                        const T si = (pi ^ e.pxz) ? (T)-1 : (T)1;
                        const T sp = pi ? (T)-1 : (T)1;
                        const Amp<T> x = vi[k], y = vp[k];
                        vi[k] = Amp<T>{c * x.x + si * (kre * y.x - kim * y.y), c * x.y + si * (kre * y.y + kim * y.x)};
                        vp[k] = Amp<T>{c * y.x + sp * (kre * x.x - kim * x.y), c * y.y + sp * (kre * x.y + kim * x.x)};
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                if (ok[k]) {
                    st[idx[k]] = vi[k];
                    st[lx ? (idx[k] ^ lx) : (idx[k] + half)] = vp[k];
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if (active && (k << 8) < dim) a[base | dk[k]] = st[threadIdx.x + (k << 8)];
        __syncthreads();  // the next tile overwrites the LDS copy
    }
}

constexpr int kMaxBlockOps = 1 << 16;  // rotations per block (bounds the host's records only)

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
        // partner signs: parity(lx & lz) with each pass's final exchange mask
        for (int ps = blk.firstPass; ps < (int)passes.size(); ++ps)
            for (int o = 0; o < passes[ps].count; ++o) {
                PauliExpDev &d = ops[blk.firstOp + passes[ps].first + o];
                d.pxz = (unsigned)__builtin_popcount(passes[ps].lx & d.lz) & 1u;
            }
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
            // one workgroup per tile measured best (a persistent grid of 3 per CU
            // that prefetches its next tile was no faster: the passes are
            // VALU-bound, not waiting on HBM)
            const unsigned grid = tiles;
            // pairs per thread: 2^tb / 2 / kBlock at the full tile size (8 FP32, 4 FP64)
            if (q->prec == FP64)
                hipLaunchKernelGGL((kPauliBlock<double, 4>), dim3(grid), dim3(kBlock),
                                   (size_t)sizeof(Amp<double>) << b.tile.tb, hstream(), (Amp<double> *)q->amps, tiles,
                                   b.tile, dPasses + b.firstPass, b.numPasses, dOps + b.firstOp);
            else
                hipLaunchKernelGGL((kPauliBlock<float, 8>), dim3(grid), dim3(kBlock),
                                   (size_t)sizeof(Amp<float>) << b.tile.tb, hstream(), (Amp<float> *)q->amps, tiles,
                                   b.tile, dPasses + b.firstPass, b.numPasses, dOps + b.firstOp);
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
