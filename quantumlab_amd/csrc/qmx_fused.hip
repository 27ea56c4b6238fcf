// Fused dense 5-qubit blocks on FP32 amplitudes, as GEMMs on MFMA.
//
// A block U (32x32 complex) acting on 5 target qubits maps every group of 32
// amplitudes (one per target combination, the other qubits fixed) to U times
// that group. For 16 groups at once, with X = Xr + i Xi (32x16), the product
// is formed from three real 32x32 by 32x16 products (Gauss):
//   P1 = Ur Xr,  P2 = Ui Xi,  P3 = (Ur + Ui)(Xr + Xi),
//   Re(UX) = P1 - P2,  Im(UX) = P3 - P1 - P2,
// each as 2 row tiles x 8 steps of v_mfma_f32_16x16x4f32: 48 MFMAs per 16
// groups, where the real 64x64 form [[Ur, -Ui], [Ui, Ur]] takes 64. Ur + Ui is
// formed on the host in FP64 and rounded once; Xr + Xi costs 8 FP32 adds per
// lane. A row tile's real and imaginary outputs come from the same three
// products, so a tile is finished and stored before the next one starts.
//
// Operand placement (16x16x4 f32 MFMA, wave64):
//   A (16x4): lane l holds A[l%16][l/16]  -> matrix fragments, staged once per
//             workgroup in LDS as [product][row tile][row][k lane][step] with
//             rows 36 floats apart (so 8 consecutive lanes cover the 32 banks),
//             a lane reading its 8 values of a product and row tile as two
//             16 B vectors;
//   B (4x16): lane l holds B[l/16][l%16]  -> lane l loads amplitude
//             (group l%16, combination 4s + l/16) for s = 0..7: its .x feeds
//             P1's step s, its .y P2's, their sum P3's;
//   D (16x16): lane l holds D[4(l/16)+v][l%16], v = 0..3.
// Amplitudes go from HBM straight into MFMA operands and back. The block is
// indexed by ascending target (combination bit j = j-th lowest target), so a
// load's 4 K lanes run over the two lowest targets and its 16 columns over
// consecutive groups: whole cache lines whenever the low qubits are targets or
// groups. Each wave takes 4 consecutive 16-group chunks and loads the next
// chunk's amplitudes before multiplying the current one.
//
// Cost per amplitude: 192 FP32 MFMA flops for 16 B of HBM traffic (the 4-product
// form: 256), which keeps the matrix cores well inside the time MI355X's HBM
// needs for the sweep; every fused block replaces ~10 per-gate sweeps.
#include <algorithm>
#include <cstdio>
#include <vector>

#include "qmx_device.h"

namespace qmx {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

struct Fused5Spec {
    BitList targetsSorted;   // 5 ascending target positions (for group -> base index)
    uint64_t off[32];        // offset of combination c: sum_j bit_j(c) << targs[j]
};

constexpr int kChunksPerWave = 4;  // 16-group chunks handled by one wave

// index of group g's combination 0: g with zeros inserted at the 5 target
// positions (a constant count, so the loop unrolls over scalar positions)
__device__ __forceinline__ uint64_t fusedGroupBase(uint64_t g, const BitList &t) {
#pragma unroll
    for (int j = 0; j < 5; ++j) g = insertZero(g, t.pos[j]);
    return g;
}

// matrix fragments in LDS: [product][row tile][row: kRowStride floats][k lane][step]
constexpr int kRowStride = 36;
constexpr int kTileStride = 16 * kRowStride;
constexpr int kProductStride = 2 * kTileStride;

__global__ void __launch_bounds__(kBlock, 5) kFused5(Amp<float> *__restrict__ a, long long numChunks,
                                                  Fused5Spec spec, const float *__restrict__ M) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col = lane & 15, kq = lane >> 4;
    __shared__ __attribute__((aligned(16))) float sM[3 * kProductStride];
    for (int e = threadIdx.x; e < 3 * 1024; e += kBlock) {
        // e = ((p * 2 + rt) * 16 + row) * 32 + kk * 8 + ss: M[p][16 rt + row][4 ss + kk]
        const int p = e >> 10, rt = (e >> 9) & 1, row = (e >> 5) & 15, kk = (e >> 3) & 3, ss = e & 7;
        sM[p * kProductStride + rt * kTileStride + row * kRowStride + kk * 8 + ss] =
            M[p * 1024 + (16 * rt + row) * 32 + 4 * ss + kk];
    }
    __syncthreads();

    const long long firstChunk = (xcdBlock() * (kBlock / 64) + wave) * kChunksPerWave;
    // Combination offsets split into a per-lane part and a wave-uniform part:
    // input c = 4 s + kq has off[c] = off[kq] | off[4 s], output c = 16 h +
    // 4 kq + v has off[c] = off[4 kq] | off[16 h + v] (disjoint bits). The
    // uniform parts stay in scalar registers, so no lane reads the offset
    // table from memory: such reads queue behind the prefetched amplitudes,
    // and waiting for them would wait for the prefetch too.
    const uint64_t inLane = kq == 0 ? 0 : kq == 1 ? spec.off[1] : kq == 2 ? spec.off[2] : spec.off[3];
    const uint64_t outLane = kq == 0 ? 0 : kq == 1 ? spec.off[4] : kq == 2 ? spec.off[8] : spec.off[12];
    const float *lm = sM + col * kRowStride + kq * 8;
    // the next chunk's amplitudes are loaded while the current chunk multiplies
    Amp<float> x[8], xn[8];
    uint64_t base = 0, nbase = 0;
    if (firstChunk < numChunks) {
        nbase = fusedGroupBase((uint64_t)firstChunk * 16 + col, spec.targetsSorted);
#pragma unroll
        for (int s = 0; s < 8; ++s) xn[s] = ldA<float, false>(a + ((nbase | inLane) + spec.off[4 * s]));
    }
    for (int it = 0; it < kChunksPerWave; ++it) {
        const long long chunk = firstChunk + it;
        if (chunk >= numChunks) break;
        base = nbase;
#pragma unroll
        for (int s = 0; s < 8; ++s) x[s] = xn[s];
        if (it + 1 < kChunksPerWave && chunk + 1 < numChunks) {
            nbase = fusedGroupBase((uint64_t)(chunk + 1) * 16 + col, spec.targetsSorted);
#pragma unroll
            for (int s = 0; s < 8; ++s) xn[s] = ldA<float, false>(a + ((nbase | inLane) + spec.off[4 * s]));
        }
        float xs[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) xs[s] = x[s].x + x[s].y;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            f4 p1 = f4{0.f, 0.f, 0.f, 0.f}, p2 = p1, p3 = p1;
            const f4 *ar = reinterpret_cast<const f4 *>(lm + h * kTileStride);
            const f4 *ai = reinterpret_cast<const f4 *>(lm + kProductStride + h * kTileStride);
            const f4 *as = reinterpret_cast<const f4 *>(lm + 2 * kProductStride + h * kTileStride);
            const f4 r0 = ar[0], r1 = ar[1], i0 = ai[0], i1 = ai[1], s0 = as[0], s1 = as[1];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                p1 = __builtin_amdgcn_mfma_f32_16x16x4f32(r0[s], x[s].x, p1, 0, 0, 0);
                p2 = __builtin_amdgcn_mfma_f32_16x16x4f32(i0[s], x[s].y, p2, 0, 0, 0);
                p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(s0[s], xs[s], p3, 0, 0, 0);
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                p1 = __builtin_amdgcn_mfma_f32_16x16x4f32(r1[s], x[s + 4].x, p1, 0, 0, 0);
                p2 = __builtin_amdgcn_mfma_f32_16x16x4f32(i1[s], x[s + 4].y, p2, 0, 0, 0);
                p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(s1[s], xs[s + 4], p3, 0, 0, 0);
            }
            // lane holds output rows 4 kq + v of tile h: combination c = 16 h + 4 kq + v
#pragma unroll
            for (int v = 0; v < 4; ++v)
                stA<float, false>(a + ((base | outLane) + spec.off[16 * h + v]),
                                  Amp<float>{p1[v] - p2[v], p3[v] - p1[v] - p2[v]});
        }
    }
}

// ---- tile form (registers of 11 qubits and more) -----------------------------
// A workgroup owns tiles of 64 groups x 32 combinations = 2048 amplitudes
// (16 KiB). A tile's index bits are P = the 5 targets and the 6 lowest other
// qubits, so P always holds qubits 0-5 and the tile is a union of contiguous
// runs of at least 64 amplitudes (512 B): it moves between HBM and LDS as
// coalesced nontemporal 16 B accesses whatever the targets, like the in-place
// streams. Each wave multiplies 16 groups out of LDS (the MFMA forms above) and
// writes the results back in place; the workgroup then stores the tile. LDS
// holds two tiles (double buffering, one barrier fewer), and the next two
// tiles' loads are in flight while the current one multiplies and is stored.
//
// Each workgroup takes a few consecutive tiles (kTilesPerGroup), and the
// hardware dispatches workgroups in index order, so at any moment the chip
// works on a compact window of the register. A persistent grid (one long
// range of tiles per workgroup, 768 ranges streamed at once) measured 7-8 %
// slower at 32 qubits, with or without the products: the pure load/store
// skeleton of the tile pattern took 12.9 ms per 32 GiB sweep persistent and
// 11.9 ms with one tile per workgroup. What a workgroup needs besides its
// tiles comes precomputed on the host (the tile form below), so its prologue
// is 13 coalesced loads.
//
// LDS slot of tile element e: a linear XOR swizzle (slotBit[i] = slot of
// element 2^i) chosen on the host so that the 16 lanes of an MFMA operand read
// (one combination, 16 consecutive groups) hit 16 different 8 B bank pairs
// whichever bit positions the groups take.
constexpr int kTileLog = 11;
constexpr int kTileGroups = 64;
constexpr int kTileAmps = 32 * kTileGroups;
constexpr int kTilesPerGroup = 4;

struct Fused5TileSpec {
    int pos[kTileLog];          // P, ascending
    int tIdx[5];                // index in P of the j-th lowest target (combination bit j)
    int gIdx[6];                // index in P of group bit b
    unsigned slotBit[kTileLog];  // swizzle: slot of element 2^i
    // workgroup-uniform slots and offsets (see kFused5Tile)
    unsigned sIn[8], sOut[8], jSlot[4];
    uint64_t jAddr[4];
    long long numTiles, tilesPerGroup;
};

template <class Spec>
__host__ __device__ __forceinline__ unsigned tileSlot(unsigned e, const Spec &sp) {
    unsigned s = 0;
#pragma unroll
    for (int i = 0; i < kTileLog; ++i)
        if ((e >> i) & 1) s ^= sp.slotBit[i];
    return s;
}

template <class Spec>
__host__ __device__ __forceinline__ unsigned tileCombo(unsigned c, const Spec &sp) {
    unsigned e = 0;
#pragma unroll
    for (int j = 0; j < 5; ++j) e |= ((c >> j) & 1u) << sp.tIdx[j];
    return e;
}

// state index of tile t's element 0: t with zeros inserted at the P positions
template <class Spec>
__device__ __forceinline__ uint64_t tileBase(uint64_t t, const Spec &sp) {
#pragma unroll
    for (int i = 0; i < kTileLog; ++i) t = insertZero(t, sp.pos[i]);
    return t;
}

// The tile form of a block, built on the host (kFused5TileFormBytes, 16 KiB):
//   frag[i][lane] (i < 12, float4): lane (col = lane % 16, kq = lane / 16)'s
//     48 matrix fragments M_p[16 h + col][4 st + kq], fragment f = (3 h + p) 8 +
//     st in component f % 4 of vector f / 4 (one coalesced 1 KiB load each);
//   lanes[tid] (uint4): {inLane | outLane << 16, laneSlot, laneAddr low, high}
//     for thread tid of a workgroup (LDS slots of its MFMA operands and of its
//     transfer elements, and the state offset of its transfer elements).
constexpr int kFragVecs = 12;
constexpr size_t kFused5TileFormBytes = kFragVecs * 64 * 16 + kBlock * 16;

typedef float f4v __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(kBlock, 3) kFused5Tile(Amp<float> *__restrict__ a, Fused5TileSpec sp,
                                                      const f4v *__restrict__ form) {
    __shared__ __attribute__((aligned(16))) Amp<float> buf[2][kTileAmps];
    const int tid = threadIdx.x, lane = tid & 63;
    const long long first = xcdBlock() * sp.tilesPerGroup;
    const long long last = min(first + sp.tilesPerGroup, sp.numTiles);
    if (first >= last) return;  // the same for every thread of the workgroup

    // matrix fragments: product p, row tile h, step st -> M_p[16h + col][4st + kq]
    f4v fv[kFragVecs];
#pragma unroll
    for (int i = 0; i < kFragVecs; ++i) fv[i] = form[i * 64 + lane];
    const uint4 L = reinterpret_cast<const uint4 *>(form + kFragVecs * 64)[tid];
    // LDS slots of this lane's operands (group g = 16 wave + col; input
    // combination 4s + kq, output combination 16h + 4kq + v; linear swizzle:
    // slot(x | y) = slot(x) ^ slot(y) for disjoint bits) and of its transfer
    // elements e = 2 tid + 512 j and e + 1 (j < 4): bits 0-5 of e are qubits
    // 0-5, bits 6-10 go to P[6..10]; the parts from tid (bits 0-8) and from j
    // (bits 9-10) are disjoint, so address and slot split into a per-lane and
    // a uniform term.
    const unsigned inLane = L.x & 0xffffu, outLane = L.x >> 16, laneSlot = L.y;
    const uint64_t laneAddr = (uint64_t)L.z | ((uint64_t)L.w << 32);
    const unsigned slot1 = sp.slotBit[0];  // slot(e + 1) = slot(e) ^ slot(1) for even e

    // Two tiles' loads in flight: tile t + 2 is requested while tile t
    // multiplies and is stored, into the register set tile t came in (two
    // sets, the loop unrolled by two; LDS buffer = tile parity).
    auto load = [&](f4v (&R)[4], long long t) {
        const uint64_t base = tileBase((uint64_t)t, sp);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            R[j] = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(a + ((base | laneAddr) | sp.jAddr[j])));
    };
    auto step = [&](f4v (&R)[4], long long t, Amp<float> *T) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            T[laneSlot ^ sp.jSlot[j]] = Amp<float>{R[j].x, R[j].y};
            T[laneSlot ^ sp.jSlot[j] ^ slot1] = Amp<float>{R[j].z, R[j].w};
        }
        __syncthreads();
        if (t + 2 < last) load(R, t + 2);
        // both row tiles accumulate together, so an operand is live for one step
        f4 acc[2][3];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int p = 0; p < 3; ++p) acc[h][p] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int st = 0; st < 8; ++st) {
            const Amp<float> x = T[inLane ^ sp.sIn[st]];
            const float xs = x.x + x.y;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int f0 = (3 * h) * 8 + st, f1 = (3 * h + 1) * 8 + st, f2 = (3 * h + 2) * 8 + st;
                acc[h][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(fv[f0 >> 2][f0 & 3], x.x, acc[h][0], 0, 0, 0);
                acc[h][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(fv[f1 >> 2][f1 & 3], x.y, acc[h][1], 0, 0, 0);
                acc[h][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(fv[f2 >> 2][f2 & 3], xs, acc[h][2], 0, 0, 0);
            }
        }
        // the wave's reads of these 16 groups are done: results go back in place
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                T[outLane ^ sp.sOut[4 * h + v]] = Amp<float>{acc[h][0][v] - acc[h][1][v],
                                                             acc[h][2][v] - acc[h][0][v] - acc[h][1][v]};
        __syncthreads();
        const uint64_t cur = tileBase((uint64_t)t, sp);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const Amp<float> lo = T[laneSlot ^ sp.jSlot[j]], hi = T[laneSlot ^ sp.jSlot[j] ^ slot1];
            __builtin_nontemporal_store(f4v{lo.x, lo.y, hi.x, hi.y},
                                        reinterpret_cast<f4v *>(a + ((cur | laneAddr) | sp.jAddr[j])));
        }
    };
    f4v R0[4], R1[4];
    load(R0, first);
    if (first + 1 < last) load(R1, first + 1);
    for (long long t = first; t < last; t += 2) {
        step(R0, t, buf[0]);
        if (t + 1 < last) step(R1, t + 1, buf[1]);
    }
}

// ---- FP64 tile form ------------------------------------------------------------
// The same tiles on FP64 registers (2048 amplitudes = 32 KiB, two per
// workgroup in LDS, so two workgroups per CU), multiplied with
// v_mfma_f64_16x16x4f64 in the same three-product form. Its result element v
// of row tile h is combination 16 h + 4 v + kq, the input combination of
// step 4 h + v: a lane writes each result to the LDS slot it read an input
// from. A thread moves elements e = tid + 256 j (j < 8), one amplitude per 16
// B access: bits 0-7 of e from tid (qubits 0-5, then P[6], P[7]), bits 8-10
// from j. The block's complex matrix (ascending-target order, 16 KiB) is the
// only per-block input; each lane reads its 16 fragments from it once.
// tiles per workgroup (32 KiB each): at 32 qubits 8, 16 and 32 measured
// 22.4-23.1 ms per block, 2 and 4 26.0-26.1 ms (repeat runs of one setting
// vary by up to 4 ms with the register's placement in memory)
constexpr int kTilesPerGroup64 = 8;

struct Fused5TileSpec64 {
    int pos[kTileLog];
    int tIdx[5];
    int gIdx[6];
    unsigned slotBit[kTileLog];
    unsigned sIn[8], jSlot[8];
    uint64_t jAddr[8];
    long long numTiles, tilesPerGroup;
};

typedef double d2v __attribute__((ext_vector_type(2)));
typedef double d4v __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(kBlock, 2) kFused5Tile64(Amp<double> *__restrict__ a, Fused5TileSpec64 sp,
                                                        const d2v *__restrict__ M) {
    __shared__ __attribute__((aligned(16))) Amp<double> buf[2][kTileAmps];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 15, kq = lane >> 4;
    const long long first = xcdBlock() * sp.tilesPerGroup;
    const long long last = min(first + sp.tilesPerGroup, sp.numTiles);
    if (first >= last) return;  // the same for every thread of the workgroup

    // fragments of row tile h, step st: M[16h + col][4st + kq] (real and
    // imaginary; their sum, the third product's operand, is formed per use)
    double fr[2][8], fi[2][8];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int st = 0; st < 8; ++st) {
            const d2v z = M[(16 * h + col) * 32 + 4 * st + kq];
            fr[h][st] = z.x;
            fi[h][st] = z.y;
        }
    const unsigned g = 16 * wave + col;
    unsigned eG = 0;
#pragma unroll
    for (int b = 0; b < 6; ++b) eG |= ((g >> b) & 1u) << sp.gIdx[b];
    const unsigned inLane = tileSlot(eG | tileCombo(kq, sp), sp);
    uint64_t laneAddr = tid & 63u;
#pragma unroll
    for (int i = 0; i < 2; ++i) laneAddr |= (uint64_t)((tid >> (6 + i)) & 1u) << sp.pos[6 + i];
    const unsigned laneSlot = tileSlot((unsigned)tid, sp);

    auto load = [&](d2v (&R)[8], long long t) {
        const uint64_t base = tileBase((uint64_t)t, sp);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            R[j] = __builtin_nontemporal_load(reinterpret_cast<const d2v *>(a + ((base | laneAddr) | sp.jAddr[j])));
    };
    auto step = [&](d2v (&R)[8], long long t, Amp<double> *T) {
#pragma unroll
        for (int j = 0; j < 8; ++j) T[laneSlot ^ sp.jSlot[j]] = Amp<double>{R[j].x, R[j].y};
        __syncthreads();
        if (t + 1 < last) load(R, t + 1);
        d4v acc[2][3];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int p = 0; p < 3; ++p) acc[h][p] = d4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int st = 0; st < 8; ++st) {
            const Amp<double> x = T[inLane ^ sp.sIn[st]];
            const double xs = x.x + x.y;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                acc[h][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(fr[h][st], x.x, acc[h][0], 0, 0, 0);
                acc[h][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(fi[h][st], x.y, acc[h][1], 0, 0, 0);
                acc[h][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(fr[h][st] + fi[h][st], xs, acc[h][2], 0, 0, 0);
            }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                T[inLane ^ sp.sIn[4 * h + v]] = Amp<double>{acc[h][0][v] - acc[h][1][v],
                                                            acc[h][2][v] - acc[h][0][v] - acc[h][1][v]};
        __syncthreads();
        const uint64_t cur = tileBase((uint64_t)t, sp);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const Amp<double> o = T[laneSlot ^ sp.jSlot[j]];
            __builtin_nontemporal_store(d2v{o.x, o.y}, reinterpret_cast<d2v *>(a + ((cur | laneAddr) | sp.jAddr[j])));
        }
    };
    // one tile's loads in flight (the next tile is requested as soon as the
    // current one is in LDS); LDS buffer = tile parity
    d2v R[8];
    load(R, first);
    for (long long t = first; t < last; t += 2) {
        step(R, t, buf[0]);
        if (t + 1 < last) step(R, t + 1, buf[1]);
    }
}

struct Fused5Prepared {
    Fused5Spec spec;
    float *M;     // device copy of the block's three real 32x32 matrices (Ur, Ui, Ur + Ui)
    void *form;   // device copy of its tile form
};

constexpr int kFused5MatrixFloats = 3 * 32 * 32;

// The kernel's form of a 5-qubit block (targs[0] = least significant bit of
// u's index, u row-major 32x32): the spec, and the three real 32x32 matrices
// Ur, Ui and Ur + Ui (row-major, one after the other) written to hm.
void fused5Order(const int *targs, Fused5Spec &spec, int toOld[32]) {
    // index the block by ascending target: bit j of a combination is the j-th
    // lowest target, so a wave's 4 combinations (MFMA K lanes) run over the
    // lowest targets and its loads cover whole cache lines
    int order[5] = {0, 1, 2, 3, 4};
    std::sort(order, order + 5, [&](int x, int y) { return targs[x] < targs[y]; });
    spec.targetsSorted.n = 5;
    for (int j = 0; j < 5; ++j) spec.targetsSorted.pos[j] = targs[order[j]];
    for (int c = 0; c < 32; ++c) {
        uint64_t o = 0;
        int old = 0;
        for (int j = 0; j < 5; ++j)
            if ((c >> j) & 1) {
                o |= 1ull << targs[order[j]];
                old |= 1 << order[j];
            }
        spec.off[c] = o;
        toOld[c] = old;
    }
}

void fused5Build(const int *targs, const Cplx *u, Fused5Spec &spec, float *hm) {
    int toOld[32];
    fused5Order(targs, spec, toOld);
    for (int r = 0; r < 32; ++r)
        for (int c = 0; c < 32; ++c) {
            const Cplx z = u[toOld[r] * 32 + toOld[c]];
            hm[r * 32 + c] = (float)z.re;
            hm[1024 + r * 32 + c] = (float)z.im;
            hm[2048 + r * 32 + c] = (float)(z.re + z.im);
        }
}

// The tile form's index positions, LDS swizzle and workgroup-uniform slots and
// offsets for a block whose ascending targets are t[0..4].
Fused5TileSpec tileSpecOf(const BitList &targets) {
    Fused5TileSpec sp{};
    uint64_t targ = 0;
    for (int j = 0; j < 5; ++j) targ |= 1ull << targets.pos[j];
    int groupPos[6], ng = 0;
    for (int p = 0; ng < 6; ++p)
        if (!((targ >> p) & 1)) groupPos[ng++] = p;
    uint64_t pm = targ;
    for (int b = 0; b < 6; ++b) pm |= 1ull << groupPos[b];
    int k = 0;
    for (int p = 0; p < 64 && k < kTileLog; ++p)
        if ((pm >> p) & 1) sp.pos[k++] = p;
    auto indexOf = [&](int position) {
        for (int i = 0; i < kTileLog; ++i)
            if (sp.pos[i] == position) return i;
        return -1;
    };
    for (int j = 0; j < 5; ++j) sp.tIdx[j] = indexOf(targets.pos[j]);
    for (int b = 0; b < 6; ++b) sp.gIdx[b] = indexOf(groupPos[b]);
    // swizzle: group bits 0-3 (the lanes of one operand read) onto distinct slot
    // bits 0-3; those already below 4 stay, the others take the free low bits
    for (int i = 0; i < kTileLog; ++i) sp.slotBit[i] = 1u << i;
    bool used[4] = {false, false, false, false};
    for (int b = 0; b < 4; ++b)
        if (sp.gIdx[b] < 4) used[sp.gIdx[b]] = true;
    for (int b = 0; b < 4; ++b) {
        if (sp.gIdx[b] < 4) continue;
        int f = 0;
        while (used[f]) ++f;
        used[f] = true;
        sp.slotBit[sp.gIdx[b]] |= 1u << f;
    }
    for (int st = 0; st < 8; ++st) sp.sIn[st] = tileSlot(tileCombo(4 * st, sp), sp);
    for (int h = 0; h < 2; ++h)
        for (int v = 0; v < 4; ++v) sp.sOut[4 * h + v] = tileSlot(tileCombo(16 * h + v, sp), sp);
    for (int j = 0; j < 4; ++j) {
        sp.jAddr[j] = ((uint64_t)(j & 1) << sp.pos[9]) | ((uint64_t)(j >> 1) << sp.pos[10]);
        sp.jSlot[j] = tileSlot(512u * j, sp);
    }
    return sp;
}

// The tile form (layout at kFused5Tile) of a block from its row-major matrices
// hm (fused5Build) and its tile spec.
void tileFormBuild(const float *hm, const Fused5TileSpec &sp, void *out) {
    float *frag = static_cast<float *>(out);
    for (int lane = 0; lane < 64; ++lane) {
        const int col = lane & 15, kq = lane >> 4;
        for (int f = 0; f < 4 * kFragVecs; ++f) {
            const int h = f / 24, p = (f / 8) % 3, st = f % 8;
            frag[((f >> 2) * 64 + lane) * 4 + (f & 3)] = hm[p * 1024 + (16 * h + col) * 32 + 4 * st + kq];
        }
    }
    uint32_t *lanes = reinterpret_cast<uint32_t *>(frag + kFragVecs * 64 * 4);
    for (int tid = 0; tid < kBlock; ++tid) {
        const int wave = tid >> 6, lane = tid & 63, col = lane & 15, kq = lane >> 4;
        const unsigned g = 16 * wave + col;
        unsigned eG = 0;
        for (int b = 0; b < 6; ++b) eG |= ((g >> b) & 1u) << sp.gIdx[b];
        const unsigned inLane = tileSlot(eG | tileCombo(kq, sp), sp);
        const unsigned outLane = tileSlot(eG | tileCombo(4 * kq, sp), sp);
        const unsigned eLane = 2 * tid;
        uint64_t laneAddr = eLane & 63u;
        for (int i = 0; i < 3; ++i) laneAddr |= (uint64_t)((eLane >> (6 + i)) & 1u) << sp.pos[6 + i];
        lanes[4 * tid + 0] = inLane | (outLane << 16);
        lanes[4 * tid + 1] = tileSlot(eLane, sp);
        lanes[4 * tid + 2] = (uint32_t)laneAddr;
        lanes[4 * tid + 3] = (uint32_t)(laneAddr >> 32);
    }
}

void launchFused5Tile(QuregDev *q, const Fused5Spec &spec, const void *form) {
    Fused5TileSpec sp = tileSpecOf(spec.targetsSorted);
    sp.numTiles = q->numAmps / kTileAmps;
    sp.tilesPerGroup = kTilesPerGroup;
    // a multiple of 8 workgroups (xcdBlock's XCD order); the extra ones exit
    const long long g = ((sp.numTiles + kTilesPerGroup - 1) / kTilesPerGroup + 7) / 8 * 8;
    hipLaunchKernelGGL(kFused5Tile, dim3((unsigned)g), dim3(kBlock), 0, hstream(), (Amp<float> *)q->amps, sp,
                       static_cast<const f4v *>(form));
    checkLaunch("fused5Apply");
}

// The FP64 tile form's spec: the positions of tileSpecOf and a swizzle for
// 16 B LDS accesses (8 lanes per 128 B bank cycle): the slots of elements 1, 2
// and 4 (eight consecutive threads' transfer elements) and those of group bits
// 0-2 (eight lanes of one operand read) each span all 8 slots modulo 8.
// slotBit[i] = 2^i plus bits below min(i, 3) keeps the map invertible.
Fused5TileSpec64 tileSpecOf64(const BitList &targets) {
    const Fused5TileSpec base = tileSpecOf(targets);
    Fused5TileSpec64 sp{};
    for (int i = 0; i < kTileLog; ++i) {
        sp.pos[i] = base.pos[i];
        sp.slotBit[i] = 1u << i;
    }
    for (int j = 0; j < 5; ++j) sp.tIdx[j] = base.tIdx[j];
    for (int b = 0; b < 6; ++b) sp.gIdx[b] = base.gIdx[b];
    auto rank3 = [](unsigned x, unsigned y, unsigned z) {
        return x && y && z && (x ^ y) && (x ^ z) && (y ^ z) && (x ^ y ^ z);
    };
    bool found = false;
    for (unsigned x0 = 0; x0 < 8 && !found; ++x0)
        for (unsigned x1 = 0; x1 < 8 && !found; ++x1)
            for (unsigned x2 = 0; x2 < 8 && !found; ++x2) {
                const unsigned xs[3] = {x0, x1, x2};
                bool ok = true;
                unsigned img[3];
                for (int b = 0; b < 3; ++b) {
                    const int i = sp.gIdx[b];
                    const unsigned allowed = (1u << std::min(i, 3)) - 1u;  // bits below min(i, 3)
                    if (xs[b] & ~allowed) ok = false;
                    img[b] = ((1u << i) & 7u) ^ xs[b];
                }
                if (ok && rank3(img[0], img[1], img[2])) {
                    for (int b = 0; b < 3; ++b) sp.slotBit[sp.gIdx[b]] ^= xs[b];
                    found = true;
                }
            }
    if (!found) {
        std::fprintf(stderr, "libquest_mi355x: no FP64 tile swizzle for these targets\n");
        std::abort();
    }
    for (int st = 0; st < 8; ++st) sp.sIn[st] = tileSlot(tileCombo(4 * st, sp), sp);
    for (int j = 0; j < 8; ++j) {
        sp.jAddr[j] = ((uint64_t)(j & 1) << sp.pos[8]) | ((uint64_t)((j >> 1) & 1) << sp.pos[9]) |
                      ((uint64_t)(j >> 2) << sp.pos[10]);
        sp.jSlot[j] = tileSlot(256u * j, sp);
    }
    return sp;
}

void launchFused5Direct(QuregDev *q, const Fused5Spec &spec, const float *M) {
    const long long numChunks = (q->numAmps >> 5) / 16;
    const long long wavesNeeded = (numChunks + kChunksPerWave - 1) / kChunksPerWave;
    const long long grid = (wavesNeeded + (kBlock / 64) - 1) / (kBlock / 64);
    hipLaunchKernelGGL(kFused5, dim3((unsigned)grid), dim3(kBlock), 0, hstream(), (Amp<float> *)q->amps,
                       numChunks, spec, M);
    checkLaunch("fused5Apply");
}

}  // namespace

// Build the device-resident forms of a 5-qubit block once, for blocks applied
// many times (circuits): the row-major matrices (registers of 9-10 qubits) and
// the tile form (11 qubits and more).
void *fused5Prepare(const int *targs, const Cplx *u) {
    auto *p = new Fused5Prepared;
    std::vector<float> hm(kFused5MatrixFloats);
    fused5Build(targs, u, p->spec, hm.data());
    std::vector<char> form(kFused5TileFormBytes);
    tileFormBuild(hm.data(), tileSpecOf(p->spec.targetsSorted), form.data());
    p->M = static_cast<float *>(devAlloc(hm.size() * sizeof(float)));
    hostToDevice(p->M, hm.data(), hm.size() * sizeof(float));
    p->form = devAlloc(form.size());
    hostToDevice(p->form, form.data(), form.size());
    return p;
}

// One application of a block (multiQubitUnitary on an FP32 register): the
// form the register size needs goes through the staging ring, with no
// allocation or synchronisation.
void fused5ApplyOnce(QuregDev *q, const int *targs, const Cplx *u) {
    constexpr size_t bytes = kFused5MatrixFloats * sizeof(float);
    static_assert(bytes <= kStageSlotBytes && kFused5TileFormBytes <= kStageSlotBytes,
                  "a fused block fits a staging slot");
    Fused5Spec spec;
    hipStream_t st = hstream();
    if (q->numQubits >= kTileLog) {
        float hm[kFused5MatrixFloats];
        fused5Build(targs, u, spec, hm);
        const Staged s = stageBegin(kFused5TileFormBytes);
        tileFormBuild(hm, tileSpecOf(spec.targetsSorted), s.host);
        stageCommit(s, kFused5TileFormBytes, st);
        launchFused5Tile(q, spec, s.dev);
        return;
    }
    const Staged s = stageBegin(bytes);
    fused5Build(targs, u, spec, static_cast<float *>(s.host));
    stageCommit(s, bytes, st);
    launchFused5Direct(q, spec, static_cast<const float *>(s.dev));
}

void fused5Free(void *prepared) {
    auto *p = static_cast<Fused5Prepared *>(prepared);
    if (!p) return;
    devFree(p->M);
    devFree(p->form);
    delete p;
}

// Apply a prepared block to an FP32 register of >= 9 qubits (one 16-group
// chunk is 512 amplitudes).
void fused5Apply(QuregDev *q, const void *prepared) {
    const auto *p = static_cast<const Fused5Prepared *>(prepared);
    if (q->numQubits >= kTileLog) launchFused5Tile(q, p->spec, p->form);
    else launchFused5Direct(q, p->spec, p->M);
}

// One application of an uncontrolled 5-qubit matrix to an FP64 register of
// >= 11 qubits: the block's complex matrix in ascending-target order goes
// through the staging ring (16 KiB), with no allocation or synchronisation.
void fused5ApplyOnce64(QuregDev *q, const int *targs, const Cplx *u) {
    constexpr size_t bytes = 32 * 32 * sizeof(Amp<double>);
    static_assert(bytes <= kStageSlotBytes, "an FP64 block fits a staging slot");
    Fused5Spec spec;
    int toOld[32];
    fused5Order(targs, spec, toOld);
    hipStream_t st = hstream();
    const Staged sg = stageBegin(bytes);
    Amp<double> *hm = static_cast<Amp<double> *>(sg.host);
    for (int r = 0; r < 32; ++r)
        for (int c = 0; c < 32; ++c) {
            const Cplx z = u[toOld[r] * 32 + toOld[c]];
            hm[r * 32 + c] = Amp<double>{z.re, z.im};
        }
    stageCommit(sg, bytes, st);
    Fused5TileSpec64 sp = tileSpecOf64(spec.targetsSorted);
    sp.numTiles = q->numAmps / kTileAmps;
    sp.tilesPerGroup = kTilesPerGroup64;
    const long long g = ((sp.numTiles + kTilesPerGroup64 - 1) / kTilesPerGroup64 + 7) / 8 * 8;
    hipLaunchKernelGGL(kFused5Tile64, dim3((unsigned)g), dim3(kBlock), 0, st, (Amp<double> *)q->amps, sp,
                       static_cast<const d2v *>(sg.dev));
    checkLaunch("fused5ApplyOnce64");
}

}  // namespace qmx
