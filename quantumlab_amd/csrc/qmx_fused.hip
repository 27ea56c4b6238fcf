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

struct Fused5Prepared {
    Fused5Spec spec;
    float *M;  // device copy of the block's three real 32x32 matrices (Ur, Ui, Ur + Ui)
};

constexpr int kFused5MatrixFloats = 3 * 32 * 32;

// The kernel's form of a 5-qubit block (targs[0] = least significant bit of
// u's index, u row-major 32x32): the spec, and the three real 32x32 matrices
// Ur, Ui and Ur + Ui (row-major, one after the other) written to hm.
void fused5Build(const int *targs, const Cplx *u, Fused5Spec &spec, float *hm) {
    // index the block by ascending target: bit j of a combination is the j-th
    // lowest target, so a wave's 4 combinations (MFMA K lanes) run over the
    // lowest targets and its loads cover whole cache lines
    int order[5] = {0, 1, 2, 3, 4};
    std::sort(order, order + 5, [&](int x, int y) { return targs[x] < targs[y]; });
    spec.targetsSorted.n = 5;
    for (int j = 0; j < 5; ++j) spec.targetsSorted.pos[j] = targs[order[j]];
    int toOld[32];
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
    for (int r = 0; r < 32; ++r)
        for (int c = 0; c < 32; ++c) {
            const Cplx z = u[toOld[r] * 32 + toOld[c]];
            hm[r * 32 + c] = (float)z.re;
            hm[1024 + r * 32 + c] = (float)z.im;
            hm[2048 + r * 32 + c] = (float)(z.re + z.im);
        }
}

void launchFused5(QuregDev *q, const Fused5Spec &spec, const float *M) {
    const long long numChunks = (q->numAmps >> 5) / 16;
    const long long wavesNeeded = (numChunks + kChunksPerWave - 1) / kChunksPerWave;
    const long long grid = (wavesNeeded + (kBlock / 64) - 1) / (kBlock / 64);
    hipLaunchKernelGGL(kFused5, dim3((unsigned)grid), dim3(kBlock), 0, hstream(), (Amp<float> *)q->amps,
                       numChunks, spec, M);
    checkLaunch("fused5Apply");
}

}  // namespace

// Build the device-resident form of a 5-qubit block once, for blocks applied
// many times (circuits).
void *fused5Prepare(const int *targs, const Cplx *u) {
    auto *p = new Fused5Prepared;
    std::vector<float> hm(kFused5MatrixFloats);
    fused5Build(targs, u, p->spec, hm.data());
    p->M = static_cast<float *>(devAlloc(hm.size() * sizeof(float)));
    hostToDevice(p->M, hm.data(), hm.size() * sizeof(float));
    return p;
}

// One application of a block (multiQubitUnitary on an FP32 register): the
// matrix goes through the staging ring, with no allocation or synchronisation.
void fused5ApplyOnce(QuregDev *q, const int *targs, const Cplx *u) {
    constexpr size_t bytes = kFused5MatrixFloats * sizeof(float);
    static_assert(bytes <= kStageSlotBytes, "a fused block fits a staging slot");
    Fused5Spec spec;
    hipStream_t st = hstream();
    const Staged s = stageBegin(bytes);
    fused5Build(targs, u, spec, static_cast<float *>(s.host));
    stageCommit(s, bytes, st);
    launchFused5(q, spec, static_cast<const float *>(s.dev));
}

void fused5Free(void *prepared) {
    auto *p = static_cast<Fused5Prepared *>(prepared);
    if (!p) return;
    devFree(p->M);
    delete p;
}

// Apply a prepared block to an FP32 register of >= 9 qubits (one 16-group
// chunk is 512 amplitudes).
void fused5Apply(QuregDev *q, const void *prepared) {
    const auto *p = static_cast<const Fused5Prepared *>(prepared);
    launchFused5(q, p->spec, p->M);
}

}  // namespace qmx
