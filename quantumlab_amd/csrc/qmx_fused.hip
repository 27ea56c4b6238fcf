// Fused dense 5-qubit blocks on FP32 amplitudes, as a GEMM on MFMA.
//
// A block U (32x32 complex) acting on 5 target qubits maps every group of 32
// amplitudes (one per target combination, the other qubits fixed) to U times
// that group. Written on real and imaginary parts, a column x = [Re; Im] of
// 64 floats goes to M x with M = [[Ur, -Ui], [Ui, Ur]] (64x64). For 16 groups
// at once that is OUT(64x16) = M(64x64) * X(64x16): 4 output tiles of 16x16,
// each the sum of 16 v_mfma_f32_16x16x4f32 steps over K = 64.
//
// Operand placement (16x16x4 f32 MFMA, wave64):
//   A (16x4): lane l holds A[l%16][l/16]  -> M fragments, staged once per
//             workgroup in LDS as [row tile][row][k lane][step] so a lane
//             reads its 16 values of a row tile as four 16 B vectors;
//   B (4x16): lane l holds B[l/16][l%16]  -> lane l loads amplitude
//             (group l%16, combination 4s + l/16) for s = 0..7: its .x feeds
//             K-slice s (real rows), its .y K-slice s+8 (imaginary rows);
//   D (16x16): lane l holds D[4(l/16)+v][l%16], v = 0..3.
// Amplitudes go from HBM straight into MFMA operands and back. The block is
// indexed by ascending target (combination bit j = j-th lowest target), so a
// load's 4 K lanes run over the two lowest targets and its 16 columns over
// consecutive groups: whole cache lines whenever the low qubits are targets or
// groups. Each wave takes 4 consecutive 16-group chunks and loads the next
// chunk's amplitudes before multiplying the current one.
//
// Cost per amplitude: 256 FP32 MFMA flops for 16 B of HBM traffic, which on
// MI355X (157.3 TFLOP/s FP32 matrix, 8 TB/s HBM) keeps the block near the
// memory bound; every fused block replaces ~10 per-gate sweeps.
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

__global__ void __launch_bounds__(kBlock, 5) kFused5(Amp<float> *__restrict__ a, long long numChunks,
                                                  Fused5Spec spec, const float *__restrict__ M) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col = lane & 15, kq = lane >> 4;
    // A fragments staged in LDS as [rt][col][kq][s] so a lane's 16 values of a
    // row tile are contiguous
    __shared__ float sM[4 * 16 * 4 * 16];
    for (int e = threadIdx.x; e < 4096; e += kBlock) {
        const int rt = e >> 10, cc = (e >> 6) & 15, kk = (e >> 4) & 3, ss = e & 15;
        sM[e] = M[(16 * rt + cc) * 64 + 4 * ss + kk];
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
        f4 acc[4];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            float Af[1][16];
            const f4 *src = reinterpret_cast<const f4 *>(sM + (((rt * 16 + col) * 4 + kq) << 4));
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const f4 v = src[u];
                Af[0][4 * u] = v[0]; Af[0][4 * u + 1] = v[1]; Af[0][4 * u + 2] = v[2]; Af[0][4 * u + 3] = v[3];
            }
            acc[rt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 8; ++s)
                acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[0][s], x[s].x, acc[rt], 0, 0, 0);
#pragma unroll
            for (int s = 0; s < 8; ++s)
                acc[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[0][s + 8], x[s].y, acc[rt], 0, 0, 0);
        }
        // lane holds output rows 4 kq + v of each tile: combination c = 16 h + 4 kq + v,
        // real part from tile h, imaginary part from tile h + 2
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                stA<float, false>(a + ((base | outLane) + spec.off[16 * h + v]),
                                  Amp<float>{acc[h][v], acc[h + 2][v]});
            }
    }
}

struct Fused5Prepared {
    Fused5Spec spec;
    float *M;  // device copy of the 64x64 real block matrix
};

// The kernel's form of a 5-qubit block (targs[0] = least significant bit of
// u's index, u row-major 32x32): the spec, and the 64x64 real matrix
// [[Ur, -Ui], [Ui, Ur]] written to hm.
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
            hm[r * 64 + c] = (float)z.re;
            hm[r * 64 + 32 + c] = (float)-z.im;
            hm[(32 + r) * 64 + c] = (float)z.im;
            hm[(32 + r) * 64 + 32 + c] = (float)z.re;
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
    std::vector<float> hm(64 * 64);
    fused5Build(targs, u, p->spec, hm.data());
    p->M = static_cast<float *>(devAlloc(hm.size() * sizeof(float)));
    hostToDevice(p->M, hm.data(), hm.size() * sizeof(float));
    return p;
}

// One application of a block (multiQubitUnitary on an FP32 register): the
// matrix goes through the staging ring, with no allocation or synchronisation.
void fused5ApplyOnce(QuregDev *q, const int *targs, const Cplx *u) {
    constexpr size_t bytes = 64 * 64 * sizeof(float);
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
