// Batched swap tests for quantum k-means (QKMeans): thousands of registers of
// uneven size in one persistent launch.
//
// Reference construction (QuantumLab): a point v1 and a centroid v2 of length
// 2^q are amplitude-encoded into q-qubit registers by a tree of uniformly
// controlled Ry rotations (AmplitudeEncodeCuda.cu:61-150 angles,
// :498-533 ExchangeToYGate, FRGate.cu:154-252 FRy), and swap-tested on a
// (2q+1)-qubit register: H(anc); AE(v1) on qubits 1..q; AE(v2) on q+1..2q;
// q controlled-SWAPs (anc, i, i+q); H(anc) (SwapTest.cpp:70-123). P(anc = 0)
// gives the overlap.
//
// MI355X form:
//  * kAEState: the encoding tree applied to |0...0> for every vector at once,
//    one workgroup per vector: level-by-level angles theta = 2 atan2(right,
//    left) (the reference's recursion), then each amplitude as the product of
//    cos/sin(theta/2) along its path — the state the Ry tree prepares, in
//    O(2^q) instead of 2^q gates of O(2^(2q+1)) each.
//  * kSwapTiles: after H and the two encodings the register is |+>|a>|b>; the
//    CSWAP ladder and the final H leave ancilla-0 amplitudes
//    (a_x b_y + a_y b_x)/2. The kernel evaluates every one of those 2^(2q)
//    amplitudes (nothing of the 2^(2q+1)-amplitude register is stored) and
//    reduces |amp|^2 in FP64. Work is a static list of tiles (test, first
//    row, rows) over which a persistent grid strides — no atomics, every wave
//    exits when the list ends; the rows per tile are chosen per q by an
//    autotuner that times each candidate on the device.
//  * one k-means iteration = swap tiles -> per-test P0 -> argmax assignment ->
//    centroid means -> re-encoded centroids, captured once as a hipGraph and
//    replayed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>

#include "qmx_device.h"
#include "quest_mi355x.h"

namespace qmx {

namespace {

constexpr int kMinQ = 1, kMaxQ = 12;

struct Tile {
    int test;   // point * numCentroids + centroid
    int x0;     // first row
    int nx;     // rows
    int pad;
};

// ---- amplitude encoding -----------------------------------------------------
// vectors: packed real FP64 inputs (length 2^q each, offsets in off[]); out:
// the encoded states in the storage precision T (same packing). The encoding
// is O(2^q) per vector, so it is computed in FP64 for both storage precisions
// and each amplitude is rounded once when stored. One workgroup per vector;
// LDS holds the magnitudes and the cos/sin factors of the 2^q - 1 angles.
template <typename T>
__global__ void __launch_bounds__(kBlock) kAEState(const double *__restrict__ vecs, const long long *__restrict__ off,
                                                   const int *__restrict__ qs, T *__restrict__ out, int count) {
    extern __shared__ unsigned char smem[];
    const int v = blockIdx.x;
    if (v >= count) return;
    const int q = qs[v];
    const int D = 1 << q;
    double *mag = reinterpret_cast<double *>(smem);  // D values
    double *cs = mag + D;                            // D-1 cos(theta/2), level-major like the reference's list
    double *sn = cs + D;                             // D-1 sin(theta/2)
    const double *in = vecs + off[v];
    for (int i = threadIdx.x; i < D; i += kBlock) mag[i] = in[i];
    __syncthreads();
    // level l: pairs (left = k 2^(l+1), right = left + 2^l) of the running
    // magnitudes; angle index = D - 2^(q-l) + k (AmplitudeEncodeCuda.cu:126-129)
    for (int l = 0; l < q; ++l) {
        const int pairs = D >> (l + 1);
        const int base = D - (D >> l);
        for (int k = threadIdx.x; k < pairs; k += kBlock) {
            const int left = k << (l + 1), right = left + (1 << l);
            const double c = mag[left], s = mag[right];
            const double theta = 2.0 * atan2(s, c);
            cs[base + k] = cos(theta / 2.0);
            sn[base + k] = sin(theta / 2.0);
            if (pairs != 1) {
                // magnitudes at or below the arithmetic's epsilon are zeroed, as
                // the reference does with _QL_FLT_EPSILON (QLFloat.h:120,139: the
                // epsilon of the precision the angles are computed in)
                const double m2 = c * c + s * s;
                mag[left] = m2 > std::numeric_limits<double>::epsilon() ? sqrt(m2) : 0.0;
            }
        }
        __syncthreads();
    }
    // amplitude of basis state x: product over levels of cos (bit 0) / sin (bit 1)
    T *o = out + off[v];
    for (int x = threadIdx.x; x < D; x += kBlock) {
        double amp = 1.0;
        for (int l = 0; l < q; ++l) {
            const int idx = D - (D >> l) + (x >> (l + 1));
            amp *= ((x >> l) & 1) ? sn[idx] : cs[idx];
        }
        o[x] = (T)amp;
    }
}

// centroid registers at every size: cent (C x 2^maxQ, FP64) truncated to 2^q
// and L2-normalised into cin (FP64, packed per (q, c))
__global__ void __launch_bounds__(kBlock) kCentroidInputs(const double *__restrict__ cent, int maxQ, int C, int qLo,
                                                          int qHi, const long long *__restrict__ cOff,
                                                          double *__restrict__ cin) {
    const int q = qLo + blockIdx.x / C, c = blockIdx.x % C;
    if (q > qHi) return;
    const int D = 1 << q;
    const double *src = cent + (size_t)c * (1 << maxQ);
    Acc acc[1];
    for (int i = threadIdx.x; i < D; i += kBlock) acc[0].add(src[i] * src[i]);
    double tot[1];
    blockReduce<1>(acc, tot);
    __shared__ double norm;
    if (threadIdx.x == 0) norm = tot[0] > 0 ? 1.0 / sqrt(tot[0]) : 0.0;
    __syncthreads();
    double *dst = cin + cOff[blockIdx.x];
    for (int i = threadIdx.x; i < D; i += kBlock) dst[i] = src[i] * norm;
}

// ---- swap-test tiles ------------------------------------------------------------
// Tile (test, x0, nx): rows x in [x0, x0+nx), all columns y. A lane owns
// CPL = clamp(D/256, 1, 4) columns per pass (registers a_y, b_y); lanes split
// into RG = 256/TPR row groups of TPR = min(D/CPL, 256) lanes, each walking
// its rows with a_x, b_x read from LDS (uniform within the row group). Sums
// of 8 rows are FP32, accumulated in FP64 with compensation.
constexpr int kMaxTileRows = 1024;

template <typename T>
__global__ void __launch_bounds__(kBlock) kSwapTiles(const Tile *__restrict__ tiles, int numTiles,
                                                     const T *__restrict__ pointAE, const long long *__restrict__ pOff,
                                                     const int *__restrict__ pQ, const T *__restrict__ centAE,
                                                     const long long *__restrict__ cOff, int C, int qLo,
                                                     double *__restrict__ partial) {
    __shared__ T lx[2][kMaxTileRows];  // a_x, b_x of the tile's rows
    for (int w = blockIdx.x; w < numTiles; w += gridDim.x) {
        const Tile t = tiles[w];
        const int p = t.test / C, c = t.test % C;
        const int q = pQ[p];
        const int D = 1 << q;
        const T *a = pointAE + pOff[p];
        const T *b = centAE + cOff[(q - qLo) * C + c];
        __syncthreads();
        for (int i = threadIdx.x; i < t.nx; i += kBlock) {
            lx[0][i] = a[t.x0 + i];
            lx[1][i] = b[t.x0 + i];
        }
        __syncthreads();
        const int CPL = D >= 4 * kBlock ? 4 : (D >= kBlock ? D / kBlock : 1);
        const int TPR = D / CPL < kBlock ? D / CPL : kBlock;
        const int RG = kBlock / TPR;
        const int col0 = threadIdx.x % TPR, rg = threadIdx.x / TPR;
        Acc acc[1];
        for (int y0 = col0; y0 < D; y0 += CPL * TPR) {
            T ay[4], by[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int y = y0 + k * TPR;
                const bool ok = k < CPL && y < D;
                ay[k] = ok ? a[y] : (T)0;
                by[k] = ok ? b[y] : (T)0;
            }
            T s8 = (T)0;
            int inChunk = 0;
            for (int r = rg; r < t.nx; r += RG) {
                const T ax = lx[0][r], bx = lx[1][r];
                T v = ax * by[0] + ay[0] * bx;
                T s = v * v;
                if (CPL > 1) {
#pragma unroll
                    for (int k = 1; k < 4; ++k) {
                        if (k < CPL) {
                            v = ax * by[k] + ay[k] * bx;
                            s += v * v;
                        }
                    }
                }
                s8 += s;
                if (++inChunk == 8) {
                    acc[0].add((double)s8);
                    s8 = (T)0;
                    inChunk = 0;
                }
            }
            acc[0].add((double)s8);
        }
        double tot[1];
        blockReduce<1>(acc, tot);
        if (threadIdx.x == 0) partial[w] = tot[0];
    }
}

// FP32 plans: the same sums on the matrix cores, one wave per tile. A tile's
// rows x0 .. x0+nx (nx a multiple of 32, or the whole register when D < 32)
// against all columns, in 32 x 32 blocks of M[x][y] = a_x b_y + b_x a_y (twice
// the ancilla-0 amplitude) from one v_mfma_f32_32x32x2f32 each:
//   A (32 x 2): lane l holds A[l%32][l/32] = (l < 32 ? a : b)[x0 + l%32]
//   B (2 x 32): lane l holds B[l/32][l%32] = (l < 32 ? b : a)[y0 + l%32]
// so K = 2 is exactly the two products (and M is symmetric, so only blocks on
// or above the diagonal are formed). Each of the 16 accumulator values per
// lane is one element of the block; their squares are summed in FP32 per column
// block (1, 2, 4 or 8 row blocks share one B operand) and in FP64 across blocks.
constexpr int kMaxRowBlocks = 8;  // 256 rows per tile
typedef float floatx16 __attribute__((ext_vector_type(16)));

// one tile of NB row blocks. M is symmetric (M[x][y] = M[y][x]), so the sum
// over all blocks is the diagonal blocks once plus the blocks above the
// diagonal twice: row block X only meets column blocks Y >= X.
template <int NB>
__device__ __forceinline__ double swapTileMfma(const float *sx, const float *sy, int x0, int D, int r) {
    const floatx16 zero = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float av[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int x = x0 + 32 * k + r;
        av[k] = x < D ? sx[x] : 0.f;
    }
    double acc = 0.0;
    float bv = x0 + r < D ? sy[x0 + r] : 0.f;
    for (int y0 = x0; y0 < D; y0 += 32) {
        const int yn = y0 + 32 + r;
        const float bnext = yn < D ? sy[yn] : 0.f;  // next column block, loaded ahead
        float s = 0.f;
        if (y0 >= x0 + 32 * (NB - 1)) {
            // every row block of the tile is at or left of this column block
            floatx16 d[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) d[k] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[k], bv, zero, 0, 0, 0);
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                float sk = 0.f;
#pragma unroll
                for (int v = 0; v < 16; ++v) sk = fmaf(d[k][v], d[k][v], sk);
                s += (y0 == x0 + 32 * k) ? sk : 2.f * sk;
            }
        } else {
            // the tile's first column blocks: only row blocks X <= Y
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                if (y0 >= x0 + 32 * k) {
                    const floatx16 d = __builtin_amdgcn_mfma_f32_32x32x2f32(av[k], bv, zero, 0, 0, 0);
                    float sk = 0.f;
#pragma unroll
                    for (int v = 0; v < 16; ++v) sk = fmaf(d[v], d[v], sk);
                    s += (y0 == x0 + 32 * k) ? sk : 2.f * sk;
                }
            }
        }
        acc += (double)s;
        bv = bnext;
    }
    return acc;
}

__global__ void __launch_bounds__(kBlock) kSwapTilesMfma(const Tile *__restrict__ tiles, int numTiles,
                                                         const float *__restrict__ pointAE,
                                                         const long long *__restrict__ pOff, const int *__restrict__ pQ,
                                                         const float *__restrict__ centAE,
                                                         const long long *__restrict__ cOff, int C, int qLo,
                                                         double *__restrict__ partial) {
    const int lane = threadIdx.x & 63, r = lane & 31, hi = lane >> 5;
    const int wavesTotal = gridDim.x * (kBlock / 64);
    for (int w = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); w < numTiles; w += wavesTotal) {
        const Tile t = tiles[w];
        const int p = t.test / C, c = t.test % C;
        const int q = pQ[p];
        const int D = 1 << q;
        const float *a = pointAE + pOff[p];
        const float *b = centAE + cOff[(q - qLo) * C + c];
        const float *sx = hi ? b : a;  // rows
        const float *sy = hi ? a : b;  // columns
        const int nb = (t.nx + 31) >> 5;
        double acc;
        switch (nb) {  // wave-uniform
            case 1: acc = swapTileMfma<1>(sx, sy, t.x0, D, r); break;
            case 2: acc = swapTileMfma<2>(sx, sy, t.x0, D, r); break;
            case 4: acc = swapTileMfma<4>(sx, sy, t.x0, D, r); break;
            default: acc = swapTileMfma<8>(sx, sy, t.x0, D, r); break;
        }
        for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) partial[w] = acc;
    }
}

// P0 of each test: (1/4) sum of its tiles' partials (tiles of a test are contiguous)
__global__ void __launch_bounds__(kBlock) kSwapFinish(const double *__restrict__ partial, const int *__restrict__ tStart,
                                                      const int *__restrict__ tCount, int numTests, double *__restrict__ p0) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= numTests) return;
    Acc acc;
    const int s = tStart[t], n = tCount[t];
    for (int i = 0; i < n; ++i) acc.add(partial[s + i]);
    p0[t] = 0.25 * acc.value();
}

// assignment: the centroid of largest P0 (= largest overlap); ties -> lowest index
__global__ void __launch_bounds__(kBlock) kAssign(const double *__restrict__ p0, int P, int C, int *__restrict__ assign) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P) return;
    int best = 0;
    double bv = p0[(size_t)p * C];
    for (int c = 1; c < C; ++c) {
        const double v = p0[(size_t)p * C + c];
        if (v > bv) {
            bv = v;
            best = c;
        }
    }
    assign[p] = best;
}

// centroid means over assigned points, coordinate by coordinate (points only
// contribute their own 2^q coordinates); empty clusters keep their centroid.
// Stage 1: one workgroup per (centroid, 256-coordinate slice, chunk of points)
// writes the chunk's sum and count; stage 2 adds the chunks in order. FP64,
// deterministic.
constexpr int kMeanChunks = 16;

__global__ void __launch_bounds__(kBlock) kCentroidPartial(const double *__restrict__ pointsV,
                                                           const long long *__restrict__ pOff,
                                                           const int *__restrict__ pQ, const int *__restrict__ assign,
                                                           int P, int C, int maxQ, double *__restrict__ sums,
                                                           int *__restrict__ counts, int *__restrict__ members) {
    const int dim = 1 << maxQ;
    const int slices = dim / kBlock > 0 ? dim / kBlock : 1;
    const int chunk = blockIdx.x % kMeanChunks;
    const int cs = blockIdx.x / kMeanChunks;
    const int c = cs / slices;
    const int j = (cs % slices) * kBlock + threadIdx.x;
    const int per = (P + kMeanChunks - 1) / kMeanChunks;
    const int p0 = chunk * per, p1 = p0 + per < P ? p0 + per : P;
    if (c >= C) return;
    double sum = 0.0;
    int cnt = 0, mem = 0;
    for (int p = p0; p < p1; ++p) {
        if (assign[p] != c) continue;
        ++mem;
        if (j < (1 << pQ[p])) {
            sum += pointsV[pOff[p] + j];
            ++cnt;
        }
    }
    if (j < dim) {
        const size_t o = ((size_t)chunk * C + c) * dim + j;
        sums[o] = sum;
        counts[o] = cnt;
    }
    if (threadIdx.x == 0 && (cs % slices) == 0) members[chunk * C + c] = mem;
}

__global__ void __launch_bounds__(kBlock) kCentroidCombine(const double *__restrict__ sums, const int *__restrict__ counts,
                                                           const int *__restrict__ members, int C, int maxQ,
                                                           double *__restrict__ cent) {
    const int dim = 1 << maxQ;
    const int slices = dim / kBlock > 0 ? dim / kBlock : 1;
    const int c = blockIdx.x / slices;
    const int j = (blockIdx.x % slices) * kBlock + threadIdx.x;
    if (c >= C || j >= dim) return;
    double sum = 0.0;
    int cnt = 0, mem = 0;
    for (int k = 0; k < kMeanChunks; ++k) {
        const size_t o = ((size_t)k * C + c) * dim + j;
        sum += sums[o];
        cnt += counts[o];
        mem += members[k * C + c];
    }
    if (mem > 0 && cnt > 0) cent[(size_t)c * dim + j] = sum / cnt;
}

}  // namespace

// ---- the plan ------------------------------------------------------------------

struct SwapPlanImpl {
    int prec;
    int P, C, maxQ, qLo, qHi;
    std::vector<int> hQ;
    std::vector<long long> hPOff, hCOff;
    void *pointsV = nullptr, *pointAE = nullptr, *cent = nullptr, *centIn = nullptr, *centAE = nullptr;
    long long *pOff = nullptr, *cOff = nullptr;
    int *pQ = nullptr, *cQ = nullptr, *assign = nullptr, *tStart = nullptr, *tCount = nullptr;
    Tile *tiles = nullptr;
    double *partial = nullptr, *p0 = nullptr;
    double *meanSums = nullptr;
    int *meanCounts = nullptr, *meanMembers = nullptr;
    int numTiles = 0;
    int rowsPerTile[kMaxQ + 1] = {0};
    int grid = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    size_t aeSmem = 0;
};

namespace {

// the swap-test sums of every tile: matrix cores for FP32 registers, FP64 on the VALU
template <typename T>
void launchSwapTiles(SwapPlanImpl *s) {
    if constexpr (std::is_same<T, float>::value)
        hipLaunchKernelGGL(kSwapTilesMfma, dim3(s->grid), dim3(kBlock), 0, hstream(), s->tiles, s->numTiles,
                           (const float *)s->pointAE, s->pOff, s->pQ, (const float *)s->centAE, s->cOff, s->C, s->qLo,
                           s->partial);
    else
        hipLaunchKernelGGL(kSwapTiles<T>, dim3(s->grid), dim3(kBlock), 0, hstream(), s->tiles, s->numTiles,
                           (const T *)s->pointAE, s->pOff, s->pQ, (const T *)s->centAE, s->cOff, s->C, s->qLo,
                           s->partial);
}

template <typename T>
void launchIteration(SwapPlanImpl *s) {
    const int numTests = s->P * s->C;
    const int numCent = (s->qHi - s->qLo + 1) * s->C;
    launchSwapTiles<T>(s);
    hipLaunchKernelGGL(kSwapFinish, dim3((numTests + kBlock - 1) / kBlock), dim3(kBlock), 0, hstream(), s->partial,
                       s->tStart, s->tCount, numTests, s->p0);
    hipLaunchKernelGGL(kAssign, dim3((s->P + kBlock - 1) / kBlock), dim3(kBlock), 0, hstream(), s->p0, s->P, s->C,
                       s->assign);
    const int slices = std::max(1, (1 << s->maxQ) / kBlock);
    hipLaunchKernelGGL(kCentroidPartial, dim3(s->C * slices * kMeanChunks), dim3(kBlock), 0, hstream(),
                       (const double *)s->pointsV, s->pOff, s->pQ, s->assign, s->P, s->C, s->maxQ, s->meanSums,
                       s->meanCounts, s->meanMembers);
    hipLaunchKernelGGL(kCentroidCombine, dim3(s->C * slices), dim3(kBlock), 0, hstream(), s->meanSums, s->meanCounts,
                       s->meanMembers, s->C, s->maxQ, (double *)s->cent);
    hipLaunchKernelGGL(kCentroidInputs, dim3(numCent), dim3(kBlock), 0, hstream(), (const double *)s->cent, s->maxQ,
                       s->C, s->qLo, s->qHi, s->cOff, (double *)s->centIn);
    hipLaunchKernelGGL(kAEState<T>, dim3(numCent), dim3(kBlock), s->aeSmem, hstream(), (const double *)s->centIn,
                       s->cOff, s->cQ, (T *)s->centAE, numCent);
}

// encode the current centroids at every size (used at creation)
template <typename T>
void encodeCentroids(SwapPlanImpl *s) {
    const int numCent = (s->qHi - s->qLo + 1) * s->C;
    hipLaunchKernelGGL(kCentroidInputs, dim3(numCent), dim3(kBlock), 0, hstream(), (const double *)s->cent, s->maxQ,
                       s->C, s->qLo, s->qHi, s->cOff, (double *)s->centIn);
    hipLaunchKernelGGL(kAEState<T>, dim3(numCent), dim3(kBlock), s->aeSmem, hstream(), (const double *)s->centIn,
                       s->cOff, s->cQ, (T *)s->centAE, numCent);
}

// build the tile list for the current rows-per-tile choice (per q)
void buildTiles(SwapPlanImpl *s) {
    std::vector<Tile> tl;
    std::vector<int> start(s->P * s->C), count(s->P * s->C);
    // order: largest registers first so the persistent grid's tail is small tiles
    std::vector<int> order(s->P);
    for (int p = 0; p < s->P; ++p) order[p] = p;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return s->hQ[x] > s->hQ[y]; });
    for (int p : order) {
        const int D = 1 << s->hQ[p];
        const int rows = std::min(std::min(D, s->prec == FP32 ? 32 * kMaxRowBlocks : kMaxTileRows),
                                  s->rowsPerTile[s->hQ[p]]);
        for (int c = 0; c < s->C; ++c) {
            const int test = p * s->C + c;
            start[test] = (int)tl.size();
            for (int x0 = 0; x0 < D; x0 += rows) tl.push_back(Tile{test, x0, std::min(rows, D - x0), 0});
            count[test] = (int)tl.size() - start[test];
        }
    }
    if (s->tiles) devFree(s->tiles);
    if (s->partial) devFree(s->partial);
    s->numTiles = (int)tl.size();
    s->tiles = static_cast<Tile *>(devAlloc(tl.size() * sizeof(Tile)));
    s->partial = static_cast<double *>(devAlloc(tl.size() * sizeof(double)));
    hostToDevice(s->tiles, tl.data(), tl.size() * sizeof(Tile));
    hostToDevice(s->tStart, start.data(), start.size() * sizeof(int));
    hostToDevice(s->tCount, count.data(), count.size() * sizeof(int));
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess) cus = prop.multiProcessorCount;
    // FP32: one tile per wave (4 per workgroup); FP64: one tile per workgroup
    const int perGroup = s->prec == FP32 ? kBlock / 64 : 1;
    s->grid = std::max(1, std::min((s->numTiles + perGroup - 1) / perGroup, cus * 8));
}

void dropGraph(SwapPlanImpl *s) {
    if (s->graphExec) QMX_HIP_CHECK(hipGraphExecDestroy(s->graphExec));
    if (s->graph) QMX_HIP_CHECK(hipGraphDestroy(s->graph));
    s->graphExec = nullptr;
    s->graph = nullptr;
}

}  // namespace

}  // namespace qmx

using qmx::SwapPlanImpl;

struct qmxSwapPlan {
    SwapPlanImpl impl;
};

namespace qmx {
namespace {
template <typename T>
__global__ void __launch_bounds__(kBlock) kAEWrite(Amp<T> *__restrict__ a, const double *__restrict__ v, int D) {
    const int x = blockIdx.x * kBlock + threadIdx.x;
    if (x < D) a[x] = Amp<T>{(T)v[x], (T)0};
}

// AmplitudeEncodeOneVector without its global phase gate (AmplitudeEncodeCuda.cu:
// 691-724, angles :20-30 and :61-109, circuit :452-496 with FRyz): the Y tree
// of the real encoder over |v_i|, plus a phase tree over arg v_i (Z = left
// phase - right phase, running phase = their mean). Each level's FRyz applies
// Ry(Y) and then Rz(-Z) uniformly controlled by the higher qubits, so basis
// state x ends with the product of cos / sin(Y/2) along its path and the phase
// sum of +Z/2 (bit clear) / -Z/2 (bit set). One workgroup; LDS holds the
// running magnitudes and phases and the Y and Z lists (4 x 2^q doubles).
__global__ void __launch_bounds__(kBlock) kAEComplex(const double *__restrict__ vre, const double *__restrict__ vim,
                                                     int q, double *__restrict__ ore, double *__restrict__ oim) {
    extern __shared__ unsigned char smem[];
    const int D = 1 << q;
    double *mag = reinterpret_cast<double *>(smem), *ph = mag + D, *Y = ph + D, *Z = Y + D;
    for (int i = threadIdx.x; i < D; i += kBlock) {
        mag[i] = hypot(vre[i], vim[i]);
        ph[i] = atan2(vim[i], vre[i]);
    }
    __syncthreads();
    for (int l = 0; l < q; ++l) {
        const int pairs = D >> (l + 1);
        const int base = D - (D >> l);
        for (int k = threadIdx.x; k < pairs; k += kBlock) {
            const int left = k << (l + 1), right = left + (1 << l);
            const double c = mag[left], s = mag[right];
            Y[base + k] = 2.0 * atan2(s, c);
            if (pairs != 1) {
                const double m2 = c * c + s * s;
                mag[left] = m2 > std::numeric_limits<double>::epsilon() ? sqrt(m2) : 0.0;
            }
            Z[base + k] = ph[left] - ph[right];
            ph[left] = 0.5 * (ph[left] + ph[right]);
        }
        __syncthreads();
    }
    for (int x = threadIdx.x; x < D; x += kBlock) {
        double amp = 1.0, phase = 0.0;
        for (int l = 0; l < q; ++l) {
            const int idx = D - (D >> l) + (x >> (l + 1));
            const bool bit = (x >> l) & 1;
            amp *= bit ? sin(0.5 * Y[idx]) : cos(0.5 * Y[idx]);
            phase += bit ? -0.5 * Z[idx] : 0.5 * Z[idx];
        }
        ore[x] = amp * cos(phase);
        oim[x] = amp * sin(phase);
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) kAEWriteComplex(Amp<T> *__restrict__ a, const double *__restrict__ re,
                                                          const double *__restrict__ im, int D) {
    const int x = blockIdx.x * kBlock + threadIdx.x;
    if (x < D) a[x] = Amp<T>{(T)re[x], (T)im[x]};
}
}  // namespace

// qubits 0..q-1 take the encoded complex vector, the others stay |0>
void amplitudeEncodeComplex(QuregDev *q, const double *re, const double *im, int qbits) {
    const int D = 1 << qbits;
    const size_t smem = (size_t)4 * D * sizeof(double);
    if (smem > 65536 && !allowDynamicLds(reinterpret_cast<const void *>(kAEComplex), smem,
                                         "qmxAmplitudeEncodeComplex"))
        return;
    double *buf = static_cast<double *>(devAlloc((size_t)4 * D * sizeof(double)));
    hostToDevice(buf, re, (size_t)D * sizeof(double));
    hostToDevice(buf + D, im, (size_t)D * sizeof(double));
    hipLaunchKernelGGL(kAEComplex, dim3(1), dim3(kBlock), smem, hstream(), buf, buf + D, qbits, buf + 2 * D,
                       buf + 3 * D);
    initBlank(q);
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kAEWriteComplex<T>, dim3(gridFor(D, 1)), dim3(kBlock), 0, hstream(), (Amp<T> *)q->amps,
                           buf + 2 * D, buf + 3 * D, D);
    });
    checkLaunch("qmxAmplitudeEncodeComplex");
    devFree(buf);
}

// the reference's amplitude-encoding circuit (AmplitudeEncodeCuda.cu:112-150 angles,
// Ry + uniformly controlled Ry ladder, FRGate.cu:181-252) applied to |0...0>:
// qubits 0..q-1 take the encoded vector, the others stay |0>
void amplitudeEncode(QuregDev *q, const double *vec, int qbits) {
    const int D = 1 << qbits;
    const size_t smem = (size_t)3 * D * sizeof(double);
    if (smem > 65536 && !allowDynamicLds(reinterpret_cast<const void *>(kAEState<double>), smem,
                                         "qmxAmplitudeEncodeReal"))
        return;
    struct Hdr {
        long long off;
        int q;
    };
    char *buf = static_cast<char *>(devAlloc((size_t)2 * D * sizeof(double) + 64));
    double *dIn = reinterpret_cast<double *>(buf), *dOut = dIn + D;
    long long *dOff = reinterpret_cast<long long *>(dOut + D);
    int *dQ = reinterpret_cast<int *>(dOff + 1);
    const Hdr h{0, qbits};
    hostToDevice(dIn, vec, (size_t)D * sizeof(double));
    hostToDevice(dOff, &h.off, sizeof(long long));
    hostToDevice(dQ, &h.q, sizeof(int));
    hipLaunchKernelGGL(kAEState<double>, dim3(1), dim3(kBlock), smem, hstream(), dIn, dOff, dQ, dOut, 1);
    initBlank(q);
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kAEWrite<T>, dim3(gridFor(D, 1)), dim3(kBlock), 0, hstream(), (Amp<T> *)q->amps, dOut, D);
    });
    checkLaunch("qmxAmplitudeEncodeReal");
    devFree(buf);
}
}  // namespace qmx

extern "C" {

qmxSwapPlan *qmxSwapPlanCreate(int numPoints, const int *pointQubits, const double *points, int numCentroids,
                               int maxQubits, const double *centroids, int precision) {
    using namespace qmx;
    runtimeInit();
    if (numPoints < 1 || numCentroids < 1 || maxQubits < kMinQ || maxQubits > kMaxQ || (precision != 0 && precision != 1)) {
        std::fprintf(stderr, "libquest_mi355x: qmxSwapPlanCreate: invalid sizes or precision\n");
        return nullptr;
    }
    auto *plan = new qmxSwapPlan;
    SwapPlanImpl *s = &plan->impl;
    s->prec = precision;
    s->P = numPoints;
    s->C = numCentroids;
    s->maxQ = maxQubits;
    s->qLo = maxQubits;
    s->qHi = 1;
    s->hQ.assign(pointQubits, pointQubits + numPoints);
    for (int q : s->hQ) {
        if (q < kMinQ || q > maxQubits) {
            std::fprintf(stderr, "libquest_mi355x: qmxSwapPlanCreate: point register size %d outside [1, %d]\n", q,
                         maxQubits);
            delete plan;
            return nullptr;
        }
        s->qLo = std::min(s->qLo, q);
        s->qHi = std::max(s->qHi, q);
    }
    s->hPOff.resize(numPoints);
    long long tot = 0;
    for (int p = 0; p < numPoints; ++p) {
        s->hPOff[p] = tot;
        tot += 1ll << s->hQ[p];
    }
    const int nq = s->qHi - s->qLo + 1;
    s->hCOff.resize((size_t)nq * numCentroids);
    std::vector<int> cq((size_t)nq * numCentroids);
    long long ctot = 0;
    for (int qi = 0; qi < nq; ++qi)
        for (int c = 0; c < numCentroids; ++c) {
            s->hCOff[(size_t)qi * numCentroids + c] = ctot;
            cq[(size_t)qi * numCentroids + c] = s->qLo + qi;
            ctot += 1ll << (s->qLo + qi);
        }
    // encoded registers in the storage precision; inputs, centroids and the
    // encoding arithmetic in FP64 (O(2^q) per vector, negligible next to the
    // O(2^2q) swap-test evaluation)
    const size_t eb = precision == FP64 ? 8 : 4;
    std::vector<double> pts(points, points + tot);
    for (int p = 0; p < numPoints; ++p) {
        double n2 = 0;
        const long long D = 1ll << s->hQ[p];
        for (long long i = 0; i < D; ++i) n2 += pts[s->hPOff[p] + i] * pts[s->hPOff[p] + i];
        const double f = n2 > 0 ? 1.0 / std::sqrt(n2) : 0.0;
        for (long long i = 0; i < D; ++i) pts[s->hPOff[p] + i] *= f;
    }
    s->pointsV = devAlloc((size_t)tot * 8);
    hostToDevice(s->pointsV, pts.data(), (size_t)tot * 8);
    s->pointAE = devAlloc((size_t)tot * eb);
    const size_t centCount = (size_t)numCentroids * (1ull << maxQubits);
    s->cent = devAlloc(centCount * 8);
    hostToDevice(s->cent, centroids, centCount * 8);
    s->centIn = devAlloc((size_t)ctot * 8);
    s->centAE = devAlloc((size_t)ctot * eb);
    s->pOff = static_cast<long long *>(devAlloc(numPoints * sizeof(long long)));
    s->cOff = static_cast<long long *>(devAlloc(s->hCOff.size() * sizeof(long long)));
    s->pQ = static_cast<int *>(devAlloc(numPoints * sizeof(int)));
    s->cQ = static_cast<int *>(devAlloc(cq.size() * sizeof(int)));
    s->assign = static_cast<int *>(devAlloc(numPoints * sizeof(int)));
    s->tStart = static_cast<int *>(devAlloc((size_t)numPoints * numCentroids * sizeof(int)));
    s->tCount = static_cast<int *>(devAlloc((size_t)numPoints * numCentroids * sizeof(int)));
    s->p0 = static_cast<double *>(devAlloc((size_t)numPoints * numCentroids * sizeof(double)));
    const size_t meanSlots = (size_t)kMeanChunks * numCentroids * (1ull << maxQubits);
    s->meanSums = static_cast<double *>(devAlloc(meanSlots * sizeof(double)));
    s->meanCounts = static_cast<int *>(devAlloc(meanSlots * sizeof(int)));
    s->meanMembers = static_cast<int *>(devAlloc((size_t)kMeanChunks * numCentroids * sizeof(int)));
    hostToDevice(s->pOff, s->hPOff.data(), numPoints * sizeof(long long));
    hostToDevice(s->cOff, s->hCOff.data(), s->hCOff.size() * sizeof(long long));
    hostToDevice(s->pQ, s->hQ.data(), numPoints * sizeof(int));
    hostToDevice(s->cQ, cq.data(), cq.size() * sizeof(int));
    s->aeSmem = (size_t)3 * (1u << s->qHi) * sizeof(double);
    // 11- and 12-qubit registers need more than the 64 KiB default of dynamic LDS
    bool ldsOk = true;
    QMX_DISPATCH(precision, T, {
        if (s->aeSmem > 65536)
            ldsOk = allowDynamicLds(reinterpret_cast<const void *>(kAEState<T>), s->aeSmem, "qmxSwapPlanCreate");
    });
    if (!ldsOk) {
        qmxSwapPlanDestroy(plan);
        return nullptr;
    }
    // default tile rows: FP32 (one wave per tile) 64 rows, or the register
    // when smaller; FP64 (one workgroup per tile) ~16K amplitudes a tile
    for (int q = 0; q <= kMaxQ; ++q)
        s->rowsPerTile[q] = precision == FP32 ? std::min(1 << q, 64)
                                              : std::max(1, std::min(16384 >> q, std::min(1 << q, kMaxTileRows)));
    QMX_DISPATCH(precision, T, {
        hipLaunchKernelGGL(kAEState<T>, dim3(numPoints), dim3(kBlock), s->aeSmem, hstream(),
                           (const double *)s->pointsV, s->pOff, s->pQ, (T *)s->pointAE, numPoints);
        encodeCentroids<T>(s);
    });
    checkLaunch("qmxSwapPlanCreate");
    buildTiles(s);
    deviceSync();
    return plan;
}

void qmxSwapPlanDestroy(qmxSwapPlan *plan) {
    using namespace qmx;
    if (!plan) return;
    SwapPlanImpl *s = &plan->impl;
    dropGraph(s);
    for (void *p : {s->pointsV, s->pointAE, s->cent, s->centIn, s->centAE, (void *)s->pOff, (void *)s->cOff,
                    (void *)s->pQ, (void *)s->cQ, (void *)s->assign, (void *)s->tStart, (void *)s->tCount,
                    (void *)s->tiles, (void *)s->partial, (void *)s->p0, (void *)s->meanSums,
                    (void *)s->meanCounts, (void *)s->meanMembers})
        devFree(p);
    delete plan;
}

// time candidate rows-per-tile for each register size on the device and keep
// the fastest; returns the number of (size, candidate) timings made
int qmxSwapPlanAutotune(qmxSwapPlan *plan) {
    using namespace qmx;
    SwapPlanImpl *s = &plan->impl;
    dropGraph(s);
    // FP32 tiles are row blocks of 32 on the matrix cores; FP64 tiles any rows
    const std::vector<int> cands = s->prec == FP32 ? std::vector<int>{32, 64, 128, 256}
                                                   : std::vector<int>{1, 4, 16, 64, 256, 1024, 4096};
    int timings = 0;
    for (int q = s->qLo; q <= s->qHi; ++q) {
        bool present = false;
        for (int x : s->hQ) present |= (x == q);
        if (!present) continue;
        const int D = 1 << q;
        double best = 1e30;
        int bestRows = s->rowsPerTile[q];
        int lastRows = -1;
        for (int rowsCand : cands) {
            int rows = rowsCand;
            if (s->prec == FP32) {
                if (D < 32) rows = D;           // a register smaller than one block: one tile
                else if (rows > D) continue;
            } else if (rows > D || rows > kMaxTileRows || (long long)rows * D > (1ll << 22) ||
                       ((long long)rows * D < 256 && rows != D)) {
                continue;
            }
            if (rows == lastRows) continue;
            lastRows = rows;
            // retime the whole tile list with this size's rows changed
            int saved[kMaxQ + 1];
            std::memcpy(saved, s->rowsPerTile, sizeof saved);
            s->rowsPerTile[q] = rows;
            buildTiles(s);
            timerBegin();
            QMX_DISPATCH(s->prec, T, { launchSwapTiles<T>(s); });
            checkLaunch("autotune");
            const double ms = timerEndMs();
            ++timings;
            if (ms < best) {
                best = ms;
                bestRows = rows;
            }
            std::memcpy(s->rowsPerTile, saved, sizeof saved);
        }
        s->rowsPerTile[q] = bestRows;
    }
    buildTiles(s);
    deviceSync();
    return timings;
}

int qmxSwapPlanRowsPerTile(const qmxSwapPlan *plan, int q) {
    return (q >= 0 && q <= qmx::kMaxQ) ? plan->impl.rowsPerTile[q] : 0;
}

int qmxSwapPlanNumTiles(const qmxSwapPlan *plan) { return plan->impl.numTiles; }

// 32 x 32 amplitude blocks one iteration forms on the matrix cores (FP32
// plans: blocks on or above the diagonal of every test, padded to a block for
// registers under 32 amplitudes); 0 for FP64 plans
long long qmxSwapPlanMfmaBlocks(const qmxSwapPlan *plan) {
    using namespace qmx;
    const SwapPlanImpl *s = &plan->impl;
    if (s->prec != FP32) return 0;
    long long blocks = 0;
    for (int p = 0; p < s->P; ++p) {
        const int D = 1 << s->hQ[p];
        const int rows = std::min(std::min(D, 32 * kMaxRowBlocks), s->rowsPerTile[s->hQ[p]]);
        const int nyb = (D + 31) / 32;
        for (int x0 = 0; x0 < D; x0 += rows) {
            const int xb0 = x0 / 32, nb = (std::min(rows, D - x0) + 31) / 32;
            for (int k = 0; k < nb; ++k) blocks += nyb - (xb0 + k);
        }
    }
    return blocks * s->C;
}

// device time of one launch of the swap-test tiles (ms, HIP events on the
// library stream), averaged over `launches`
double qmxSwapPlanTimeTiles(qmxSwapPlan *plan, int launches) {
    using namespace qmx;
    SwapPlanImpl *s = &plan->impl;
    if (launches < 1) return 0.0;
    deviceSync();
    timerBegin();
    for (int i = 0; i < launches; ++i) QMX_DISPATCH(s->prec, T, { launchSwapTiles<T>(s); });
    checkLaunch("qmxSwapPlanTimeTiles");
    return timerEndMs() / launches;
}

// run `iterations` k-means iterations; useGraph != 0 replays one captured
// iteration graph (captured on first use)
void qmxSwapPlanRun(qmxSwapPlan *plan, int iterations, int useGraph) {
    using namespace qmx;
    SwapPlanImpl *s = &plan->impl;
    if (useGraph) {
        if (!s->graphExec) {
            QMX_HIP_CHECK(hipStreamBeginCapture(hstream(), hipStreamCaptureModeThreadLocal));
            QMX_DISPATCH(s->prec, T, { launchIteration<T>(s); });
            QMX_HIP_CHECK(hipStreamEndCapture(hstream(), &s->graph));
            QMX_HIP_CHECK(hipGraphInstantiate(&s->graphExec, s->graph, nullptr, nullptr, 0));
        }
        for (int i = 0; i < iterations; ++i) QMX_HIP_CHECK(hipGraphLaunch(s->graphExec, hstream()));
    } else {
        for (int i = 0; i < iterations; ++i) {
            QMX_DISPATCH(s->prec, T, { launchIteration<T>(s); });
            checkLaunch("qmxSwapPlanRun");
        }
    }
}

// P(ancilla = 0) of every test of the last iteration, [point][centroid]
void qmxSwapPlanP0(qmxSwapPlan *plan, double *out) {
    qmx::deviceToHost(out, plan->impl.p0, (size_t)plan->impl.P * plan->impl.C * sizeof(double));
}

void qmxSwapPlanAssignments(qmxSwapPlan *plan, int *out) {
    qmx::deviceToHost(out, plan->impl.assign, (size_t)plan->impl.P * sizeof(int));
}

// current centroids (numCentroids x 2^maxQubits), as doubles
void qmxSwapPlanCentroids(qmxSwapPlan *plan, double *out) {
    using namespace qmx;
    SwapPlanImpl *s = &plan->impl;
    deviceToHost(out, s->cent, (size_t)s->C * (1ull << s->maxQ) * sizeof(double));
}

// replace the centroids (numCentroids x 2^maxQubits doubles) and re-encode
// them; the point registers, tiles, autotuned shapes and a captured graph stay
void qmxSwapPlanSetCentroids(qmxSwapPlan *plan, const double *centroids) {
    using namespace qmx;
    SwapPlanImpl *s = &plan->impl;
    deviceSync();
    hostToDevice(s->cent, centroids, (size_t)s->C * (1ull << s->maxQ) * sizeof(double));
    QMX_DISPATCH(s->prec, T, { encodeCentroids<T>(s); });
    checkLaunch("qmxSwapPlanSetCentroids");
    deviceSync();
}

// encoded state of point p (its 2^q amplitudes), for checking the encoder
void qmxSwapPlanPointState(qmxSwapPlan *plan, int p, double *out) {
    using namespace qmx;
    SwapPlanImpl *s = &plan->impl;
    const size_t D = 1ull << s->hQ[p];
    if (s->prec == FP64) {
        deviceToHost(out, (double *)s->pointAE + s->hPOff[p], D * 8);
    } else {
        std::vector<float> f(D);
        deviceToHost(f.data(), (float *)s->pointAE + s->hPOff[p], D * 4);
        for (size_t i = 0; i < D; ++i) out[i] = f[i];
    }
}

}  // extern "C"
