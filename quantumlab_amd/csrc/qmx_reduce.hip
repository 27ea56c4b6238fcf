// Reductions. Every sum is accumulated in FP64 (also for FP32 registers) with
// Neumaier compensation per lane, a wave64 butterfly, an LDS step across the
// four waves, and one second-stage workgroup over the per-block partials, all
// on the device: no host round trip per tree level and no full-state copy to
// the host (QuEST_gpu.cu:1649-1679 copies the whole state for calcTotalProb).
// The grid size is a fixed function of the problem size, so results are
// bit-reproducible run to run.
#include <algorithm>

#include "qmx_device.h"

namespace qmx {

namespace {

constexpr int kMaxPartials = 2048;
double *g_partials = nullptr;  // kMaxPartials * 4 doubles
double *g_result = nullptr;    // 4 doubles

void ensureBuffers() {
    if (!g_partials) {
        g_partials = static_cast<double *>(devAlloc(kMaxPartials * 4 * sizeof(double)));
        g_result = static_cast<double *>(devAlloc(64 * sizeof(double)));
    }
}

unsigned redGrid(long long work) {
    long long g = (work + (long long)kBlock * 16 - 1) / ((long long)kBlock * 16);
    return (unsigned)std::max<long long>(1, std::min<long long>(g, kMaxPartials));
}

// second stage: one workgroup sums numPartials x NV values (stride kMaxPartials)
template <int NV>
__global__ void __launch_bounds__(kBlock) kFinish(const double *part, int numPartials, double *res) {
    Acc acc[NV];
    for (int i = threadIdx.x; i < numPartials; i += kBlock)
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v].add(part[v * kMaxPartials + i]);
    double out[NV];
    blockReduce<NV>(acc, out);
    if (threadIdx.x == 0)
#pragma unroll
        for (int v = 0; v < NV; ++v) res[v] = out[v];
}

template <int NV>
__device__ __forceinline__ void writePartial(Acc (&acc)[NV], double *part) {
    double out[NV];
    blockReduce<NV>(acc, out);
    if (threadIdx.x == 0)
#pragma unroll
        for (int v = 0; v < NV; ++v) part[v * kMaxPartials + blockIdx.x] = out[v];
}

// Σ |a_i|^2 over i = insertZeros(w, fixed) | setMask (statevec P0 and total)
template <typename T>
__global__ void __launch_bounds__(kBlock) kSumSq(const Amp<T> *__restrict__ a, long long work,
                                                 BitList fixed, uint64_t setMask, double *part) {
    // four loads in flight per thread before the compensated sums consume them
    constexpr int U = 4;
    Acc acc[1];
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long w0 = (long long)blockIdx.x * kBlock + threadIdx.x; w0 < work; w0 += U * stride) {
        Amp<T> v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long w = w0 + u * stride;
            v[u] = w < work ? ldA<T, true>(a + (insertZeros((uint64_t)w, fixed) | setMask)) : Amp<T>{0, 0};
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double re = v[u].x, im = v[u].y;
            acc[0].add(re * re);
            acc[0].add(im * im);
        }
    }
    writePartial<1>(acc, part);
}

// Σ conj(bra_i) ket_i  (QuEST_cpu.c:1089-1129) — Re and Im in one pass
template <typename TB, typename TK>
__global__ void __launch_bounds__(kBlock) kInner(const Amp<TB> *__restrict__ b, const Amp<TK> *__restrict__ k,
                                                 long long n, double *part) {
    Acc acc[2];
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock) {
        Amp<TB> x = b[i];
        Amp<TK> y = k[i];
        double br = x.x, bi = x.y, kr = y.x, ki = y.y;
        acc[0].add(br * kr + bi * ki);
        acc[1].add(br * ki - bi * kr);
    }
    writePartial<2>(acc, part);
}

// <D> for a diagonal operator D (FP64 op arrays): Σ |ψ_i|^2 D_i on a state vector,
// Σ ρ_ii D_i on a density matrix (QuEST_cpu.c:4107-4209), Re and Im in one pass
template <typename T, bool DENS>
__global__ void __launch_bounds__(kBlock) kExpecDiag(const Amp<T> *__restrict__ a, long long dim,
                                                     const double *__restrict__ opRe,
                                                     const double *__restrict__ opIm, double *part) {
    Acc acc[2];
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < dim; i += (long long)gridDim.x * kBlock) {
        const double dr = opRe[i], di = opIm[i];
        if (DENS) {
            const Amp<T> v = a[i * (dim + 1)];
            acc[0].add((double)v.x * dr - (double)v.y * di);
            acc[1].add((double)v.x * di + (double)v.y * dr);
        } else {
            const Amp<T> v = a[i];
            const double p = (double)v.x * v.x + (double)v.y * v.y;
            acc[0].add(p * dr);
            acc[1].add(p * di);
        }
    }
    writePartial<2>(acc, part);
}

// density diagonal: Σ Re ρ_rr over rows r with bit q == 0 (or all rows)
// QuEST_cpu.c:3385-3431, QuEST_cpu_local.c:121-142
template <typename T>
__global__ void __launch_bounds__(kBlock) kDiag(const Amp<T> *__restrict__ a, long long dim, BitList fixed,
                                                long long work, double *part) {
    Acc acc[1];
    for (long long w = (long long)blockIdx.x * kBlock + threadIdx.x; w < work;
         w += (long long)gridDim.x * kBlock) {
        uint64_t r = insertZeros((uint64_t)w, fixed);
        acc[0].add((double)a[r * (dim + 1)].x);
    }
    writePartial<1>(acc, part);
}

// Σ |a_i|^2 over all elements (purity), Σ Re(conj(a) b), Σ |a - b|^2
template <typename T, int MODE>
__global__ void __launch_bounds__(kBlock) kPair(const Amp<T> *__restrict__ a, const Amp<T> *__restrict__ b,
                                                long long n, double *part) {
    Acc acc[1];
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock) {
        Amp<T> x = a[i];
        double xr = x.x, xi = x.y;
        if (MODE == 0) {
            acc[0].add(xr * xr + xi * xi);
        } else {
            Amp<T> y = b[i];
            double yr = y.x, yi = y.y;
            if (MODE == 1) acc[0].add(xr * yr + xi * yi);
            else {
                double dr = xr - yr, di = xi - yi;
                acc[0].add(dr * dr + di * di);
            }
        }
    }
    writePartial<1>(acc, part);
}

// <psi|rho|psi> = Σ_{r,c} conj(psi_r) rho[r + c*dim] psi_c, real part.
// Every element contributes independently, so the O(dim) serial row loop of
// the reference (QuEST_gpu.cu:2486-2572) becomes one flat coalesced sweep.
template <typename TR, typename TP>
__global__ void __launch_bounds__(kBlock) kFidelity(const Amp<TR> *__restrict__ rho,
                                                    const Amp<TP> *__restrict__ psi, int nRep,
                                                    double *part) {
    const long long dim = 1ll << nRep, n = dim * dim;
    Acc acc[1];
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock) {
        long long r = i & (dim - 1), c = i >> nRep;
        Amp<TR> e = rho[i];
        Amp<TP> pr = psi[r], pc = psi[c];
        double er = e.x, ei = e.y;
        // t = rho_rc * psi_c
        double tr = er * pc.x - ei * pc.y, ti = er * pc.y + ei * pc.x;
        // Re(conj(psi_r) * t)
        acc[0].add((double)pr.x * tr + (double)pr.y * ti);
    }
    writePartial<1>(acc, part);
}

// Outcome histogram in one pass over the register (QuEST_cpu.c
// calcProbOfAllOutcomesLocal makes one pass too). Index i = (hi << 6) | lane:
// a lane keeps the index bits below 6 fixed, so the outcome bits among them
// (lowMask) are fixed per lane, and a wave walks the hi values that carry one
// combination g of the outcome bits at or above 6 (highPos, in hi space),
// split into 2^chunkLog chunks of 2^lenLog values. Each lane sums its
// probabilities with Neumaier compensation; lanes sharing their low outcome
// bits merge by a butterfly over the other lane bits, and the lane whose other
// bits are zero writes the chunk's partial of that outcome. Loads are
// contiguous 64-amplitude runs for any choice of qubits. Deterministic.
struct OutcomePlan {
    BitList highPos;        // outcome qubits >= 6, minus 6, ascending
    int qpos[kMaxBits];     // qubits[j] in the caller's order: outcome bit j
    int k;
    unsigned lowMask;       // lane bits that are outcome qubits
    int chunkLog, lenLog;
    long long units;        // 2^(high outcome bits) x 2^chunkLog
};

template <typename T, bool DENS>
__global__ void __launch_bounds__(kBlock) kOutcomeHist(const Amp<T> *__restrict__ a, long long work, long long dim,
                                                       OutcomePlan pl, double *part) {
    const unsigned lane = threadIdx.x & 63u;
    const long long numWaves = ((long long)gridDim.x * kBlock) >> 6;
    for (long long u = (((long long)blockIdx.x * kBlock) + threadIdx.x) >> 6; u < pl.units; u += numWaves) {
        const uint64_t g = (uint64_t)u >> pl.chunkLog;
        const uint64_t ch = (uint64_t)u & ((1ull << pl.chunkLog) - 1ull);
        uint64_t gbits = 0;  // g deposited at the high outcome positions
        for (int j = 0; j < pl.highPos.n; ++j) gbits |= ((g >> j) & 1ull) << pl.highPos.pos[j];
        Acc acc[1];
        const uint64_t len = 1ull << pl.lenLog;
        const uint64_t first = ((insertZeros(ch << pl.lenLog, pl.highPos) | gbits) << 6) | lane;
        for (uint64_t t = 0; t < len; ++t) {
            const uint64_t i = ((insertZeros((ch << pl.lenLog) | t, pl.highPos) | gbits) << 6) | lane;
            if ((long long)i >= work) continue;
            if (DENS) {
                acc[0].add((double)a[i * (uint64_t)(dim + 1)].x);
            } else {
                const Amp<T> v = a[i];
                acc[0].add((double)v.x * v.x);
                acc[0].add((double)v.y * v.y);
            }
        }
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            if ((pl.lowMask >> b) & 1u) continue;
            const double os = __shfl_xor(acc[0].s, 1 << b, 64);
            const double oc = __shfl_xor(acc[0].c, 1 << b, 64);
            acc[0].merge(os, oc);
        }
        if ((lane & ~pl.lowMask) == 0u && (long long)first < work) {
            uint64_t o = 0;
            for (int j = 0; j < pl.k; ++j) o |= ((first >> pl.qpos[j]) & 1ull) << j;
            part[(o << pl.chunkLog) | ch] = acc[0].value();
        }
    }
}

// per outcome, the chunks' partials in chunk order
__global__ void __launch_bounds__(kBlock) kOutcomeFinish(const double *__restrict__ part, int chunkLog,
                                                         long long numOut, double *out) {
    const long long o = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (o >= numOut) return;
    Acc acc;
    for (long long c = 0; c < (1ll << chunkLog); ++c) acc.add(part[(o << chunkLog) + c]);
    out[o] = acc.value();
}

template <int NV>
void finish(unsigned grid, double *dst) {
    hipLaunchKernelGGL(kFinish<NV>, dim3(1), dim3(kBlock), 0, hstream(), g_partials, (int)grid, g_result);
    checkLaunch("reduce finish");
    double *h = pinnedScratch();
    QMX_HIP_CHECK(hipMemcpyAsync(h, g_result, NV * sizeof(double), hipMemcpyDeviceToHost, hstream()));
    QMX_HIP_CHECK(hipStreamSynchronize(hstream()));
    for (int v = 0; v < NV; ++v) dst[v] = h[v];
}

BitList bitsOf(uint64_t mask) {
    BitList b;
    b.n = 0;
    for (int p = 0; p < 64; ++p)
        if ((mask >> p) & 1ull) b.pos[b.n++] = p;
    return b;
}

double sumSq(const QuregDev *q, uint64_t fixedMask) {
    ensureBuffers();
    const long long work = q->numAmps >> __builtin_popcountll(fixedMask);
    const unsigned grid = redGrid(work);
    BitList fixed = bitsOf(fixedMask);
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kSumSq<T>, dim3(grid), dim3(kBlock), 0, hstream(), (const Amp<T> *)q->amps,
                           work, fixed, (uint64_t)0, g_partials);
    });
    checkLaunch("sumSq");
    double r;
    finish<1>(grid, &r);
    return r;
}

double diagSum(const QuregDev *q, uint64_t fixedMask) {
    ensureBuffers();
    const long long dim = 1ll << q->numQubitsRep;
    const long long work = dim >> __builtin_popcountll(fixedMask);
    const unsigned grid = redGrid(work);
    BitList fixed = bitsOf(fixedMask);
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kDiag<T>, dim3(grid), dim3(kBlock), 0, hstream(), (const Amp<T> *)q->amps, dim,
                           fixed, work, g_partials);
    });
    checkLaunch("diagSum");
    double r;
    finish<1>(grid, &r);
    return r;
}

template <int MODE>
double pairSum(const QuregDev *a, const QuregDev *b) {
    ensureBuffers();
    const unsigned grid = redGrid(a->numAmps);
    QMX_DISPATCH(a->prec, T, {
        hipLaunchKernelGGL((kPair<T, MODE>), dim3(grid), dim3(kBlock), 0, hstream(), (const Amp<T> *)a->amps,
                           (const Amp<T> *)(b ? b->amps : a->amps), a->numAmps, g_partials);
    });
    checkLaunch("pairSum");
    double r;
    finish<1>(grid, &r);
    return r;
}

}  // namespace

double probOfZero(const QuregDev *q, int qubit) { return sumSq(q, 1ull << qubit); }

double totalProb(const QuregDev *q) { return sumSq(q, 0); }

Cplx expecDiagonal(const QuregDev *q, const double *opRe, const double *opIm) {
    ensureBuffers();
    const long long dim = 1LL << q->numQubitsRep;
    const unsigned grid = redGrid(dim);
    QMX_DISPATCH(q->prec, T, {
        if (q->isDensity)
            hipLaunchKernelGGL((kExpecDiag<T, true>), dim3(grid), dim3(kBlock), 0, hstream(), (const Amp<T> *)q->amps,
                               dim, opRe, opIm, g_partials);
        else
            hipLaunchKernelGGL((kExpecDiag<T, false>), dim3(grid), dim3(kBlock), 0, hstream(),
                               (const Amp<T> *)q->amps, dim, opRe, opIm, g_partials);
    });
    checkLaunch("calcExpecDiagonalOp");
    double r[2];
    finish<2>(grid, r);
    return Cplx{r[0], r[1]};
}

Cplx innerProduct(const QuregDev *bra, const QuregDev *ket) {
    ensureBuffers();
    const unsigned grid = redGrid(bra->numAmps);
    QMX_DISPATCH(bra->prec, TB, {
        QMX_DISPATCH(ket->prec, TK, {
            hipLaunchKernelGGL((kInner<TB, TK>), dim3(grid), dim3(kBlock), 0, hstream(),
                               (const Amp<TB> *)bra->amps, (const Amp<TK> *)ket->amps, bra->numAmps,
                               g_partials);
        });
    });
    checkLaunch("innerProduct");
    double r[2];
    finish<2>(grid, r);
    return Cplx{r[0], r[1]};
}

double densProbOfZero(const QuregDev *q, int qubit) { return diagSum(q, 1ull << qubit); }

double densTotalProb(const QuregDev *q) { return diagSum(q, 0); }

double purity(const QuregDev *q) { return pairSum<0>(q, nullptr); }

double densInnerProduct(const QuregDev *a, const QuregDev *b) { return pairSum<1>(a, b); }

double hsDistanceSquared(const QuregDev *a, const QuregDev *b) { return pairSum<2>(a, b); }

double densFidelity(const QuregDev *rho, const QuregDev *psi) {
    ensureBuffers();
    const unsigned grid = redGrid(rho->numAmps);
    QMX_DISPATCH(rho->prec, TR, {
        QMX_DISPATCH(psi->prec, TP, {
            hipLaunchKernelGGL((kFidelity<TR, TP>), dim3(grid), dim3(kBlock), 0, hstream(),
                               (const Amp<TR> *)rho->amps, (const Amp<TP> *)psi->amps, rho->numQubitsRep,
                               g_partials);
        });
    });
    checkLaunch("densFidelity");
    double r;
    finish<1>(grid, &r);
    return r;
}

namespace {

// waves the histogram aims for (16 per CU on 256 CUs), and the grid cap
constexpr int kOutcomeUnitsLog = 12;
constexpr long long kOutcomeMaxBlocks = 8192;

OutcomePlan outcomePlan(const QuregDev *q, const int *qubits, int k) {
    OutcomePlan pl{};
    const long long work = q->isDensity ? (1ll << q->numQubitsRep) : q->numAmps;
    const int workBits = 63 - __builtin_clzll((unsigned long long)work);
    pl.k = k;
    int high[kMaxBits], nh = 0;
    for (int j = 0; j < k; ++j) {
        pl.qpos[j] = qubits[j];
        if (qubits[j] < 6) pl.lowMask |= 1u << qubits[j];
        else high[nh++] = qubits[j] - 6;
    }
    std::sort(high, high + nh);
    pl.highPos.n = nh;
    for (int j = 0; j < nh; ++j) pl.highPos.pos[j] = high[j];
    const int hiBits = std::max(0, workBits - 6 - nh);
    pl.chunkLog = std::min(hiBits, std::max(0, kOutcomeUnitsLog - nh));
    pl.lenLog = hiBits - pl.chunkLog;
    pl.units = 1ll << (nh + pl.chunkLog);
    return pl;
}

}  // namespace

long long outcomeScratchDoubles(const QuregDev *q, const int *qubits, int k) {
    const OutcomePlan pl = outcomePlan(q, qubits, k);
    const long long numOut = 1ll << k;
    return pl.chunkLog ? (numOut << pl.chunkLog) + numOut : numOut;
}

const double *probOfAllOutcomesDevice(const QuregDev *q, const int *qubits, int k, double *scratch) {
    const OutcomePlan pl = outcomePlan(q, qubits, k);
    const bool dens = q->isDensity != 0;
    const long long dim = 1ll << q->numQubitsRep;
    const long long work = dens ? dim : q->numAmps;
    const long long numOut = 1ll << k;
    double *res = scratch;                            // numOut results
    double *part = pl.chunkLog ? scratch + numOut : scratch;
    const long long blocks = std::min(kOutcomeMaxBlocks, (pl.units + 3) / 4);
    QMX_DISPATCH(q->prec, T, {
        if (dens)
            hipLaunchKernelGGL((kOutcomeHist<T, true>), dim3((unsigned)blocks), dim3(kBlock), 0, hstream(),
                               (const Amp<T> *)q->amps, work, dim, pl, part);
        else
            hipLaunchKernelGGL((kOutcomeHist<T, false>), dim3((unsigned)blocks), dim3(kBlock), 0, hstream(),
                               (const Amp<T> *)q->amps, work, dim, pl, part);
    });
    checkLaunch("probOfAllOutcomes");
    if (pl.chunkLog) {
        hipLaunchKernelGGL(kOutcomeFinish, dim3((unsigned)((numOut + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           hstream(), part, pl.chunkLog, numOut, res);
        checkLaunch("probOfAllOutcomes finish");
    }
    return res;
}

void probOfAllOutcomesInto(const QuregDev *q, const int *qubits, int k, double *scratch, double *out) {
    const double *res = probOfAllOutcomesDevice(q, qubits, k, scratch);
    QMX_HIP_CHECK(hipMemcpyAsync(out, res, ((size_t)1 << k) * sizeof(double), hipMemcpyDeviceToHost, hstream()));
    QMX_HIP_CHECK(hipStreamSynchronize(hstream()));
}

void probOfAllOutcomes(const QuregDev *q, const int *qubits, int k, double *out) {
    static double *buf = nullptr;
    static long long bufDoubles = 0;
    const long long need = outcomeScratchDoubles(q, qubits, k);
    if (need > bufDoubles) {
        if (buf) {
            deviceSync();
            devFree(buf);
        }
        buf = static_cast<double *>(devAlloc((size_t)need * sizeof(double)));
        bufDoubles = need;
    }
    probOfAllOutcomesInto(q, qubits, k, buf, out);
}

double normSquaredThenScale(QuregDev *q, bool rescale) {
    double n2 = q->isDensity ? densTotalProb(q) : totalProb(q);
    if (rescale && n2 > 0) scaleAll(q, q->isDensity ? 1.0 / n2 : 1.0 / sqrt(n2));
    return n2;
}

}  // namespace qmx

namespace qmx {

namespace {

// max_i |a_i - b_i| (complex modulus), any precisions
template <typename TA, typename TB>
__global__ void __launch_bounds__(kBlock) kMaxDiff(const Amp<TA> *__restrict__ a, const Amp<TB> *__restrict__ b,
                                                   long long n, double *part) {
    double m = 0.0;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock) {
        double dr = (double)a[i].x - (double)b[i].x, di = (double)a[i].y - (double)b[i].y;
        m = fmax(m, sqrt(dr * dr + di * di));
    }
    for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    __shared__ double lds[kBlock / 64];
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) r = fmax(r, lds[w]);
        part[blockIdx.x] = r;
    }
}

__global__ void __launch_bounds__(kBlock) kMaxFinish(const double *part, int n, double *res) {
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) m = fmax(m, part[i]);
    for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    __shared__ double lds[kBlock / 64];
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) r = fmax(r, lds[w]);
        res[0] = r;
    }
}

}  // namespace

double maxAbsDiff(const QuregDev *a, const QuregDev *b) {
    ensureBuffers();
    const unsigned grid = redGrid(a->numAmps);
    QMX_DISPATCH(a->prec, TA, {
        QMX_DISPATCH(b->prec, TB, {
            hipLaunchKernelGGL((kMaxDiff<TA, TB>), dim3(grid), dim3(kBlock), 0, hstream(), (const Amp<TA> *)a->amps,
                               (const Amp<TB> *)b->amps, a->numAmps, g_partials);
        });
    });
    checkLaunch("maxAbsDiff");
    hipLaunchKernelGGL(kMaxFinish, dim3(1), dim3(kBlock), 0, hstream(), g_partials, (int)grid, g_result);
    checkLaunch("maxAbsDiff finish");
    double *h = pinnedScratch();
    QMX_HIP_CHECK(hipMemcpyAsync(h, g_result, sizeof(double), hipMemcpyDeviceToHost, hstream()));
    QMX_HIP_CHECK(hipStreamSynchronize(hstream()));
    return h[0];
}

}  // namespace qmx
