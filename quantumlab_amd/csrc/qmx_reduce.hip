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
    Acc acc[1];
    for (long long w = (long long)blockIdx.x * kBlock + threadIdx.x; w < work;
         w += (long long)gridDim.x * kBlock) {
        uint64_t i = insertZeros((uint64_t)w, fixed) | setMask;
        Amp<T> v = a[i];
        double re = v.x, im = v.y;
        acc[0].add(re * re);
        acc[0].add(im * im);
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

// histogram of probabilities over the outcome bits (statevec or density diag)
// QuEST_cpu.c calcProbOfAllOutcomesLocal; one pass per 64 outcomes
template <typename T, bool DENS>
__global__ void __launch_bounds__(kBlock) kAllOutcomes(const Amp<T> *__restrict__ a, long long work,
                                                       long long dim, BitList qubits, int first,
                                                       int count, double *part) {
    Acc h[1];
    for (int o = 0; o < count; ++o) {
        h[0] = Acc();
        const int outcome = first + o;
        for (long long w = (long long)blockIdx.x * kBlock + threadIdx.x; w < work;
             w += (long long)gridDim.x * kBlock) {
            uint64_t idx = (uint64_t)w;
            int got = 0;
            for (int j = 0; j < qubits.n; ++j) got |= (int)((idx >> qubits.pos[j]) & 1ull) << j;
            if (got != outcome) continue;
            if (DENS) {
                h[0].add((double)a[idx * (dim + 1)].x);
            } else {
                Amp<T> v = a[idx];
                h[0].add((double)v.x * v.x);
                h[0].add((double)v.y * v.y);
            }
        }
        double out[1];
        blockReduce<1>(h, out);
        if (threadIdx.x == 0) part[(size_t)o * kMaxPartials + blockIdx.x] = out[0];
        __syncthreads();
    }
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

void probOfAllOutcomes(const QuregDev *q, const int *qubits, int k, double *out) {
    ensureBuffers();
    const bool dens = q->isDensity != 0;
    const long long dim = 1ll << q->numQubitsRep;
    const long long work = dens ? dim : q->numAmps;
    const unsigned grid = std::min<unsigned>(redGrid(work), 256);
    BitList qb;
    qb.n = k;
    for (int j = 0; j < k; ++j) qb.pos[j] = qubits[j];
    const int numOut = 1 << k;
    for (int first = 0; first < numOut; first += 4) {
        int count = std::min(4, numOut - first);
        QMX_DISPATCH(q->prec, T, {
            if (dens)
                hipLaunchKernelGGL((kAllOutcomes<T, true>), dim3(grid), dim3(kBlock), 0, hstream(),
                                   (const Amp<T> *)q->amps, work, dim, qb, first, count, g_partials);
            else
                hipLaunchKernelGGL((kAllOutcomes<T, false>), dim3(grid), dim3(kBlock), 0, hstream(),
                                   (const Amp<T> *)q->amps, work, dim, qb, first, count, g_partials);
        });
        checkLaunch("probOfAllOutcomes");
        for (int o = 0; o < count; ++o) {
            hipLaunchKernelGGL(kFinish<1>, dim3(1), dim3(kBlock), 0, hstream(),
                               g_partials + (size_t)o * kMaxPartials, (int)grid, g_result + o);
            checkLaunch("probOfAllOutcomes finish");
        }
        double *h = pinnedScratch();
        QMX_HIP_CHECK(hipMemcpyAsync(h, g_result, count * sizeof(double), hipMemcpyDeviceToHost, hstream()));
        QMX_HIP_CHECK(hipStreamSynchronize(hstream()));
        for (int o = 0; o < count; ++o) out[first + o] = h[o];
    }
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
