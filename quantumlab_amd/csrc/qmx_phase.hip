// Diagonal operators and phase functions: the reference's generic element-wise
// sweeps over the register (QuEST_cpu.c:4030-4105 diagonal operators,
// 4251-4620 phase functions), one thread per amplitude. These are not on a
// workload's hot path; they are one streaming pass each, like the gate kernels.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "qmx_device.h"

namespace qmx {

namespace {

constexpr int kPer = 4;
constexpr int kMaxRegs = 100;  // MAX_NUM_REGS_APPLY_ARBITRARY_PHASE

// amp_i *= D_{i & rowMask}: rowMask = all ones (state vector) or 2^n - 1 (the
// row index of a density matrix)
template <typename T>
__global__ void __launch_bounds__(kBlock) kDiagApply(Amp<T> *__restrict__ a, long long n, unsigned long long rowMask,
                                                     const double *__restrict__ opRe,
                                                     const double *__restrict__ opIm) {
    long long i = flatBlockLinear() * kBlock * kPer + threadIdx.x;
#pragma unroll
    for (int j = 0; j < kPer; ++j, i += kBlock) {
        if (i < n) {
            const long long r = (long long)((unsigned long long)i & rowMask);
            const double c = opRe[r], d = opIm[r];
            const Amp<T> v = a[i];
            const double x = v.x, y = v.y;
            a[i] = Amp<T>{(T)(x * c - y * d), (T)(x * d + y * c)};
        }
    }
}

// D_i = Σ_t c_t (-1)^parity(i & zMask_t) (QuEST_cpu.c:1368-1428)
__global__ void __launch_bounds__(kBlock) kDiagFromZ(double *__restrict__ re, double *__restrict__ im, long long n,
                                                     const unsigned long long *__restrict__ z,
                                                     const double *__restrict__ c, int numTerms) {
    const long long i = flatBlockLinear() * kBlock + threadIdx.x;
    if (i >= n) return;
    double e = 0;
    for (int t = 0; t < numTerms; ++t) e += (__popcll((unsigned long long)i & z[t]) & 1) ? -c[t] : c[t];
    re[i] = e;
    im[i] = 0;
}

struct PhaseDev {
    int numRegs, encoding, mode, funcCode, numOverrides, conj;
    const int *qubits, *numQubitsPerReg, *numTermsPerReg;
    const double *coeffs, *exponents, *params, *overridePhases;
    const long long *overrideInds;
};

// The phase arithmetic runs in qreal with every operation rounded separately,
// as QuEST's loops do (QuEST_cpu.c:4268-4323, 4347-4424, 4449-4584: qreal
// phase, norm, prod and dist, integer sub-register values converted where C
// converts them). In the float build (QuEST_PREC=1) phases reach 10^5 rad,
// where a float step is 0.01 rad, so accumulating in double instead would move
// a phase by more than QuEST's own tests allow against a float evaluation.
using PhaseReal = qreal;

__device__ PhaseReal namedPhase(const PhaseDev &p, const long long *x) {
#pragma clang fp contract(off)
    typedef PhaseReal R;
    const int f = p.funcCode;
    const double *prm = p.params;  // qreal values, held exactly in double
    if (f == NORM || f == INVERSE_NORM || f == SCALED_NORM || f == SCALED_INVERSE_NORM ||
        f == SCALED_INVERSE_SHIFTED_NORM) {
        R norm = 0;
        for (int r = 0; r < p.numRegs; ++r) {
            if (f == SCALED_INVERSE_SHIFTED_NORM) {
                const R d = (R)x[r] - (R)prm[2 + r];
                norm += d * d;
            } else {
                norm += (R)(x[r] * x[r]);
            }
        }
        norm = (R)sqrt((double)norm);
        if (f == NORM) return norm;
        if (f == INVERSE_NORM) return norm == 0. ? (R)prm[0] : (R)1 / norm;
        if (f == SCALED_NORM) return (R)prm[0] * norm;
        return norm <= REAL_EPS ? (R)prm[1] : (R)prm[0] / norm;
    }
    if (f == PRODUCT || f == INVERSE_PRODUCT || f == SCALED_PRODUCT || f == SCALED_INVERSE_PRODUCT) {
        R prod = 1;
        for (int r = 0; r < p.numRegs; ++r) prod *= (R)x[r];
        if (f == PRODUCT) return prod;
        if (f == INVERSE_PRODUCT) return prod == 0. ? (R)prm[0] : (R)1 / prod;
        if (f == SCALED_PRODUCT) return (R)prm[0] * prod;
        return prod == 0. ? (R)prm[1] : (R)prm[0] / prod;
    }
    R dist = 0;  // distances between consecutive register pairs
    for (int r = 0; r + 1 < p.numRegs; r += 2) {
        if (f == SCALED_INVERSE_SHIFTED_DISTANCE) {
            const R d = (R)(x[r] - x[r + 1]) - (R)prm[2 + r / 2];
            dist += d * d;
        } else {
            dist += (R)((x[r + 1] - x[r]) * (x[r + 1] - x[r]));
        }
    }
    dist = (R)sqrt((double)dist);
    if (f == DISTANCE) return dist;
    if (f == INVERSE_DISTANCE) return dist == 0. ? (R)prm[0] : (R)1 / dist;
    if (f == SCALED_DISTANCE) return (R)prm[0] * dist;
    return dist <= REAL_EPS ? (R)prm[1] : (R)prm[0] / dist;
}

// one thread per amplitude: sub-register values, override lookup, phase,
// amp *= exp(i phase) (QuEST_cpu.c:4251-4620)
template <typename T>
__global__ void __launch_bounds__(kBlock) kPhaseFunc(Amp<T> *__restrict__ a, long long n, PhaseDev p) {
#pragma clang fp contract(off)
    typedef PhaseReal R;
    // amplitudes update in float only when both the register and qreal are float
    typedef typename std::conditional<sizeof(T) == 4 && sizeof(R) == 4, float, double>::type A;
    const long long i = flatBlockLinear() * kBlock + threadIdx.x;
    if (i >= n) return;
    long long x[kMaxRegs];
    int flat = 0;
    for (int r = 0; r < p.numRegs; ++r) {
        const int nq = p.numQubitsPerReg[r];
        long long v = 0;
        const int valueBits = p.encoding == TWOS_COMPLEMENT ? nq - 1 : nq;
        for (int b = 0; b < valueBits; ++b) v += (long long)((i >> p.qubits[flat++]) & 1) << b;
        if (p.encoding == TWOS_COMPLEMENT && ((i >> p.qubits[flat++]) & 1)) v -= 1LL << (nq - 1);
        x[r] = v;
    }
    int o = 0;
    for (; o < p.numOverrides; ++o) {
        bool hit = true;
        for (int r = 0; r < p.numRegs && hit; ++r) hit = x[r] == p.overrideInds[(long long)o * p.numRegs + r];
        if (hit) break;
    }
    R phase = 0;
    if (o < p.numOverrides) {
        phase = (R)p.overridePhases[o];
    } else if (p.mode == 0) {
        // phase += coeff * pow(x, exponent): the product in double, each sum rounded to qreal
        int t = 0;
        for (int r = 0; r < p.numRegs; ++r)
            for (int k = 0; k < p.numTermsPerReg[r]; ++k, ++t)
                phase = (R)((double)phase + p.coeffs[t] * pow((double)x[r], p.exponents[t]));
    } else {
        phase = namedPhase(p, x);
    }
    if (p.conj) phase = -phase;
    const A c = (A)(R)cos((double)phase), s = (A)(R)sin((double)phase);
    const Amp<T> v = a[i];
    const A re = v.x, im = v.y;
    a[i] = Amp<T>{(T)(re * c - im * s), (T)(re * s + im * c)};
}

template <typename V>
V *uploadVec(const V *h, size_t n, std::vector<void *> &owned) {
    void *d = devAlloc(std::max<size_t>(1, n) * sizeof(V));
    if (n) hostToDevice(d, h, n * sizeof(V));
    owned.push_back(d);
    return static_cast<V *>(d);
}

}  // namespace

void applyDiagonal(QuregDev *q, const double *opRe, const double *opIm) {
    const unsigned long long rowMask = q->isDensity ? ((1ull << q->numQubitsRep) - 1ull) : ~0ull;
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kDiagApply<T>, dim3(gridFor(q->numAmps, kPer)), dim3(kBlock), 0, hstream(),
                           (Amp<T> *)q->amps, q->numAmps, rowMask, opRe, opIm);
    });
    checkLaunch("applyDiagonalOp");
}

void diagonalFromZHamil(double *opRe, double *opIm, int numQubits, const uint64_t *zMasks, const double *coeffs,
                        int numTerms) {
    std::vector<void *> owned;
    const unsigned long long *dz =
        uploadVec(reinterpret_cast<const unsigned long long *>(zMasks), (size_t)numTerms, owned);
    const double *dc = uploadVec(coeffs, (size_t)numTerms, owned);
    const long long n = 1LL << numQubits;
    hipLaunchKernelGGL(kDiagFromZ, dim3(gridFor(n, 1)), dim3(kBlock), 0, hstream(), opRe, opIm, n, dz, dc, numTerms);
    checkLaunch("initDiagonalOpFromPauliHamil");
    for (void *d : owned) devFree(d);
}

void applyPhaseFunction(QuregDev *q, const PhaseFuncArgs &args, int qubitShift, bool conj) {
    int totalQubits = 0, totalTerms = 0;
    for (int r = 0; r < args.numRegs; ++r) {
        totalQubits += args.numQubitsPerReg[r];
        if (args.mode == 0) totalTerms += args.numTermsPerReg[r];
    }
    std::vector<int> qs(args.qubits, args.qubits + totalQubits);
    for (int &b : qs) b += qubitShift;
    std::vector<void *> owned;
    PhaseDev p{};
    p.numRegs = args.numRegs;
    p.encoding = args.encoding;
    p.mode = args.mode;
    p.funcCode = args.funcCode;
    p.numOverrides = args.numOverrides;
    p.conj = conj ? 1 : 0;
    p.qubits = uploadVec(qs.data(), qs.size(), owned);
    p.numQubitsPerReg = uploadVec(args.numQubitsPerReg, (size_t)args.numRegs, owned);
    if (args.mode == 0) {
        p.numTermsPerReg = uploadVec(args.numTermsPerReg, (size_t)args.numRegs, owned);
        p.coeffs = uploadVec(args.coeffs, (size_t)totalTerms, owned);
        p.exponents = uploadVec(args.exponents, (size_t)totalTerms, owned);
    } else {
        p.params = uploadVec(args.params, (size_t)args.numParams, owned);
    }
    p.overrideInds = uploadVec(args.overrideInds, (size_t)args.numOverrides * args.numRegs, owned);
    p.overridePhases = uploadVec(args.overridePhases, (size_t)args.numOverrides, owned);
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kPhaseFunc<T>, dim3(gridFor(q->numAmps, 1)), dim3(kBlock), 0, hstream(),
                           (Amp<T> *)q->amps, q->numAmps, p);
    });
    checkLaunch("applyPhaseFunc");
    for (void *d : owned) devFree(d);
}

}  // namespace qmx
