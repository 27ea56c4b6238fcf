// State set-up, host-mirror transfer and element-wise combination kernels.
// Each is one coalesced streaming sweep over the interleaved amplitudes. The
// whole-register sweeps (fill, copy, scale) take the form of the 2x2 lane
// sweep: one amplitude per thread, nontemporal 16 B (FP64) / 8 B (FP32)
// accesses, XCD-ordered workgroups.
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "qmx_device.h"

namespace qmx {

namespace {

template <typename T>
__global__ void __launch_bounds__(kBlock) kFill(Amp<T> *a, long long n, T re, T im) {
    const long long i = flatBlock() * kBlock + threadIdx.x;
    if (i < n) stA<T, true>(a + i, Amp<T>{re, im});
}

template <typename T>
__global__ void kSetOne(Amp<T> *a, long long idx, T re, T im) {
    a[idx] = Amp<T>{re, im};
}

// QuEST_cpu.c:1655-1660: amp k = (2k + i(2k+1)) / 10
template <typename T>
__global__ void kDebug(Amp<T> *a, long long n) {
    long long i = flatBlockLinear() * kBlock * 4 + threadIdx.x;
#pragma unroll
    for (int j = 0; j < 4; ++j, i += kBlock)
        if (i < n) a[i] = Amp<T>{(T)((i * 2.0) / 10.0), (T)((i * 2.0 + 1.0) / 10.0)};
}

// QuEST_cpu.c:1617-1628: qubit fixed to `outcome`, others in |+>
template <typename T>
__global__ void kSingleQubit(Amp<T> *a, long long n, int qubit, int outcome, double norm) {
    long long i = flatBlockLinear() * kBlock * 4 + threadIdx.x;
#pragma unroll
    for (int j = 0; j < 4; ++j, i += kBlock)
        if (i < n) {
            int bit = (int)((i >> qubit) & 1);
            a[i] = Amp<T>{bit == outcome ? (T)norm : (T)0, (T)0};
        }
}

// rho[r + c*dim] = psi[r] * conj(psi[c]) (QuEST_cpu.c:1216-1253)
template <typename TR, typename TP>
__global__ void kPureDensity(Amp<TR> *rho, const Amp<TP> *psi, int nRep) {
    long long dim = 1ll << nRep;
    long long n = dim * dim;
    long long i = flatBlockLinear() * kBlock * 4 + threadIdx.x;
#pragma unroll
    for (int j = 0; j < 4; ++j, i += kBlock)
        if (i < n) {
            long long r = i & (dim - 1), c = i >> nRep;
            double kr = psi[r].x, ki = psi[r].y;
            double br = psi[c].x, bi = -(double)psi[c].y;
            rho[i] = Amp<TR>{(TR)(kr * br - ki * bi), (TR)(kr * bi + ki * br)};
        }
}

template <typename TD, typename TS>
__global__ void __launch_bounds__(kBlock) kCopy(Amp<TD> *d, const Amp<TS> *s, long long n) {
    const long long i = flatBlock() * kBlock + threadIdx.x;
    if (i >= n) return;
    const Amp<TS> v = ldA<TS, true>(s + i);
    stA<TD, true>(d + i, Amp<TD>{(TD)v.x, (TD)v.y});
}

template <typename T, typename H>
__global__ void kInterleave(Amp<T> *a, const H *re, const H *im, long long n) {
    long long i = flatBlockLinear() * kBlock * 4 + threadIdx.x;
#pragma unroll
    for (int j = 0; j < 4; ++j, i += kBlock)
        if (i < n) a[i] = Amp<T>{(T)re[i], (T)im[i]};
}

template <typename T, typename H>
__global__ void kDeinterleave(const Amp<T> *a, H *re, H *im, long long n) {
    long long i = flatBlockLinear() * kBlock * 4 + threadIdx.x;
#pragma unroll
    for (int j = 0; j < 4; ++j, i += kBlock)
        if (i < n) {
            Amp<T> v = a[i];
            re[i] = (H)v.x;
            im[i] = (H)v.y;
        }
}

// out = fac1*q1 + fac2*q2 + facOut*out with QuEST's grouping and every
// operation rounded separately, (facOut out) + (fac1 q1) + (fac2 q2), each
// product's two terms formed first (QuEST_cpu.c:3988-4027, built without
// contraction): the same inputs give the same bits. In qreal, in float only
// when both the registers and qreal are float.
template <typename T>
__global__ void kWeighted(Cplx f1, const Amp<T> *a, Cplx f2, const Amp<T> *b, Cplx fo, Amp<T> *o,
                          long long n) {
#pragma clang fp contract(off)
    using A = typename std::conditional<sizeof(T) == 4 && sizeof(qreal) == 4, float, double>::type;
    const A f1r = (A)f1.re, f1i = (A)f1.im, f2r = (A)f2.re, f2i = (A)f2.im, for_ = (A)fo.re, foi = (A)fo.im;
    long long i = flatBlockLinear() * kBlock * 4 + threadIdx.x;
#pragma unroll
    for (int j = 0; j < 4; ++j, i += kBlock)
        if (i < n) {
            const A ar = a[i].x, ai = a[i].y, br = b[i].x, bi = b[i].y, orr = o[i].x, oi = o[i].y;
            const A re = (for_ * orr - foi * oi) + (f1r * ar - f1i * ai) + (f2r * br - f2i * bi);
            const A im = (for_ * oi + foi * orr) + (f1r * ai + f1i * ar) + (f2r * bi + f2i * br);
            o[i] = Amp<T>{(T)re, (T)im};
        }
}

// combine = (1-p) combine + p other (QuEST_cpu.c:908-936)
template <typename T>
__global__ void kMixDens(Amp<T> *c, const Amp<T> *o, double p, long long n) {
    // in qreal, each operation rounded separately, as QuEST's loop
#pragma clang fp contract(off)
    using A = typename std::conditional<sizeof(T) == 4 && sizeof(qreal) == 4, float, double>::type;
    const A pa = (A)p, q = (A)1 - pa;
    long long i = flatBlockLinear() * kBlock * 4 + threadIdx.x;
#pragma unroll
    for (int j = 0; j < 4; ++j, i += kBlock)
        if (i < n) {
            A cr = (A)c[i].x * q, ci = (A)c[i].y * q;
            cr += pa * (A)o[i].x;
            ci += pa * (A)o[i].y;
            c[i] = Amp<T>{(T)cr, (T)ci};
        }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) kScale(Amp<T> *a, long long n, double f) {
    const long long i = flatBlock() * kBlock + threadIdx.x;
    if (i >= n) return;
    const Amp<T> v = ldA<T, true>(a + i);
    stA<T, true>(a + i, Amp<T>{(T)(v.x * f), (T)(v.y * f)});
}

// staging buffer for SoA host <-> interleaved device transfers
constexpr long long kStageAmps = 1ll << 22;
std::mutex g_stageMu;
double *g_stage = nullptr;  // 2 * kStageAmps doubles

double *stage() {
    if (!g_stage) g_stage = static_cast<double *>(devAlloc(2 * kStageAmps * sizeof(double)));
    return g_stage;
}

}  // namespace

void initBlank(QuregDev *q) {
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kFill<T>, gridFor(q->numAmps, 1), dim3(kBlock), 0, hstream(),
                           (Amp<T> *)q->amps, q->numAmps, (T)0, (T)0);
    });
    checkLaunch("initBlank");
}

void initClassical(QuregDev *q, long long index) {
    initBlank(q);
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kSetOne<T>, dim3(1), dim3(1), 0, hstream(), (Amp<T> *)q->amps, index,
                           (T)1, (T)0);
    });
    checkLaunch("initClassical");
}

void initUniform(QuregDev *q, double re) {
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kFill<T>, gridFor(q->numAmps, 1), dim3(kBlock), 0, hstream(),
                           (Amp<T> *)q->amps, q->numAmps, (T)re, (T)0);
    });
    checkLaunch("initUniform");
}

void initDebug(QuregDev *q) {
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kDebug<T>, dim3(gridFor(q->numAmps, 4)), dim3(kBlock), 0, hstream(),
                           (Amp<T> *)q->amps, q->numAmps);
    });
    checkLaunch("initDebug");
}

void initSingleQubit(QuregDev *q, int qubit, int outcome) {
    double norm = 1.0 / sqrt((double)q->numAmps / 2.0);
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kSingleQubit<T>, dim3(gridFor(q->numAmps, 4)), dim3(kBlock), 0,
                           hstream(), (Amp<T> *)q->amps, q->numAmps, qubit, outcome, norm);
    });
    checkLaunch("initSingleQubit");
}

void initPureDensity(QuregDev *rho, const QuregDev *psi) {
    QMX_DISPATCH(rho->prec, TR, {
        QMX_DISPATCH(psi->prec, TP, {
            hipLaunchKernelGGL((kPureDensity<TR, TP>), dim3(gridFor(rho->numAmps, 4)),
                               dim3(kBlock), 0, hstream(), (Amp<TR> *)rho->amps,
                               (const Amp<TP> *)psi->amps, rho->numQubitsRep);
        });
    });
    checkLaunch("initPureDensity");
}

// A copy kernel rather than a runtime device-to-device copy: the runtime's copy
// moved 128 GiB (32 qubits, FP64) in 27.4 ms, the nontemporal stream needs ~20.
void copyState(QuregDev *dst, const QuregDev *src) {
    QMX_DISPATCH(dst->prec, TD, {
        QMX_DISPATCH(src->prec, TS, {
            hipLaunchKernelGGL((kCopy<TD, TS>), gridFor(src->numAmps, 1), dim3(kBlock), 0,
                               hstream(), (Amp<TD> *)dst->amps, (const Amp<TS> *)src->amps,
                               src->numAmps);
        });
    });
    checkLaunch("copyState");
}

// Host SoA arrays of H (double or float) <-> the interleaved device amplitudes,
// through the staging buffer in chunks of kStageAmps.
template <typename H>
void uploadT(QuregDev *q, long long start, const H *re, const H *im, long long count) {
    std::lock_guard<std::mutex> lk(g_stageMu);
    H *s = reinterpret_cast<H *>(stage());
    for (long long off = 0; off < count; off += kStageAmps) {
        long long m = std::min(kStageAmps, count - off);
        QMX_HIP_CHECK(hipMemcpyAsync(s, re + off, m * sizeof(H), hipMemcpyHostToDevice, hstream()));
        QMX_HIP_CHECK(hipMemcpyAsync(s + kStageAmps, im + off, m * sizeof(H), hipMemcpyHostToDevice, hstream()));
        QMX_DISPATCH(q->prec, T, {
            hipLaunchKernelGGL((kInterleave<T, H>), dim3(gridFor(m, 4)), dim3(kBlock), 0, hstream(),
                               (Amp<T> *)q->amps + start + off, s, s + kStageAmps, m);
        });
        checkLaunch("uploadAmps");
    }
    QMX_HIP_CHECK(hipStreamSynchronize(hstream()));
}

template <typename H>
void downloadT(const QuregDev *q, long long start, H *re, H *im, long long count) {
    std::lock_guard<std::mutex> lk(g_stageMu);
    H *s = reinterpret_cast<H *>(stage());
    for (long long off = 0; off < count; off += kStageAmps) {
        long long m = std::min(kStageAmps, count - off);
        QMX_DISPATCH(q->prec, T, {
            hipLaunchKernelGGL((kDeinterleave<T, H>), dim3(gridFor(m, 4)), dim3(kBlock), 0, hstream(),
                               (const Amp<T> *)q->amps + start + off, s, s + kStageAmps, m);
        });
        checkLaunch("downloadAmps");
        QMX_HIP_CHECK(hipMemcpyAsync(re + off, s, m * sizeof(H), hipMemcpyDeviceToHost, hstream()));
        QMX_HIP_CHECK(hipMemcpyAsync(im + off, s + kStageAmps, m * sizeof(H), hipMemcpyDeviceToHost, hstream()));
        QMX_HIP_CHECK(hipStreamSynchronize(hstream()));
    }
}

void uploadAmps(QuregDev *q, long long start, const double *re, const double *im, long long count) {
    uploadT(q, start, re, im, count);
}
void uploadAmps(QuregDev *q, long long start, const float *re, const float *im, long long count) {
    uploadT(q, start, re, im, count);
}
void downloadAmps(const QuregDev *q, long long start, double *re, double *im, long long count) {
    downloadT(q, start, re, im, count);
}
void downloadAmps(const QuregDev *q, long long start, float *re, float *im, long long count) {
    downloadT(q, start, re, im, count);
}

void weightedSum(Cplx f1, const QuregDev *q1, Cplx f2, const QuregDev *q2, Cplx fo, QuregDev *out) {
    QMX_DISPATCH(out->prec, T, {
        hipLaunchKernelGGL(kWeighted<T>, dim3(gridFor(out->numAmps, 4)), dim3(kBlock), 0,
                           hstream(), f1, (const Amp<T> *)q1->amps, f2, (const Amp<T> *)q2->amps,
                           fo, (Amp<T> *)out->amps, out->numAmps);
    });
    checkLaunch("weightedSum");
}

void mixDensityMatrix(QuregDev *combine, double otherProb, const QuregDev *other) {
    QMX_DISPATCH(combine->prec, T, {
        hipLaunchKernelGGL(kMixDens<T>, dim3(gridFor(combine->numAmps, 4)), dim3(kBlock), 0,
                           hstream(), (Amp<T> *)combine->amps, (const Amp<T> *)other->amps,
                           otherProb, combine->numAmps);
    });
    checkLaunch("mixDensityMatrix");
}

void scaleAll(QuregDev *q, double factor) {
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kScale<T>, gridFor(q->numAmps, 1), dim3(kBlock), 0, hstream(),
                           (Amp<T> *)q->amps, q->numAmps, factor);
    });
    checkLaunch("scaleAll");
}

}  // namespace qmx
