// Batched Trotter-Suzuki evolution of many small registers, one register per
// workgroup held in LDS (labelling data for the Trotter planner).
//
// applyTrotterCircuit semantics (QuEST_common.c:762-844): reps repetitions of
// the order's formula at tau = time / reps; the formula is a list of stages
// (factor f, direction) — order 1: (1, fwd); order 2: (1/2, fwd), (1/2, rev);
// order 2k: the order-(2k-2) list at p tau, p tau, (1-4p) tau, p tau, p tau with
// p = 1 / (4 - 4^(1/(2k-1))). A stage applies exp(-i f tau c_j P_j) for every
// term j in that direction (identity strings skipped, as multiRotatePauli does
// when its mask is empty). Each Pauli exponential is one pass over the
// register: P|i> = i^nY (-1)^{|i & z|} |i ^ x>, so out_i = cos(phi) psi_i -
// i sin(phi) (P psi)_i pairs i with i ^ x, or multiplies by exp(-i phi s_i) when
// x = 0.
//
// Registers of n <= 12 qubits (64 KiB of FP64 amplitudes) live in LDS for the
// whole evolution: no global traffic per term, one barrier per exchange term.
// The ladder kernel evolves a register from its product state for r = 1, 2,
// 4, ... repetitions and records 1 - |<psi_r|psi_4r>|^2 until it reaches eps;
// the two previous rungs wait in a per-register global ring of 3 slots.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "qmx_device.h"
#include "quest_mi355x.h"

namespace qmx {

namespace {

constexpr int kMaxLdsQubits = 12;

struct PTerm {
    unsigned x, z;  // Pauli masks (Y sets both)
    int ny;         // number of Y factors
    int pad;
    double coeff;
};

struct Stage {
    double fac;
    int reverse;
    int pad;
};

struct BatchJob {
    int n, numTerms, termOff, order;
    int stageOff, numStages, angleOff, reps;  // reps: evolve kernel only
    double time;
    long long outOff;                         // evolve: amplitude offset; ladder: scratch offset
};

void stageList(int order, double f, std::vector<Stage> &out) {
    if (order == 1) {
        out.push_back(Stage{f, 0, 0});
    } else if (order == 2) {
        out.push_back(Stage{f / 2, 0, 0});
        out.push_back(Stage{f / 2, 1, 0});
    } else {
        const double p = 1. / (4 - std::pow(4, 1. / (order - 1)));
        const int lower = order - 2;
        stageList(lower, p * f, out);
        stageList(lower, p * f, out);
        stageList(lower, (1 - 4 * p) * f, out);
        stageList(lower, p * f, out);
        stageList(lower, p * f, out);
    }
}

// product state: amplitude of i = prod_q (bit q ? sin(th/2) e^{i ph/2} : cos(th/2) e^{-i ph/2})
__device__ void initProduct(double2 *st, int n, const double2 *f0, const double2 *f1) {
    const int dim = 1 << n;
    for (int i = threadIdx.x; i < dim; i += kBlock) {
        double re = 1.0, im = 0.0;
        for (int q = 0; q < n; ++q) {
            const double2 f = ((i >> q) & 1) ? f1[q] : f0[q];
            const double r = re * f.x - im * f.y;
            im = re * f.y + im * f.x;
            re = r;
        }
        st[i] = make_double2(re, im);
    }
}

// one Pauli exponential exp(-i phi P) on the LDS register
__device__ __forceinline__ void pauliExpLds(double2 *st, int n, const PTerm &t, double cs, double sn) {
    const int dim = 1 << n;
    if (t.x == 0) {
        for (int i = threadIdx.x; i < dim; i += kBlock) {
            const double s = (__popc((unsigned)i & t.z) & 1) ? -sn : sn;
            const double2 v = st[i];
            st[i] = make_double2(cs * v.x + s * v.y, cs * v.y - s * v.x);
        }
        return;
    }
    const unsigned lowX = t.x & (~t.x + 1u);
    const int lb = __ffs(t.x) - 1;
    // i^ny
    const double phr = (t.ny & 3) == 0 ? 1.0 : ((t.ny & 3) == 2 ? -1.0 : 0.0);
    const double phi = (t.ny & 3) == 1 ? 1.0 : ((t.ny & 3) == 3 ? -1.0 : 0.0);
    const double kre = sn * phi, kim = -sn * phr;  // -i sn i^ny
    for (int w = threadIdx.x; w < (dim >> 1); w += kBlock) {
        const unsigned i = (((unsigned)w >> lb) << (lb + 1)) | ((unsigned)w & (lowX - 1u));
        const unsigned p = i ^ t.x;
        const double2 vi = st[i], vp = st[p];
        const double si = (__popc(p & t.z) & 1) ? -1.0 : 1.0;
        const double sp = (__popc(i & t.z) & 1) ? -1.0 : 1.0;
        st[i] = make_double2(cs * vi.x + si * (kre * vp.x - kim * vp.y), cs * vi.y + si * (kre * vp.y + kim * vp.x));
        st[p] = make_double2(cs * vp.x + sp * (kre * vi.x - kim * vi.y), cs * vp.y + sp * (kre * vi.y + kim * vi.x));
    }
}

// reps repetitions of the job's stage list; ends with a barrier
__device__ void evolveLds(double2 *st, const BatchJob &j, const PTerm *terms, const Stage *stages, int reps) {
    const double tau = j.time / reps;
    bool pendingDiag = false;
    for (int r = 0; r < reps; ++r) {
        for (int s = 0; s < j.numStages; ++s) {
            const Stage sg = stages[j.stageOff + s];
            const double ft = sg.fac * tau;
            for (int k = 0; k < j.numTerms; ++k) {
                const PTerm t = terms[j.termOff + (sg.reverse ? j.numTerms - 1 - k : k)];
                if ((t.x | t.z) == 0) continue;
                const double phi = ft * t.coeff;
                const double cs = cos(phi), sn = sin(phi);
                if (t.x == 0) {
                    // diagonal terms touch only the thread's own amplitudes
                    pauliExpLds(st, j.n, t, cs, sn);
                    pendingDiag = true;
                } else {
                    if (pendingDiag) __syncthreads();
                    pauliExpLds(st, j.n, t, cs, sn);
                    __syncthreads();
                    pendingDiag = false;
                }
            }
        }
    }
    __syncthreads();
}

__device__ void loadFactors(const BatchJob &j, const double *theta, const double *phiA, double2 *f0, double2 *f1) {
    if (threadIdx.x < j.n) {
        const double th = theta[j.angleOff + threadIdx.x], ph = phiA[j.angleOff + threadIdx.x];
        const double c = cos(th / 2), s = sin(th / 2), pc = cos(ph / 2), ps = sin(ph / 2);
        f0[threadIdx.x] = make_double2(c * pc, -c * ps);
        f1[threadIdx.x] = make_double2(s * pc, s * ps);
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kBlock) kTrotterEvolve(const BatchJob *jobs, const PTerm *terms, const Stage *stages,
                                                         const double *theta, const double *phiA, double2 *out) {
    extern __shared__ double2 st[];
    __shared__ double2 f0[kMaxLdsQubits], f1[kMaxLdsQubits];
    const BatchJob j = jobs[blockIdx.x];
    loadFactors(j, theta, phiA, f0, f1);
    initProduct(st, j.n, f0, f1);
    __syncthreads();
    evolveLds(st, j, terms, stages, j.reps);
    for (int i = threadIdx.x; i < (1 << j.n); i += kBlock) out[j.outOff + i] = st[i];
}

__global__ void __launch_bounds__(kBlock) kTrotterLadder(const BatchJob *jobs, const PTerm *terms, const Stage *stages,
                                                         const double *theta, const double *phiA, double2 *ring,
                                                         int maxLog2, double eps, double *infid, int *pass) {
    extern __shared__ double2 st[];
    __shared__ double2 f0[kMaxLdsQubits], f1[kMaxLdsQubits];
    __shared__ double result;
    const BatchJob j = jobs[blockIdx.x];
    const int dim = 1 << j.n;
    double2 *slots = ring + j.outOff;  // 3 x dim
    loadFactors(j, theta, phiA, f0, f1);
    int passed = -1;
    for (int k = 0; k <= maxLog2 + 2; ++k) {
        initProduct(st, j.n, f0, f1);
        __syncthreads();
        evolveLds(st, j, terms, stages, 1 << k);
        if (k >= 2) {
            // candidate r = 2^(k-2) against the 4x finer rung just computed
            const double2 *prev = slots + (size_t)((k - 2) % 3) * dim;
            Acc acc[2];
            for (int i = threadIdx.x; i < dim; i += kBlock) {
                const double2 a = prev[i], b = st[i];
                acc[0].add(a.x * b.x + a.y * b.y);
                acc[1].add(a.x * b.y - a.y * b.x);
            }
            double ip[2];
            blockReduce<2>(acc, ip);
            if (threadIdx.x == 0) {
                result = 1.0 - (ip[0] * ip[0] + ip[1] * ip[1]);
                infid[(size_t)blockIdx.x * (maxLog2 + 1) + (k - 2)] = result;
            }
            __syncthreads();
            if (result <= eps) {
                passed = k - 2;
                break;
            }
        }
        double2 *dst = slots + (size_t)(k % 3) * dim;
        for (int i = threadIdx.x; i < dim; i += kBlock) dst[i] = st[i];
        __syncthreads();
    }
    if (threadIdx.x == 0) pass[blockIdx.x] = passed;
}

// 1 - |<psi_r|psi_4r>|^2 for the job's r (BatchJob::reps); psi_r waits in a
// global slot of the register's size while psi_4r is evolved in LDS
__global__ void __launch_bounds__(kBlock) kTrotterInfidelity(const BatchJob *jobs, const PTerm *terms,
                                                             const Stage *stages, const double *theta,
                                                             const double *phiA, double2 *slot, double *infid) {
    extern __shared__ double2 st[];
    __shared__ double2 f0[kMaxLdsQubits], f1[kMaxLdsQubits];
    const BatchJob j = jobs[blockIdx.x];
    const int dim = 1 << j.n;
    double2 *keep = slot + j.outOff;
    loadFactors(j, theta, phiA, f0, f1);
    initProduct(st, j.n, f0, f1);
    __syncthreads();
    evolveLds(st, j, terms, stages, j.reps);
    for (int i = threadIdx.x; i < dim; i += kBlock) keep[i] = st[i];
    __syncthreads();
    initProduct(st, j.n, f0, f1);
    __syncthreads();
    evolveLds(st, j, terms, stages, 4 * j.reps);
    Acc acc[2];
    for (int i = threadIdx.x; i < dim; i += kBlock) {
        const double2 a = keep[i], b = st[i];
        acc[0].add(a.x * b.x + a.y * b.y);
        acc[1].add(a.x * b.y - a.y * b.x);
    }
    double ip[2];
    blockReduce<2>(acc, ip);
    if (threadIdx.x == 0) infid[blockIdx.x] = 1.0 - (ip[0] * ip[0] + ip[1] * ip[1]);
}

// host-side packing shared by the entry points
struct Packed {
    std::vector<BatchJob> jobs;
    std::vector<PTerm> terms;
    std::vector<Stage> stages;
    std::vector<double> theta, phi;
};

template <typename R>
bool pack(int numJobs, const PauliHamil *hamils, const R *times, const R *theta, const R *phi,
          const int *orders, Packed &pk, const char *caller) {
    std::vector<int> stageOff(64, -1), stageCnt(64, 0);
    long long angle = 0;
    for (int j = 0; j < numJobs; ++j) {
        const PauliHamil &h = hamils[j];
        const int o = orders[j];
        if (h.numQubits < 1 || h.numQubits > kMaxLdsQubits || h.numSumTerms < 1 || o < 1 || o > 8 ||
            (o != 1 && (o & 1)) || !(times[j] == times[j])) {
            std::fprintf(stderr, "libquest_mi355x: %s: job %d: need 1..%d qubits, >= 1 term, order 1 or even <= 8\n",
                         caller, j, kMaxLdsQubits);
            return false;
        }
        if (stageOff[o] < 0) {
            stageOff[o] = (int)pk.stages.size();
            stageList(o, 1.0, pk.stages);
            stageCnt[o] = (int)pk.stages.size() - stageOff[o];
        }
        BatchJob b{};
        b.n = h.numQubits;
        b.numTerms = h.numSumTerms;
        b.termOff = (int)pk.terms.size();
        b.order = o;
        b.stageOff = stageOff[o];
        b.numStages = stageCnt[o];
        b.angleOff = (int)angle;
        b.time = times[j];
        for (int t = 0; t < h.numSumTerms; ++t) {
            PTerm pt{0u, 0u, 0, 0, h.termCoeffs[t]};
            for (int q = 0; q < h.numQubits; ++q) {
                const int c = h.pauliCodes[(size_t)t * h.numQubits + q];
                if (c == PAULI_X) pt.x |= 1u << q;
                if (c == PAULI_Y) { pt.x |= 1u << q; pt.z |= 1u << q; ++pt.ny; }
                if (c == PAULI_Z) pt.z |= 1u << q;
            }
            pk.terms.push_back(pt);
        }
        for (int q = 0; q < h.numQubits; ++q) {
            pk.theta.push_back(theta[angle + q]);
            pk.phi.push_back(phi[angle + q]);
        }
        angle += h.numQubits;
        pk.jobs.push_back(b);
    }
    return true;
}

template <typename V>
V *upload(const std::vector<V> &v) {
    V *d = static_cast<V *>(devAlloc(std::max<size_t>(1, v.size()) * sizeof(V)));
    if (!v.empty()) hostToDevice(d, v.data(), v.size() * sizeof(V));
    return d;
}

}  // namespace

}  // namespace qmx

extern "C" {

int qmxTrotterBatchEvolve(int numJobs, const PauliHamil *hamils, const qreal *times, const qreal *theta,
                          const qreal *phi, const int *orders, const int *reps, qreal *outReal, qreal *outImag) {
    using namespace qmx;
    runtimeInit();
    Packed pk;
    if (numJobs < 1 || !pack(numJobs, hamils, times, theta, phi, orders, pk, "qmxTrotterBatchEvolve")) return 0;
    long long off = 0;
    for (int j = 0; j < numJobs; ++j) {
        if (reps[j] < 1) {
            std::fprintf(stderr, "libquest_mi355x: qmxTrotterBatchEvolve: job %d: reps must be >= 1\n", j);
            return 0;
        }
        pk.jobs[j].reps = reps[j];
        pk.jobs[j].outOff = off;
        off += 1ll << pk.jobs[j].n;
    }
    int maxN = 1;
    for (auto &b : pk.jobs) maxN = std::max(maxN, b.n);
    const size_t lds = (size_t)(1u << maxN) * sizeof(double2);
    if (!allowDynamicLds(reinterpret_cast<const void *>(kTrotterEvolve), lds, "qmxTrotterBatchEvolve")) return 0;
    BatchJob *dJobs = upload(pk.jobs);
    PTerm *dTerms = upload(pk.terms);
    Stage *dStages = upload(pk.stages);
    double *dTh = upload(pk.theta), *dPh = upload(pk.phi);
    double2 *dOut = static_cast<double2 *>(devAlloc((size_t)off * sizeof(double2)));
    hipLaunchKernelGGL(kTrotterEvolve, dim3(numJobs), dim3(kBlock), lds, hstream(),
                       dJobs, dTerms, dStages, dTh, dPh, dOut);
    checkLaunch("qmxTrotterBatchEvolve");
    std::vector<double2> h((size_t)off);
    deviceToHost(h.data(), dOut, (size_t)off * sizeof(double2));
    for (long long i = 0; i < off; ++i) {
        outReal[i] = h[i].x;
        outImag[i] = h[i].y;
    }
    for (void *p : {(void *)dJobs, (void *)dTerms, (void *)dStages, (void *)dTh, (void *)dPh, (void *)dOut}) devFree(p);
    return 1;
}

int qmxTrotterBatchInfidelity(int numJobs, const PauliHamil *hamils, const qreal *times, const qreal *theta,
                              const qreal *phi, const int *orders, const int *reps, qreal *infidelity) {
    using namespace qmx;
    runtimeInit();
    Packed pk;
    if (numJobs < 1 || !pack(numJobs, hamils, times, theta, phi, orders, pk, "qmxTrotterBatchInfidelity")) return 0;
    int allN = 1;
    for (int j = 0; j < numJobs; ++j) {
        if (reps[j] < 1 || reps[j] > (1 << 28)) {
            std::fprintf(stderr, "libquest_mi355x: qmxTrotterBatchInfidelity: job %d: reps must be in 1..2^28\n", j);
            return 0;
        }
        pk.jobs[j].reps = reps[j];
        allN = std::max(allN, pk.jobs[j].n);
    }
    if (!allowDynamicLds(reinterpret_cast<const void *>(kTrotterInfidelity), (size_t)(1u << allN) * sizeof(double2),
                  "qmxTrotterBatchInfidelity"))
        return 0;
    PTerm *dTerms = upload(pk.terms);
    Stage *dStages = upload(pk.stages);
    double *dTh = upload(pk.theta), *dPh = upload(pk.phi);
    const size_t kSlotBudget = 1ull << 30;
    int first = 0;
    while (first < numJobs) {
        long long amps = 0;
        int last = first, maxN = 1;
        while (last < numJobs) {
            const long long need = 1ll << pk.jobs[last].n;
            if (last > first && (size_t)(amps + need) * sizeof(double2) > kSlotBudget) break;
            pk.jobs[last].outOff = amps;
            amps += need;
            maxN = std::max(maxN, pk.jobs[last].n);
            ++last;
        }
        const int count = last - first;
        std::vector<BatchJob> part(pk.jobs.begin() + first, pk.jobs.begin() + last);
        BatchJob *dJobs = upload(part);
        double2 *dSlot = static_cast<double2 *>(devAlloc((size_t)amps * sizeof(double2)));
        double *dInf = static_cast<double *>(devAlloc((size_t)count * sizeof(double)));
        hipLaunchKernelGGL(kTrotterInfidelity, dim3(count), dim3(kBlock), (size_t)(1u << maxN) * sizeof(double2),
                           hstream(), dJobs, dTerms, dStages, dTh, dPh, dSlot, dInf);
        checkLaunch("qmxTrotterBatchInfidelity");
        deviceToHost(infidelity + first, dInf, (size_t)count * sizeof(double));
        for (void *p : {(void *)dJobs, (void *)dSlot, (void *)dInf}) devFree(p);
        first = last;
    }
    for (void *p : {(void *)dTerms, (void *)dStages, (void *)dTh, (void *)dPh}) devFree(p);
    return 1;
}

int qmxTrotterBatchLadder(int numJobs, const PauliHamil *hamils, const qreal *times, const qreal *theta,
                          const qreal *phi, const int *orders, int maxLog2, qreal eps, qreal *infidelity,
                          int *passIndex) {
    using namespace qmx;
    runtimeInit();
    Packed pk;
    if (numJobs < 1 || maxLog2 < 0 || maxLog2 > 16 ||
        !pack(numJobs, hamils, times, theta, phi, orders, pk, "qmxTrotterBatchLadder"))
        return 0;
    const int stride = maxLog2 + 1;
    int allN = 1;
    for (auto &b : pk.jobs) allN = std::max(allN, b.n);
    if (!allowDynamicLds(reinterpret_cast<const void *>(kTrotterLadder), (size_t)(1u << allN) * sizeof(double2),
                  "qmxTrotterBatchLadder"))
        return 0;
    for (long long i = 0; i < (long long)numJobs * stride; ++i) infidelity[i] = NAN;
    PTerm *dTerms = upload(pk.terms);
    Stage *dStages = upload(pk.stages);
    double *dTh = upload(pk.theta), *dPh = upload(pk.phi);
    // launches of at most ~1 GiB of ring scratch
    const size_t kRingBudget = 1ull << 30;
    int first = 0;
    while (first < numJobs) {
        long long ringAmps = 0;
        int last = first;
        int maxN = 1;
        while (last < numJobs) {
            const long long need = 3ll << pk.jobs[last].n;
            if (last > first && (size_t)(ringAmps + need) * sizeof(double2) > kRingBudget) break;
            pk.jobs[last].outOff = ringAmps;
            ringAmps += need;
            maxN = std::max(maxN, pk.jobs[last].n);
            ++last;
        }
        const int count = last - first;
        std::vector<BatchJob> part(pk.jobs.begin() + first, pk.jobs.begin() + last);
        BatchJob *dJobs = upload(part);
        double2 *dRing = static_cast<double2 *>(devAlloc((size_t)ringAmps * sizeof(double2)));
        double *dInf = static_cast<double *>(devAlloc((size_t)count * stride * sizeof(double)));
        int *dPass = static_cast<int *>(devAlloc((size_t)count * sizeof(int)));
        hostToDevice(dInf, infidelity + (size_t)first * stride, (size_t)count * stride * sizeof(double));
        hipLaunchKernelGGL(kTrotterLadder, dim3(count), dim3(kBlock), (size_t)(1u << maxN) * sizeof(double2),
                           hstream(), dJobs, dTerms, dStages, dTh, dPh, dRing, maxLog2, (double)eps, dInf, dPass);
        checkLaunch("qmxTrotterBatchLadder");
        deviceToHost(infidelity + (size_t)first * stride, dInf, (size_t)count * stride * sizeof(double));
        deviceToHost(passIndex + first, dPass, (size_t)count * sizeof(int));
        for (void *p : {(void *)dJobs, (void *)dRing, (void *)dInf, (void *)dPass}) devFree(p);
        first = last;
    }
    for (void *p : {(void *)dTerms, (void *)dStages, (void *)dTh, (void *)dPh}) devFree(p);
    return 1;
}

}  // extern "C"
