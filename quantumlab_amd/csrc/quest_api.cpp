// QuEST 3.4 C-API front-end of libquest_mi355x.
//
// Mirrors the reference front-end (QuEST/src/QuEST.c) and its hardware-agnostic
// compositions (QuEST/src/QuEST_common.c): validate, then dispatch to the HIP
// backend; for density matrices every unitary U on the row qubits is followed
// by conj(U) on the column qubits (q + n), the Choi-Jamiolkowski form the
// reference uses (QuEST.c:3-16). Pauli-string rotations run as one fused sweep
// (qmx::applyPauliRotation) instead of QuEST's basis-change sandwich.
#include <sys/mman.h>
#include <sys/time.h>
#include <unistd.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "QuEST.h"
#include "mt19937.h"
#include "qmx_internal.h"
#include "quest_mi355x.h"
#include "quest_validation.h"

using qmx::Cplx;
using qmx::QuregDev;

// defined in qmx_runtime.hip: hipMalloc that returns nullptr on failure
void *qmxTryDevAlloc(size_t bytes);

#define API_TRY try {
#define API_CATCH(ret)                    \
    }                                     \
    catch (const qmx::InputError &) {     \
        return ret;                       \
    }
#define API_CATCH_VOID                    \
    }                                     \
    catch (const qmx::InputError &) {     \
        return;                           \
    }

namespace {

// A qreal array read by code that takes double: the array itself when qreal is
// double (QuEST_PREC 2), else an FP64 copy (QuEST_PREC 1). NULL stays NULL.
class Doubles {
    std::vector<double> copy_;
    const double *p_ = nullptr;

public:
    Doubles(const qreal *x, size_t n) {
        if (!x) return;
        if constexpr (std::is_same<qreal, double>::value) {
            p_ = reinterpret_cast<const double *>(x);
        } else {
            copy_.assign(x, x + n);
            p_ = copy_.data();
        }
    }
    const double *get() const { return p_; }
};

// Device copies of a DiagonalOp hold FP64 elements in both flavours (the
// kernels read double); deviceOperator.real / .imag point at them.
double *devDiag(qreal *p) { return reinterpret_cast<double *>(p); }

// createQureg / createDensityQureg store amplitudes at qreal's precision, as the
// reference's GPU backend does; createQuregEx chooses explicitly
constexpr int kDefaultPrec = sizeof(qreal) == sizeof(float) ? qmx::FP32 : qmx::FP64;

void diagToDevice(qreal *dev, const qreal *host, size_t n) {
    Doubles d(host, n);
    qmx::hostToDevice(devDiag(dev), d.get(), n * sizeof(double));
}

void diagToHost(qreal *host, qreal *dev, size_t n) {
    if constexpr (std::is_same<qreal, double>::value) {
        qmx::deviceToHost(host, devDiag(dev), n * sizeof(double));
    } else {
        std::vector<double> d(n);
        qmx::deviceToHost(d.data(), devDiag(dev), n * sizeof(double));
        for (size_t i = 0; i < n; ++i) host[i] = (qreal)d[i];
    }
}

uint64_t bitMask(const int *qubits, int n) {
    uint64_t m = 0;
    for (int i = 0; i < n; ++i) m |= 1ull << qubits[i];
    return m;
}

QuregDev *D(const Qureg &q) { return qmx::dev(q); }

Cplx conjC(Cplx z) { return Cplx{z.re, -z.im}; }

void m2ToCplx(const ComplexMatrix2 &u, Cplx out[4], bool conj = false) {
    double s = conj ? -1.0 : 1.0;
    out[0] = Cplx{u.real[0][0], s * u.imag[0][0]};
    out[1] = Cplx{u.real[0][1], s * u.imag[0][1]};
    out[2] = Cplx{u.real[1][0], s * u.imag[1][0]};
    out[3] = Cplx{u.real[1][1], s * u.imag[1][1]};
}

template <int DIM, typename M>
std::vector<Cplx> matToCplx(const M &u, bool conj) {
    std::vector<Cplx> v(DIM * DIM);
    double s = conj ? -1.0 : 1.0;
    for (int r = 0; r < DIM; ++r)
        for (int c = 0; c < DIM; ++c) v[r * DIM + c] = Cplx{u.real[r][c], s * u.imag[r][c]};
    return v;
}

std::vector<Cplx> matNToCplx(const ComplexMatrixN &u, bool conj) {
    int dim = 1 << u.numQubits;
    std::vector<Cplx> v((size_t)dim * dim);
    double s = conj ? -1.0 : 1.0;
    for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c) v[(size_t)r * dim + c] = Cplx{u.real[r][c], s * u.imag[r][c]};
    return v;
}

// compact unitary [[a, -conj(b)], [b, conj(a)]]
void compactToCplx(Complex a, Complex b, Cplx out[4]) {
    out[0] = Cplx{a.real, a.imag};
    out[1] = Cplx{-b.real, b.imag};
    out[2] = Cplx{b.real, b.imag};
    out[3] = Cplx{a.real, -a.imag};
}

// QuEST_common.c:117-124
void rotationPair(double angle, Vector axis, Complex *alpha, Complex *beta) {
    double mag = std::sqrt(axis.x * axis.x + axis.y * axis.y + axis.z * axis.z);
    double ux = axis.x / mag, uy = axis.y / mag, uz = axis.z / mag;
    alpha->real = std::cos(angle / 2.0);
    alpha->imag = -std::sin(angle / 2.0) * uz;
    beta->real = std::sin(angle / 2.0) * uy;
    beta->imag = -std::sin(angle / 2.0) * ux;
}

// 2x2 on target with controls; density: conj on shifted qubits
// a one-qubit matrix under one control as a 4x4 matrix on (target, control):
// index bit 0 is the target, bit 1 the control; active when the control bit
// equals `on`
void controlled4(const Cplx u[4], int on, Cplx out[16]) {
    for (int e = 0; e < 16; ++e) out[e] = Cplx{(e % 5 == 0) ? 1.0 : 0.0, 0};
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) out[(on * 2 + r) * 4 + (on * 2 + c)] = u[r * 2 + c];
}

void gate2(const Qureg &q, int target, uint64_t ctrl, uint64_t flip, const Cplx u[4], const Cplx uc[4]) {
    if (q.isDensityMatrix && ctrl == 0) {
        qmx::densUnitary(D(q), &target, 1, u, uc);
        return;
    }
    if (q.isDensityMatrix && __builtin_popcountll(ctrl) == 1) {
        int t[2] = {target, __builtin_ctzll(ctrl)};
        const int on = (flip & ctrl) ? 0 : 1;
        Cplx u4[16], uc4[16];
        controlled4(u, on, u4);
        controlled4(uc, on, uc4);
        qmx::densUnitary(D(q), t, 2, u4, uc4);
        return;
    }
    qmx::applyMatrix2(D(q), target, ctrl, flip, u);
    if (q.isDensityMatrix) {
        int s = q.numQubitsRepresented;
        qmx::applyMatrix2(D(q), target + s, ctrl << s, flip << s, uc);
    }
}

void gate2Conj(const Qureg &q, int target, uint64_t ctrl, uint64_t flip, const Cplx u[4]) {
    Cplx uc[4];
    for (int i = 0; i < 4; ++i) uc[i] = conjC(u[i]);
    gate2(q, target, ctrl, flip, u, uc);
}

void phaseAllOnes(const Qureg &q, uint64_t mask, Cplx term) {
    if (q.isDensityMatrix && mask && __builtin_popcountll(mask) <= 2) {
        int t[2], k = 0;
        for (uint64_t m = mask; m; m &= m - 1) t[k++] = __builtin_ctzll(m);
        const int d = 1 << k;
        Cplx u[16] = {}, uc[16] = {};
        for (int i = 0; i < d; ++i) {
            u[i * d + i] = (i == d - 1) ? term : Cplx{1, 0};
            uc[i * d + i] = conjC(u[i * d + i]);
        }
        qmx::densUnitary(D(q), t, k, u, uc);
        return;
    }
    qmx::applyPhaseOnAllOnes(D(q), mask, term);
    if (q.isDensityMatrix) qmx::applyPhaseOnAllOnes(D(q), mask << q.numQubitsRepresented, conjC(term));
}

void matrixN(const Qureg &q, uint64_t ctrl, const int *targs, int k, const std::vector<Cplx> &u,
             const std::vector<Cplx> &uc) {
    if (q.isDensityMatrix && ctrl == 0 && k <= 2) {
        qmx::densUnitary(D(q), targs, k, u.data(), uc.data());
        return;
    }
    if (q.isDensityMatrix && k == 1 && __builtin_popcountll(ctrl) == 1) {
        int t[2] = {targs[0], __builtin_ctzll(ctrl)};
        Cplx u4[16], uc4[16];
        controlled4(u.data(), 1, u4);
        controlled4(uc.data(), 1, uc4);
        qmx::densUnitary(D(q), t, 2, u4, uc4);
        return;
    }
    qmx::applyMatrixN(D(q), ctrl, targs, k, u.data());
    if (q.isDensityMatrix) {
        int s = q.numQubitsRepresented;
        std::vector<int> t(targs, targs + k);
        for (int &x : t) x += s;
        qmx::applyMatrixN(D(q), ctrl << s, t.data(), k, uc.data());
    }
}

// Pauli-string rotation exp(-i angle/2 P) on a state vector (no density doubling)
void pauliRotRaw(QuregDev *d, uint64_t ctrl, const int *targs, const enum pauliOpType *paulis, int n, int shift,
                 double angle) {
    uint64_t x = 0, z = 0;
    int ny = 0;
    for (int t = 0; t < n; ++t) {
        uint64_t b = 1ull << (targs[t] + shift);
        switch (paulis[t]) {
            case PAULI_X: x |= b; break;
            case PAULI_Y: x |= b; z |= b; ++ny; break;
            case PAULI_Z: z |= b; break;
            default: break;
        }
    }
    if (x == 0 && z == 0) return;  // identity string: QuEST skips the rotation (mask == 0)
    qmx::applyPauliRotation(d, ctrl, x, z, ny, angle);
}

// QuEST.c:691-709 semantics; density: conj(exp(-iθ/2 P)) = exp(-i θ'/2 P) on the
// column qubits with θ' = -θ (even #Y) or +θ (odd #Y)
void pauliRot(const Qureg &q, uint64_t ctrl, const int *targs, const enum pauliOpType *paulis, int n, double angle) {
    pauliRotRaw(D(q), ctrl, targs, paulis, n, 0, angle);
    if (q.isDensityMatrix) {
        int ny = 0;
        for (int t = 0; t < n; ++t) ny += paulis[t] == PAULI_Y;
        int s = q.numQubitsRepresented;
        pauliRotRaw(D(q), ctrl << s, targs, paulis, n, s, (ny & 1) ? angle : -angle);
    }
}

// ---- Trotter-Suzuki (QuEST_common.c:762-844) --------------------------------
// QuEST's recursion, recording the rotations of one repetition in QuEST's order
// (density matrices: each rotation on the rows, then its conjugate on the
// columns); the repetitions then run as fused sweeps (applyPauliSequence).
void pushPauliTerm(std::vector<qmx::PauliExp> &seq, const Qureg &q, const enum pauliOpType *codes, int n,
                   double angle) {
    uint64_t x = 0, z = 0;
    int ny = 0;
    for (int t = 0; t < n; ++t) {
        const uint64_t b = 1ull << t;
        switch (codes[t]) {
            case PAULI_X: x |= b; break;
            case PAULI_Y: x |= b; z |= b; ++ny; break;
            case PAULI_Z: z |= b; break;
            default: break;
        }
    }
    seq.push_back(qmx::PauliExp{x, z, ny, angle});
    if (q.isDensityMatrix) {
        const int s = q.numQubitsRepresented;
        seq.push_back(qmx::PauliExp{x << s, z << s, ny, (ny & 1) ? angle : -angle});
    }
}

void exponentiatedHamil(std::vector<qmx::PauliExp> &seq, const Qureg &q, const PauliHamil &h, double fac,
                        bool reverse) {
    for (int i = 0; i < h.numSumTerms; ++i) {
        int t = reverse ? h.numSumTerms - 1 - i : i;
        double angle = 2 * fac * h.termCoeffs[t];
        pushPauliTerm(seq, q, &h.pauliCodes[(size_t)t * h.numQubits], h.numQubits, angle);
    }
}

void symmetrizedTrotter(std::vector<qmx::PauliExp> &seq, const Qureg &q, const PauliHamil &h, double time,
                        int order) {
    if (order == 1) {
        exponentiatedHamil(seq, q, h, time, false);
    } else if (order == 2) {
        exponentiatedHamil(seq, q, h, time / 2., false);
        exponentiatedHamil(seq, q, h, time / 2., true);
    } else {
        double p = 1. / (4 - std::pow(4, 1. / (order - 1)));
        int lower = order - 2;
        symmetrizedTrotter(seq, q, h, p * time, lower);
        symmetrizedTrotter(seq, q, h, p * time, lower);
        symmetrizedTrotter(seq, q, h, (1 - 4 * p) * time, lower);
        symmetrizedTrotter(seq, q, h, p * time, lower);
        symmetrizedTrotter(seq, q, h, p * time, lower);
    }
}

// ---- Kraus superoperators (QuEST_common.c:590-700) -----------------------------
// superOp = Σ_n conj(K_n) ⊗ K_n, indexed [i*d + k][j*d + l] = conj(K_ij) K_kl
template <typename GetK>
std::vector<Cplx> krausSuperOp(int numOps, int d, GetK K) {
    std::vector<Cplx> s((size_t)d * d * d * d, Cplx{0, 0});
    const int D2 = d * d;
    for (int n = 0; n < numOps; ++n)
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j)
                for (int k = 0; k < d; ++k)
                    for (int l = 0; l < d; ++l) {
                        Cplx a = K(n, i, j), b = K(n, k, l);
                        Cplx &e = s[(size_t)(i * d + k) * D2 + (j * d + l)];
                        e.re += a.re * b.re + a.im * b.im;
                        e.im += a.re * b.im - a.im * b.re;
                    }
    return s;
}

void applySuperOp(const Qureg &q, const int *targets, int numTargets, const std::vector<Cplx> &sup) {
    qmx::densSuperOp(D(q), targets, numTargets, sup.data());
}

// ---- host mirror allocation ----------------------------------------------------
void *mirrorAlloc(size_t bytes) {
    void *p = mmap(nullptr, bytes ? bytes : 8, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE,
                   -1, 0);
    return p == MAP_FAILED ? nullptr : p;
}

Qureg makeQureg(int numQubitsInSV, int isDensity, int numRep, int prec, const char *caller) {
    qmx::runtimeInit();
    Qureg q;
    std::memset(&q, 0, sizeof q);
    const long long numAmps = 1LL << numQubitsInSV;
    auto *d = new QuregDev;
    d->magic = qmx::kMagic;
    d->prec = prec;
    d->numQubits = numQubitsInSV;
    d->isDensity = isDensity;
    d->numQubitsRep = numRep;
    d->numAmps = numAmps;
    d->hostBytes = (size_t)numAmps * sizeof(qreal);
    d->amps = nullptr;
    qreal *re = static_cast<qreal *>(mirrorAlloc(d->hostBytes));
    qreal *im = static_cast<qreal *>(mirrorAlloc(d->hostBytes));
    if (!re || !im) {
        if (re) munmap(re, d->hostBytes);
        if (im) munmap(im, d->hostBytes);
        delete d;
        qv::check(false, qv::E_QUREG_NOT_ALLOCATED, caller);
    }
    size_t bytes = (size_t)numAmps * qmx::ampBytes(prec);
    d->amps = qmx::poolAlloc(bytes);
    if (!d->amps) {
        munmap(re, d->hostBytes);
        munmap(im, d->hostBytes);
        delete d;
        qv::check(false, qv::E_QUREG_NOT_ALLOCATED_ON_GPU, caller);
    }
    q.isDensityMatrix = isDensity;
    q.numQubitsRepresented = numRep;
    q.numQubitsInStateVec = numQubitsInSV;
    q.numAmpsPerChunk = numAmps;
    q.numAmpsTotal = numAmps;
    q.chunkId = 0;
    q.numChunks = 1;
    q.stateVec.real = re;
    q.stateVec.imag = im;
    q.deviceStateVec.real = static_cast<qreal *>(d->amps);
    q.deviceStateVec.imag = nullptr;
    q.firstLevelReduction = reinterpret_cast<qreal *>(d);
    q.secondLevelReduction = nullptr;
    q.qasmLog = static_cast<QASMLogger *>(std::calloc(1, sizeof(QASMLogger)));
    qmx::initClassical(d, 0);
    return q;
}

void getDefaultSeedKey(unsigned long *key) {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    double ms = (tv.tv_sec) * 1000 + (tv.tv_usec) / 1000;
    key[0] = (unsigned long)ms;
    key[1] = (unsigned long)getpid();
}

int generateMeasurementOutcome(double zeroProb, double *outcomeProb) {
    // QuEST_common.c:165-180: no draw when the outcome is (numerically) certain
    int outcome;
    if (zeroProb < REAL_EPS)
        outcome = 1;
    else if (1 - zeroProb < REAL_EPS)
        outcome = 0;
    else
        outcome = (qmx::mt::nextReal1() > zeroProb);
    *outcomeProb = (outcome == 0) ? zeroProb : 1 - zeroProb;
    return outcome;
}

double probOfOutcome(const Qureg &q, int qubit, int outcome) {
    double p0 = q.isDensityMatrix ? qmx::densProbOfZero(D(q), qubit) : qmx::probOfZero(D(q), qubit);
    return outcome == 1 ? 1.0 - p0 : p0;
}

void collapseKnown(const Qureg &q, int qubit, int outcome, double prob) {
    if (q.isDensityMatrix)
        qmx::densCollapse(D(q), qubit, outcome, prob);
    else
        qmx::collapse(D(q), qubit, outcome, prob);
}

void pauliProd(const Qureg &ws, const int *targs, const enum pauliOpType *codes, int n) {
    // QuEST_common.c:499-511: plain X/Y/Z on the (row) qubits
    static const Cplx Y[4] = {{0, 0}, {0, -1}, {0, 1}, {0, 0}};
    for (int i = 0; i < n; ++i) {
        if (codes[i] == PAULI_X) qmx::applyNot(D(ws), 0, 1ull << targs[i]);
        if (codes[i] == PAULI_Y) qmx::applyMatrix2(D(ws), targs[i], 0, 0, Y);
        if (codes[i] == PAULI_Z) qmx::applyPhaseOnAllOnes(D(ws), 1ull << targs[i], Cplx{-1, 0});
    }
}

double expecPauliProd(const Qureg &q, const int *targs, const enum pauliOpType *codes, int n, const Qureg &ws) {
    qmx::copyState(D(ws), D(q));
    pauliProd(ws, targs, codes, n);
    if (q.isDensityMatrix) return qmx::densTotalProb(D(ws));
    return qmx::innerProduct(D(ws), D(q)).re;
}

// the phase-function passes of the reference: rows, then for a density matrix
// the same function conjugated on the column bits (QuEST.c:732-870)
void phaseFuncPasses(const Qureg &q, const qmx::PhaseFuncArgs &a) {
    qmx::applyPhaseFunction(D(q), a, 0, false);
    if (q.isDensityMatrix) qmx::applyPhaseFunction(D(q), a, q.numQubitsRepresented, true);
}

void hadamardRaw(const Qureg &q, int t) {
    const double r = 1 / std::sqrt(2.0);
    const Cplx H[4] = {{r, 0}, {r, 0}, {r, 0}, {-r, 0}};
    gate2(q, t, 0, 0, H, H);
}

void swapRaw(const Qureg &q, int a, int b) {
    if (q.isDensityMatrix) {
        Cplx sw[16] = {};
        sw[0] = sw[15] = Cplx{1, 0};
        sw[1 * 4 + 2] = sw[2 * 4 + 1] = Cplx{1, 0};
        int t[2] = {a, b};
        qmx::densUnitary(D(q), t, 2, sw, sw);
    } else {
        qmx::applySwap(D(q), a, b);
    }
}

// QuEST_common.c:846-908: from the top qubit down, H then the controlled
// phases onto every lower qubit as one SCALED_PRODUCT phase function of
// (lower qubits, this qubit) with coefficient pi / 2^q; then reverse the order
void qftOn(const Qureg &q, const int *qubits, int n) {
    for (int k = n - 1; k >= 0; --k) {
        hadamardRaw(q, qubits[k]);
        if (k == 0) break;
        const int perReg[2] = {k, 1};
        const double param = M_PI / (double)(1LL << k);
        qmx::PhaseFuncArgs a{};
        a.numRegs = 2;
        a.encoding = UNSIGNED;
        a.mode = 1;
        a.funcCode = SCALED_PRODUCT;
        a.qubits = qubits;
        a.numQubitsPerReg = perReg;
        a.params = &param;
        a.numParams = 1;
        phaseFuncPasses(q, a);
    }
    for (int i = 0; i < n / 2; ++i) swapRaw(q, qubits[i], qubits[n - 1 - i]);
}

DiagonalOp emptyDiagonalOp() {
    DiagonalOp op;
    std::memset(&op, 0, sizeof op);
    return op;
}

}  // namespace

extern "C" {

/* ---- environment -------------------------------------------------------------- */

QuESTEnv createQuESTEnv(void) {
    qmx::runtimeInit();
    QuESTEnv env;
    env.rank = 0;
    env.numRanks = 1;
    env.seeds = nullptr;
    env.numSeeds = 0;
    seedQuESTDefault(&env);
    return env;
}

void destroyQuESTEnv(QuESTEnv env) { std::free(env.seeds); }

void syncQuESTEnv(QuESTEnv) { qmx::deviceSync(); }

int syncQuESTSuccess(int successCode) { return successCode; }

void reportQuESTEnv(QuESTEnv env) {
    char buf[512];
    qmx::deviceDescription(buf, sizeof buf);
    std::printf("EXECUTION ENVIRONMENT:\n");
    std::printf("Running locally on one node with GPU (HIP, libquest_mi355x)\n");
    std::printf("Number of ranks is %d\n", env.numRanks);
    std::printf("%s\n", buf);
    std::printf("Precision: size of qreal is %ld bytes\n", (long)sizeof(qreal));
}

void getEnvironmentString(QuESTEnv env, char str[200]) {
    (void)env;
    char buf[160];
    qmx::deviceDescription(buf, sizeof buf);
    std::snprintf(str, 200, "HIP=1 OpenMP=0 MPI=0 threads=1 ranks=1 %s", buf);
}

void seedQuEST(QuESTEnv *env, unsigned long *seedArray, int numSeeds) {
    std::free(env->seeds);
    env->seeds = static_cast<unsigned long *>(std::malloc(numSeeds * sizeof(unsigned long)));
    for (int i = 0; i < numSeeds; ++i) env->seeds[i] = seedArray[i];
    env->numSeeds = numSeeds;
    qmx::mt::initByArray(seedArray, numSeeds);
}

void seedQuESTDefault(QuESTEnv *env) {
    unsigned long keys[2];
    getDefaultSeedKey(keys);
    seedQuEST(env, keys, 2);
}

void getQuESTSeeds(QuESTEnv env, unsigned long **seeds, int *numSeeds) {
    *seeds = env.seeds;
    *numSeeds = env.numSeeds;
}

/* ---- registers ------------------------------------------------------------------ */

Qureg createQuregEx(int numQubits, QuESTEnv env, int precision) {
    Qureg empty;
    std::memset(&empty, 0, sizeof empty);
    API_TRY
    qv::precision(precision, __func__);
    qv::numQubitsInQureg(numQubits, __func__, env.numRanks);
    return makeQureg(numQubits, 0, numQubits, precision, __func__);
    API_CATCH(empty)
}

Qureg createDensityQuregEx(int numQubits, QuESTEnv env, int precision) {
    Qureg empty;
    std::memset(&empty, 0, sizeof empty);
    API_TRY
    qv::precision(precision, __func__);
    qv::numQubitsInQureg(2 * numQubits, __func__, env.numRanks);
    return makeQureg(2 * numQubits, 1, numQubits, precision, __func__);
    API_CATCH(empty)
}

Qureg createQureg(int numQubits, QuESTEnv env) {
    Qureg empty;
    std::memset(&empty, 0, sizeof empty);
    API_TRY
    qv::numQubitsInQureg(numQubits, __func__, env.numRanks);
    return makeQureg(numQubits, 0, numQubits, kDefaultPrec, __func__);
    API_CATCH(empty)
}

Qureg createDensityQureg(int numQubits, QuESTEnv env) {
    Qureg empty;
    std::memset(&empty, 0, sizeof empty);
    API_TRY
    qv::numQubitsInQureg(2 * numQubits, __func__, env.numRanks);
    return makeQureg(2 * numQubits, 1, numQubits, kDefaultPrec, __func__);
    API_CATCH(empty)
}

Qureg createCloneQureg(Qureg qureg, QuESTEnv env) {
    Qureg empty;
    std::memset(&empty, 0, sizeof empty);
    API_TRY
    (void)env;
    Qureg q = makeQureg(qureg.numQubitsInStateVec, qureg.isDensityMatrix, qureg.numQubitsRepresented, D(qureg)->prec,
                        __func__);
    qmx::copyState(D(q), D(qureg));
    return q;
    API_CATCH(empty)
}

void destroyQureg(Qureg qureg, QuESTEnv env) {
    (void)env;
    QuregDev *d = D(qureg);
    if (!d || d->magic != qmx::kMagic) return;
    qmx::densDropPending(d);
    qmx::poolFree(d->amps, (size_t)d->numAmps * qmx::ampBytes(d->prec));
    munmap(qureg.stateVec.real, d->hostBytes);
    munmap(qureg.stateVec.imag, d->hostBytes);
    std::free(qureg.qasmLog);
    d->magic = 0;
    delete d;
}

int qmxQuregPrecision(Qureg qureg) { return D(qureg)->prec; }

ComplexMatrixN createComplexMatrixN(int numQubits) {
    ComplexMatrixN m;
    std::memset(&m, 0, sizeof m);
    API_TRY
    qv::numQubitsInMatrix(numQubits, __func__);
    int rows = 1 << numQubits;
    m.numQubits = numQubits;
    m.real = static_cast<qreal **>(std::malloc(rows * sizeof(qreal *)));
    m.imag = static_cast<qreal **>(std::malloc(rows * sizeof(qreal *)));
    for (int r = 0; r < rows; ++r) {
        m.real[r] = static_cast<qreal *>(std::calloc(rows, sizeof(qreal)));
        m.imag[r] = static_cast<qreal *>(std::calloc(rows, sizeof(qreal)));
    }
    qv::matrixInit(m, __func__);
    return m;
    API_CATCH(m)
}

void destroyComplexMatrixN(ComplexMatrixN m) {
    API_TRY
    qv::matrixInit(m, __func__);
    int rows = 1 << m.numQubits;
    for (int r = 0; r < rows; ++r) {
        std::free(m.real[r]);
        std::free(m.imag[r]);
    }
    std::free(m.real);
    std::free(m.imag);
    API_CATCH_VOID
}

PauliHamil createPauliHamil(int numQubits, int numSumTerms) {
    PauliHamil h;
    std::memset(&h, 0, sizeof h);
    API_TRY
    qv::hamilParams(numQubits, numSumTerms, __func__);
    h.numQubits = numQubits;
    h.numSumTerms = numSumTerms;
    h.termCoeffs = static_cast<qreal *>(std::malloc(numSumTerms * sizeof(qreal)));
    h.pauliCodes = static_cast<enum pauliOpType *>(std::malloc((size_t)numQubits * numSumTerms * sizeof(enum pauliOpType)));
    for (int i = 0; i < numQubits * numSumTerms; ++i) h.pauliCodes[i] = PAULI_I;
    return h;
    API_CATCH(h)
}

void destroyPauliHamil(PauliHamil h) {
    std::free(h.termCoeffs);
    std::free(h.pauliCodes);
}

// ---- PauliHamil text files -------------------------------------------------------
//
// The format read by QuEST's createPauliHamilFromFile (QuEST.c:1459-1541): one
// term per line, "coeff code_0 code_1 ... code_{n-1}", codes 0-3 for I, X, Y, Z.
// The rules below give the same register width, term count, values and
// errors as that reader:
//  - width: the fields of the first line, split on ' ' (tabs and '\r' belong to
//    a field), minus the coefficient; a line that begins with ' ' counts one
//    field more (an empty coefficient field before that space);
//  - terms: the non-empty lines ('\n'-separated; a line of blanks counts);
//  - values: the whole file is one stream of numbers separated by any white
//    space, read in order (coefficient, then n codes, per term), each number
//    as fscanf's "%lf" / "%d" take it (the longest prefix of a number; glibc
//    ignores a dangling exponent such as "1e"); the first that does not
//    convert is an error.
namespace {

struct HamilText {
    std::string text;
    size_t at = 0;
    void skipSpace() {
        while (at < text.size() && std::isspace((unsigned char)text[at])) ++at;
    }
    // The input item of a "%lf" / "%f" conversion as glibc's fscanf takes it:
    // the longest prefix of a number (C11 7.21.6.2p9), so a dangling exponent
    // ("1e", "2.5E-", "0x1p") is read and ignored; it converts when the item
    // holds a mantissa digit or is inf / infinity / nan (no "(chars)" suffix).
    // Returns the item's length, 0 when the conversion fails.
    static size_t realItem(const char *s) {
        auto ci = [](const char *a, const char *word) {
            for (; *word; ++a, ++word)
                if (std::tolower((unsigned char)*a) != *word) return false;
            return true;
        };
        auto digits = [](const char *a, size_t i, bool hex, size_t *count) {
            while (hex ? std::isxdigit((unsigned char)a[i]) : std::isdigit((unsigned char)a[i])) ++i, ++*count;
            return i;
        };
        size_t i = (s[0] == '+' || s[0] == '-') ? 1 : 0;
        if (ci(s + i, "infinity")) return i + 8;
        if (ci(s + i, "inf") || ci(s + i, "nan")) return i + 3;
        const bool hex = s[i] == '0' && (s[i + 1] == 'x' || s[i + 1] == 'X');
        if (hex) i += 2;
        size_t mantissa = 0;
        i = digits(s, i, hex, &mantissa);
        if (s[i] == '.') i = digits(s, i + 1, hex, &mantissa);
        if (mantissa == 0) return 0;
        const char e = (char)std::tolower((unsigned char)s[i]);
        if (e == (hex ? 'p' : 'e')) {
            ++i;
            if (s[i] == '+' || s[i] == '-') ++i;
            size_t expDigits = 0;
            i = digits(s, i, false, &expDigits);
        }
        return i;
    }
    bool real(qreal *out) {
        skipSpace();
        const char *begin = text.c_str() + at;
        const size_t len = realItem(begin);
        if (len == 0) return false;
        const std::string item(begin, len);  // strtod of the item ignores a dangling exponent
        *out = sizeof(qreal) == sizeof(float) ? (qreal)std::strtof(item.c_str(), nullptr)
                                               : (qreal)std::strtod(item.c_str(), nullptr);
        at += len;
        return true;
    }
    bool integer(int *out) {
        skipSpace();
        const char *begin = text.c_str() + at;
        char *end = nullptr;
        const long v = std::strtol(begin, &end, 10);
        if (end == begin) return false;
        at += (size_t)(end - begin);
        *out = (int)v;
        return true;
    }
};

int hamilFileWidth(const std::string &text) {
    const std::string first = text.substr(0, text.find('\n'));
    int fields = 0;
    bool inField = false;
    for (char ch : first) {
        const bool blank = ch == ' ';
        if (!blank && !inField) ++fields;
        inField = !blank;
    }
    if (!first.empty() && first[0] == ' ') ++fields;
    return fields - 1;
}

int hamilFileTerms(const std::string &text) {
    int terms = 0;
    size_t lineStart = 0;
    while (lineStart <= text.size()) {
        size_t nl = text.find('\n', lineStart);
        if (nl == std::string::npos) nl = text.size();
        terms += nl > lineStart;
        lineStart = nl + 1;
    }
    return terms;
}

}  // namespace

PauliHamil createPauliHamilFromFile(char *fn) {
    PauliHamil h;
    std::memset(&h, 0, sizeof h);
    API_TRY
    HamilText in;
    {
        FILE *file = std::fopen(fn, "rb");
        qv::fileOpened(file != nullptr, fn, __func__);
        char buf[65536];
        size_t got;
        while ((got = std::fread(buf, 1, sizeof buf, file)) > 0) in.text.append(buf, got);
        std::fclose(file);
    }
    const int numQubits = hamilFileWidth(in.text);
    const int numTerms = hamilFileTerms(in.text);
    qv::checkFmt(numQubits > 0 && numTerms > 0, qv::E_INVALID_PAULI_HAMIL_FILE_PARAMS, __func__, fn);
    PauliHamil parsed = createPauliHamil(numQubits, numTerms);
    try {
        for (int t = 0; t < numTerms; ++t) {
            qv::checkFmt(in.real(&parsed.termCoeffs[t]), qv::E_CANNOT_PARSE_PAULI_HAMIL_FILE_COEFF, __func__, fn);
            for (int q = 0; q < numQubits; ++q) {
                int code = 0;
                qv::checkFmt(in.integer(&code), qv::E_CANNOT_PARSE_PAULI_HAMIL_FILE_PAULI, __func__, fn);
                qv::checkFmtCode(qv::validPauliCode(code), qv::E_INVALID_PAULI_HAMIL_FILE_PAULI_CODE, __func__, fn,
                                 code);
                parsed.pauliCodes[t * numQubits + q] = (enum pauliOpType)code;
            }
        }
    } catch (...) {  // also a caller's throwing error hook
        destroyPauliHamil(parsed);
        throw;
    }
    return parsed;
    API_CATCH(h)
}

void initPauliHamil(PauliHamil hamil, qreal *coeffs, enum pauliOpType *codes) {
    API_TRY
    qv::hamilParams(hamil.numQubits, hamil.numSumTerms, __func__);
    qv::pauliCodes(codes, hamil.numSumTerms * hamil.numQubits, __func__);
    int i = 0;
    for (int t = 0; t < hamil.numSumTerms; ++t) {
        hamil.termCoeffs[t] = coeffs[t];
        for (int q = 0; q < hamil.numQubits; ++q, ++i) hamil.pauliCodes[i] = codes[i];
    }
    API_CATCH_VOID
}

void reportPauliHamil(PauliHamil hamil) {
    API_TRY
    qv::pauliHamil(hamil, __func__);
    for (int t = 0; t < hamil.numSumTerms; ++t) {
        std::printf(REAL_QASM_FORMAT, hamil.termCoeffs[t]);
        std::printf("\t");
        for (int q = 0; q < hamil.numQubits; ++q) std::printf("%d ", (int)hamil.pauliCodes[q + t * hamil.numQubits]);
        std::printf("\n");
    }
    API_CATCH_VOID
}

// QuEST_common.c:216-232: CSV of the host mirror
void reportState(Qureg qureg) {
    char filename[100];
    std::snprintf(filename, sizeof filename, "state_rank_%d.csv", qureg.chunkId);
    FILE *f = std::fopen(filename, "w");
    if (!f) return;
    std::fprintf(f, "real, imag\n");
    for (long long i = 0; i < qureg.numAmpsPerChunk; ++i)
        std::fprintf(f, REAL_STRING_FORMAT ", " REAL_STRING_FORMAT "\n", qureg.stateVec.real[i], qureg.stateVec.imag[i]);
    std::fclose(f);
}

void reportStateToScreen(Qureg qureg, QuESTEnv env, int reportRank) {
    (void)env;
    if (qureg.numQubitsInStateVec > 5) {
        std::printf("Error: reportStateToScreen will not print output for systems of more than 5 qubits.\n");
        return;
    }
    copyStateFromGPU(qureg);
    std::printf(reportRank ? "Reporting state from rank %d [\n" : "Reporting state [\n", qureg.chunkId);
    std::printf("real, imag\n");
    for (long long i = 0; i < qureg.numAmpsPerChunk; ++i)
        std::printf(REAL_STRING_FORMAT ", " REAL_STRING_FORMAT "\n", qureg.stateVec.real[i], qureg.stateVec.imag[i]);
    std::printf("]\n");
}

void reportQuregParams(Qureg qureg) {
    long long numAmps = 1LL << qureg.numQubitsInStateVec;
    std::printf("QUBITS:\n");
    std::printf("Number of qubits is %d.\n", qureg.numQubitsInStateVec);
    std::printf("Number of amps is %lld.\n", numAmps);
    std::printf("Number of amps per rank is %lld.\n", numAmps / qureg.numChunks);
}

int getNumQubits(Qureg qureg) { return qureg.numQubitsRepresented; }

long long int getNumAmps(Qureg qureg) {
    API_TRY
    qv::stateVecQureg(qureg, __func__);
    return qureg.numAmpsTotal;
    API_CATCH(0)
}

/* ---- initialisation ----------------------------------------------------------------- */

void initBlankState(Qureg qureg) { qmx::initBlank(D(qureg)); }

void initZeroState(Qureg qureg) { qmx::initClassical(D(qureg), 0); }

void initPlusState(Qureg qureg) {
    if (qureg.isDensityMatrix)
        qmx::initUniform(D(qureg), 1.0 / (double)(1LL << qureg.numQubitsRepresented));
    else
        qmx::initUniform(D(qureg), 1.0 / std::sqrt((double)qureg.numAmpsTotal));
}

void initClassicalState(Qureg qureg, long long int stateInd) {
    API_TRY
    qv::stateIndex(qureg, stateInd, __func__);
    if (qureg.isDensityMatrix)
        qmx::initClassical(D(qureg), ((1LL << qureg.numQubitsRepresented) + 1) * stateInd);
    else
        qmx::initClassical(D(qureg), stateInd);
    API_CATCH_VOID
}

void initPureState(Qureg qureg, Qureg pure) {
    API_TRY
    qv::secondQuregStateVec(pure, __func__);
    qv::matchingQuregDims(qureg, pure, __func__);
    if (qureg.isDensityMatrix)
        qmx::initPureDensity(D(qureg), D(pure));
    else
        qmx::copyState(D(qureg), D(pure));
    API_CATCH_VOID
}

void initDebugState(Qureg qureg) { qmx::initDebug(D(qureg)); }

void initStateFromAmps(Qureg qureg, qreal *reals, qreal *imags) {
    qmx::uploadAmps(D(qureg), 0, reals, imags, qureg.numAmpsTotal);
}

void setAmps(Qureg qureg, long long int startInd, qreal *reals, qreal *imags, long long int numAmps) {
    API_TRY
    qv::stateVecQureg(qureg, __func__);
    qv::numAmps(qureg, startInd, numAmps, __func__);
    qmx::uploadAmps(D(qureg), startInd, reals, imags, numAmps);
    API_CATCH_VOID
}

void setDensityAmps(Qureg qureg, long long int startRow, long long int startCol, qreal *reals, qreal *imags,
                    long long int numAmps) {
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    qv::numDensityAmps(qureg, startRow, startCol, numAmps, __func__);
    long long start = startRow + startCol * (1LL << qureg.numQubitsRepresented);
    qmx::uploadAmps(D(qureg), start, reals, imags, numAmps);
    API_CATCH_VOID
}

void cloneQureg(Qureg targetQureg, Qureg copyQureg) {
    API_TRY
    qv::matchingQuregTypes(targetQureg, copyQureg, __func__);
    qv::matchingQuregDims(targetQureg, copyQureg, __func__);
    qmx::copyState(D(targetQureg), D(copyQureg));
    API_CATCH_VOID
}

/* ---- host mirror ---------------------------------------------------------------------- */

void copyStateToGPU(Qureg qureg) {
    qmx::uploadAmps(D(qureg), 0, qureg.stateVec.real, qureg.stateVec.imag, qureg.numAmpsTotal);
}

void copyStateFromGPU(Qureg qureg) {
    qmx::downloadAmps(D(qureg), 0, qureg.stateVec.real, qureg.stateVec.imag, qureg.numAmpsTotal);
}

void copySubstateToGPU(Qureg qureg, long long int startInd, long long int numAmps) {
    API_TRY
    qv::numAmps(qureg, startInd, numAmps, __func__);
    qmx::uploadAmps(D(qureg), startInd, qureg.stateVec.real + startInd, qureg.stateVec.imag + startInd, numAmps);
    API_CATCH_VOID
}

void copySubstateFromGPU(Qureg qureg, long long int startInd, long long int numAmps) {
    API_TRY
    qv::numAmps(qureg, startInd, numAmps, __func__);
    qmx::downloadAmps(D(qureg), startInd, qureg.stateVec.real + startInd, qureg.stateVec.imag + startInd, numAmps);
    API_CATCH_VOID
}

static Complex ampAt(const Qureg &q, long long ind) {
    double re, im;
    qmx::downloadAmps(D(q), ind, &re, &im, 1);
    Complex c;
    c.real = re;
    c.imag = im;
    return c;
}

Complex getAmp(Qureg qureg, long long int index) {
    Complex z{0, 0};
    API_TRY
    qv::stateVecQureg(qureg, __func__);
    qv::ampIndex(qureg, index, __func__);
    return ampAt(qureg, index);
    API_CATCH(z)
}

qreal getRealAmp(Qureg qureg, long long int index) {
    API_TRY
    qv::stateVecQureg(qureg, __func__);
    qv::ampIndex(qureg, index, __func__);
    return ampAt(qureg, index).real;
    API_CATCH(0)
}

qreal getImagAmp(Qureg qureg, long long int index) {
    API_TRY
    qv::stateVecQureg(qureg, __func__);
    qv::ampIndex(qureg, index, __func__);
    return ampAt(qureg, index).imag;
    API_CATCH(0)
}

qreal getProbAmp(Qureg qureg, long long int index) {
    API_TRY
    qv::stateVecQureg(qureg, __func__);
    qv::ampIndex(qureg, index, __func__);
    Complex c = ampAt(qureg, index);
    return c.real * c.real + c.imag * c.imag;
    API_CATCH(0)
}

Complex getDensityAmp(Qureg qureg, long long int row, long long int col) {
    Complex z{0, 0};
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    qv::ampIndex(qureg, row, __func__);
    qv::ampIndex(qureg, col, __func__);
    return ampAt(qureg, row + col * (1LL << qureg.numQubitsRepresented));
    API_CATCH(z)
}

/* ---- unitaries ------------------------------------------------------------------------- */

void phaseShift(Qureg qureg, int targetQubit, qreal angle) {
    API_TRY
    qv::target(qureg, targetQubit, __func__);
    phaseAllOnes(qureg, 1ull << targetQubit, Cplx{std::cos(angle), std::sin(angle)});
    API_CATCH_VOID
}

void controlledPhaseShift(Qureg qureg, int idQubit1, int idQubit2, qreal angle) {
    API_TRY
    qv::controlTarget(qureg, idQubit1, idQubit2, __func__);
    phaseAllOnes(qureg, (1ull << idQubit1) | (1ull << idQubit2), Cplx{std::cos(angle), std::sin(angle)});
    API_CATCH_VOID
}

void multiControlledPhaseShift(Qureg qureg, int *controlQubits, int numControlQubits, qreal angle) {
    API_TRY
    qv::multiQubits(qureg, controlQubits, numControlQubits, __func__);
    phaseAllOnes(qureg, bitMask(controlQubits, numControlQubits), Cplx{std::cos(angle), std::sin(angle)});
    API_CATCH_VOID
}

void controlledPhaseFlip(Qureg qureg, int idQubit1, int idQubit2) {
    API_TRY
    qv::controlTarget(qureg, idQubit1, idQubit2, __func__);
    phaseAllOnes(qureg, (1ull << idQubit1) | (1ull << idQubit2), Cplx{-1, 0});
    API_CATCH_VOID
}

void multiControlledPhaseFlip(Qureg qureg, int *controlQubits, int numControlQubits) {
    API_TRY
    qv::multiQubits(qureg, controlQubits, numControlQubits, __func__);
    phaseAllOnes(qureg, bitMask(controlQubits, numControlQubits), Cplx{-1, 0});
    API_CATCH_VOID
}

void sGate(Qureg qureg, int targetQubit) {
    API_TRY
    qv::target(qureg, targetQubit, __func__);
    phaseAllOnes(qureg, 1ull << targetQubit, Cplx{0, 1});
    API_CATCH_VOID
}

void tGate(Qureg qureg, int targetQubit) {
    API_TRY
    qv::target(qureg, targetQubit, __func__);
    phaseAllOnes(qureg, 1ull << targetQubit, Cplx{1 / std::sqrt(2.0), 1 / std::sqrt(2.0)});
    API_CATCH_VOID
}

void compactUnitary(Qureg qureg, int targetQubit, Complex alpha, Complex beta) {
    API_TRY
    qv::target(qureg, targetQubit, __func__);
    qv::unitaryComplexPair(alpha, beta, __func__);
    Cplx u[4];
    compactToCplx(alpha, beta, u);
    gate2Conj(qureg, targetQubit, 0, 0, u);
    API_CATCH_VOID
}

void unitary(Qureg qureg, int targetQubit, ComplexMatrix2 u) {
    API_TRY
    qv::target(qureg, targetQubit, __func__);
    qv::oneQubitUnitary(u, __func__);
    Cplx m[4];
    m2ToCplx(u, m);
    gate2Conj(qureg, targetQubit, 0, 0, m);
    API_CATCH_VOID
}

static void rotateAxis(const Qureg &q, int target, uint64_t ctrl, double angle, Vector axis) {
    Complex a, b;
    rotationPair(angle, axis, &a, &b);
    Cplx u[4];
    compactToCplx(a, b, u);
    gate2Conj(q, target, ctrl, 0, u);
}

void rotateX(Qureg qureg, int rotQubit, qreal angle) {
    API_TRY
    qv::target(qureg, rotQubit, __func__);
    rotateAxis(qureg, rotQubit, 0, angle, Vector{1, 0, 0});
    API_CATCH_VOID
}

void rotateY(Qureg qureg, int rotQubit, qreal angle) {
    API_TRY
    qv::target(qureg, rotQubit, __func__);
    rotateAxis(qureg, rotQubit, 0, angle, Vector{0, 1, 0});
    API_CATCH_VOID
}

void rotateZ(Qureg qureg, int rotQubit, qreal angle) {
    API_TRY
    qv::target(qureg, rotQubit, __func__);
    rotateAxis(qureg, rotQubit, 0, angle, Vector{0, 0, 1});
    API_CATCH_VOID
}

void rotateAroundAxis(Qureg qureg, int rotQubit, qreal angle, Vector axis) {
    API_TRY
    qv::target(qureg, rotQubit, __func__);
    qv::vector(axis, __func__);
    rotateAxis(qureg, rotQubit, 0, angle, axis);
    API_CATCH_VOID
}

void controlledRotateX(Qureg qureg, int controlQubit, int targetQubit, qreal angle) {
    API_TRY
    qv::controlTarget(qureg, controlQubit, targetQubit, __func__);
    rotateAxis(qureg, targetQubit, 1ull << controlQubit, angle, Vector{1, 0, 0});
    API_CATCH_VOID
}

void controlledRotateY(Qureg qureg, int controlQubit, int targetQubit, qreal angle) {
    API_TRY
    qv::controlTarget(qureg, controlQubit, targetQubit, __func__);
    rotateAxis(qureg, targetQubit, 1ull << controlQubit, angle, Vector{0, 1, 0});
    API_CATCH_VOID
}

void controlledRotateZ(Qureg qureg, int controlQubit, int targetQubit, qreal angle) {
    API_TRY
    qv::controlTarget(qureg, controlQubit, targetQubit, __func__);
    rotateAxis(qureg, targetQubit, 1ull << controlQubit, angle, Vector{0, 0, 1});
    API_CATCH_VOID
}

void controlledRotateAroundAxis(Qureg qureg, int controlQubit, int targetQubit, qreal angle, Vector axis) {
    API_TRY
    qv::controlTarget(qureg, controlQubit, targetQubit, __func__);
    qv::vector(axis, __func__);
    rotateAxis(qureg, targetQubit, 1ull << controlQubit, angle, axis);
    API_CATCH_VOID
}

void controlledCompactUnitary(Qureg qureg, int controlQubit, int targetQubit, Complex alpha, Complex beta) {
    API_TRY
    qv::controlTarget(qureg, controlQubit, targetQubit, __func__);
    qv::unitaryComplexPair(alpha, beta, __func__);
    Cplx u[4];
    compactToCplx(alpha, beta, u);
    gate2Conj(qureg, targetQubit, 1ull << controlQubit, 0, u);
    API_CATCH_VOID
}

void controlledUnitary(Qureg qureg, int controlQubit, int targetQubit, ComplexMatrix2 u) {
    API_TRY
    qv::controlTarget(qureg, controlQubit, targetQubit, __func__);
    qv::oneQubitUnitary(u, __func__);
    Cplx m[4];
    m2ToCplx(u, m);
    gate2Conj(qureg, targetQubit, 1ull << controlQubit, 0, m);
    API_CATCH_VOID
}

void multiControlledUnitary(Qureg qureg, int *controlQubits, int numControlQubits, int targetQubit,
                            ComplexMatrix2 u) {
    API_TRY
    qv::multiControlsTarget(qureg, controlQubits, numControlQubits, targetQubit, __func__);
    qv::oneQubitUnitary(u, __func__);
    Cplx m[4];
    m2ToCplx(u, m);
    gate2Conj(qureg, targetQubit, bitMask(controlQubits, numControlQubits), 0, m);
    API_CATCH_VOID
}

void multiStateControlledUnitary(Qureg qureg, int *controlQubits, int *controlState, int numControlQubits,
                                 int targetQubit, ComplexMatrix2 u) {
    API_TRY
    qv::multiControlsTarget(qureg, controlQubits, numControlQubits, targetQubit, __func__);
    qv::oneQubitUnitary(u, __func__);
    qv::controlState(controlState, numControlQubits, __func__);
    uint64_t flip = 0;
    for (int i = 0; i < numControlQubits; ++i)
        if (controlState[i] == 0) flip |= 1ull << controlQubits[i];
    Cplx m[4];
    m2ToCplx(u, m);
    gate2Conj(qureg, targetQubit, bitMask(controlQubits, numControlQubits), flip, m);
    API_CATCH_VOID
}

void pauliX(Qureg qureg, int targetQubit) {
    API_TRY
    qv::target(qureg, targetQubit, __func__);
    if (qureg.isDensityMatrix) {
        const Cplx x[4] = {{0, 0}, {1, 0}, {1, 0}, {0, 0}};
        qmx::densUnitary(D(qureg), &targetQubit, 1, x, x);
    } else {
        qmx::applyNot(D(qureg), 0, 1ull << targetQubit);
    }
    API_CATCH_VOID
}

void pauliY(Qureg qureg, int targetQubit) {
    API_TRY
    qv::target(qureg, targetQubit, __func__);
    static const Cplx Y[4] = {{0, 0}, {0, -1}, {0, 1}, {0, 0}};
    gate2Conj(qureg, targetQubit, 0, 0, Y);
    API_CATCH_VOID
}

void pauliZ(Qureg qureg, int targetQubit) {
    API_TRY
    qv::target(qureg, targetQubit, __func__);
    phaseAllOnes(qureg, 1ull << targetQubit, Cplx{-1, 0});
    API_CATCH_VOID
}

void hadamard(Qureg qureg, int targetQubit) {
    API_TRY
    qv::target(qureg, targetQubit, __func__);
    const double r = 1 / std::sqrt(2.0);
    const Cplx H[4] = {{r, 0}, {r, 0}, {r, 0}, {-r, 0}};
    gate2(qureg, targetQubit, 0, 0, H, H);
    API_CATCH_VOID
}

void controlledNot(Qureg qureg, int controlQubit, int targetQubit) {
    API_TRY
    qv::controlTarget(qureg, controlQubit, targetQubit, __func__);
    if (qureg.isDensityMatrix) {
        const Cplx x[4] = {{0, 0}, {1, 0}, {1, 0}, {0, 0}};
        Cplx u4[16];
        controlled4(x, 1, u4);
        int t[2] = {targetQubit, controlQubit};
        qmx::densUnitary(D(qureg), t, 2, u4, u4);
    } else {
        qmx::applyNot(D(qureg), 1ull << controlQubit, 1ull << targetQubit);
    }
    API_CATCH_VOID
}

void multiControlledMultiQubitNot(Qureg qureg, int *ctrls, int numCtrls, int *targs, int numTargs) {
    API_TRY
    qv::multiControlsMultiTargets(qureg, ctrls, numCtrls, targs, numTargs, __func__);
    uint64_t c = bitMask(ctrls, numCtrls), t = bitMask(targs, numTargs);
    qmx::applyNot(D(qureg), c, t);
    if (qureg.isDensityMatrix) {
        int s = qureg.numQubitsRepresented;
        qmx::applyNot(D(qureg), c << s, t << s);
    }
    API_CATCH_VOID
}

void multiQubitNot(Qureg qureg, int *targs, int numTargs) {
    API_TRY
    qv::multiTargets(qureg, targs, numTargs, __func__);
    uint64_t t = bitMask(targs, numTargs);
    qmx::applyNot(D(qureg), 0, t);
    if (qureg.isDensityMatrix) qmx::applyNot(D(qureg), 0, t << qureg.numQubitsRepresented);
    API_CATCH_VOID
}

void controlledPauliY(Qureg qureg, int controlQubit, int targetQubit) {
    API_TRY
    qv::controlTarget(qureg, controlQubit, targetQubit, __func__);
    static const Cplx Y[4] = {{0, 0}, {0, -1}, {0, 1}, {0, 0}};
    gate2Conj(qureg, targetQubit, 1ull << controlQubit, 0, Y);
    API_CATCH_VOID
}

void swapGate(Qureg qureg, int qb1, int qb2) {
    API_TRY
    qv::uniqueTargets(qureg, qb1, qb2, __func__);
    if (qureg.isDensityMatrix) {
        Cplx sw[16] = {};
        sw[0] = sw[15] = Cplx{1, 0};
        sw[1 * 4 + 2] = sw[2 * 4 + 1] = Cplx{1, 0};
        int t[2] = {qb1, qb2};
        qmx::densUnitary(D(qureg), t, 2, sw, sw);
    } else {
        qmx::applySwap(D(qureg), qb1, qb2);
    }
    API_CATCH_VOID
}

void sqrtSwapGate(Qureg qureg, int qb1, int qb2) {
    API_TRY
    qv::uniqueTargets(qureg, qb1, qb2, __func__);
    qv::matrixFitsInNode(qureg, 2, __func__);
    // QuEST_common.c:384-395
    std::vector<Cplx> u(16, Cplx{0, 0});
    u[0] = {1, 0};
    u[15] = {1, 0};
    u[1 * 4 + 1] = {.5, .5};
    u[1 * 4 + 2] = {.5, -.5};
    u[2 * 4 + 1] = {.5, -.5};
    u[2 * 4 + 2] = {.5, .5};
    std::vector<Cplx> uc(u);
    for (auto &z : uc) z.im = -z.im;
    int t[2] = {qb1, qb2};
    matrixN(qureg, 0, t, 2, u, uc);
    API_CATCH_VOID
}

void multiRotateZ(Qureg qureg, int *qubits, int numQubits, qreal angle) {
    API_TRY
    qv::multiTargets(qureg, qubits, numQubits, __func__);
    uint64_t m = bitMask(qubits, numQubits);
    qmx::applyMultiRotateZ(D(qureg), 0, m, angle);
    if (qureg.isDensityMatrix) qmx::applyMultiRotateZ(D(qureg), 0, m << qureg.numQubitsRepresented, -angle);
    API_CATCH_VOID
}

void multiControlledMultiRotateZ(Qureg qureg, int *controlQubits, int numControls, int *targetQubits,
                                 int numTargets, qreal angle) {
    API_TRY
    qv::multiControlsMultiTargets(qureg, controlQubits, numControls, targetQubits, numTargets, __func__);
    uint64_t c = bitMask(controlQubits, numControls), t = bitMask(targetQubits, numTargets);
    qmx::applyMultiRotateZ(D(qureg), c, t, angle);
    if (qureg.isDensityMatrix) {
        int s = qureg.numQubitsRepresented;
        qmx::applyMultiRotateZ(D(qureg), c << s, t << s, -angle);
    }
    API_CATCH_VOID
}

void multiRotatePauli(Qureg qureg, int *targetQubits, enum pauliOpType *targetPaulis, int numTargets, qreal angle) {
    API_TRY
    qv::multiTargets(qureg, targetQubits, numTargets, __func__);
    qv::pauliCodes(targetPaulis, numTargets, __func__);
    pauliRot(qureg, 0, targetQubits, targetPaulis, numTargets, angle);
    API_CATCH_VOID
}

void multiControlledMultiRotatePauli(Qureg qureg, int *controlQubits, int numControls, int *targetQubits,
                                     enum pauliOpType *targetPaulis, int numTargets, qreal angle) {
    API_TRY
    qv::multiControlsMultiTargets(qureg, controlQubits, numControls, targetQubits, numTargets, __func__);
    qv::pauliCodes(targetPaulis, numTargets, __func__);
    pauliRot(qureg, bitMask(controlQubits, numControls), targetQubits, targetPaulis, numTargets, angle);
    API_CATCH_VOID
}

void twoQubitUnitary(Qureg qureg, int targetQubit1, int targetQubit2, ComplexMatrix4 u) {
    API_TRY
    int t[2] = {targetQubit1, targetQubit2};
    qv::multiTargets(qureg, t, 2, __func__);
    qv::twoQubitUnitary(qureg, u, __func__);
    matrixN(qureg, 0, t, 2, matToCplx<4>(u, false), matToCplx<4>(u, true));
    API_CATCH_VOID
}

void controlledTwoQubitUnitary(Qureg qureg, int controlQubit, int targetQubit1, int targetQubit2, ComplexMatrix4 u) {
    API_TRY
    int c[1] = {controlQubit};
    int t[2] = {targetQubit1, targetQubit2};
    qv::multiControlsMultiTargets(qureg, c, 1, t, 2, __func__);
    qv::twoQubitUnitary(qureg, u, __func__);
    matrixN(qureg, 1ull << controlQubit, t, 2, matToCplx<4>(u, false), matToCplx<4>(u, true));
    API_CATCH_VOID
}

void multiControlledTwoQubitUnitary(Qureg qureg, int *controlQubits, int numControlQubits, int targetQubit1,
                                    int targetQubit2, ComplexMatrix4 u) {
    API_TRY
    int t[2] = {targetQubit1, targetQubit2};
    qv::multiControlsMultiTargets(qureg, controlQubits, numControlQubits, t, 2, __func__);
    qv::twoQubitUnitary(qureg, u, __func__);
    matrixN(qureg, bitMask(controlQubits, numControlQubits), t, 2, matToCplx<4>(u, false), matToCplx<4>(u, true));
    API_CATCH_VOID
}

void multiQubitUnitary(Qureg qureg, int *targs, int numTargs, ComplexMatrixN u) {
    API_TRY
    qv::multiTargets(qureg, targs, numTargs, __func__);
    qv::multiQubitUnitary(qureg, u, numTargs, __func__);
    matrixN(qureg, 0, targs, numTargs, matNToCplx(u, false), matNToCplx(u, true));
    API_CATCH_VOID
}

void controlledMultiQubitUnitary(Qureg qureg, int ctrl, int *targs, int numTargs, ComplexMatrixN u) {
    API_TRY
    int c[1] = {ctrl};
    qv::multiControlsMultiTargets(qureg, c, 1, targs, numTargs, __func__);
    qv::multiQubitUnitary(qureg, u, numTargs, __func__);
    matrixN(qureg, 1ull << ctrl, targs, numTargs, matNToCplx(u, false), matNToCplx(u, true));
    API_CATCH_VOID
}

void multiControlledMultiQubitUnitary(Qureg qureg, int *ctrls, int numCtrls, int *targs, int numTargs,
                                      ComplexMatrixN u) {
    API_TRY
    qv::multiControlsMultiTargets(qureg, ctrls, numCtrls, targs, numTargs, __func__);
    qv::multiQubitUnitary(qureg, u, numTargs, __func__);
    matrixN(qureg, bitMask(ctrls, numCtrls), targs, numTargs, matNToCplx(u, false), matNToCplx(u, true));
    API_CATCH_VOID
}

/* ---- non-unitary operators --------------------------------------------------------------- */

void setWeightedQureg(Complex fac1, Qureg qureg1, Complex fac2, Qureg qureg2, Complex facOut, Qureg out) {
    API_TRY
    qv::matchingQuregTypes(qureg1, qureg2, __func__);
    qv::matchingQuregTypes(qureg1, out, __func__);
    qv::matchingQuregDims(qureg1, qureg2, __func__);
    qv::matchingQuregDims(qureg1, out, __func__);
    qv::matchingPrecision(qureg1, qureg2, __func__);
    qv::matchingPrecision(qureg1, out, __func__);
    qmx::weightedSum(Cplx{fac1.real, fac1.imag}, D(qureg1), Cplx{fac2.real, fac2.imag}, D(qureg2),
                     Cplx{facOut.real, facOut.imag}, D(out));
    API_CATCH_VOID
}

static void pauliSum(const Qureg &in, const enum pauliOpType *codes, const qreal *coeffs, int numTerms,
                     const Qureg &out) {
    // QuEST_common.c:535-557
    int n = in.numQubitsRepresented;
    std::vector<int> targs(n);
    for (int q = 0; q < n; ++q) targs[q] = q;
    qmx::initBlank(D(out));
    for (int t = 0; t < numTerms; ++t) {
        pauliProd(in, targs.data(), &codes[(size_t)t * n], n);
        qmx::weightedSum(Cplx{coeffs[t], 0}, D(in), Cplx{1, 0}, D(out), Cplx{0, 0}, D(out));
        pauliProd(in, targs.data(), &codes[(size_t)t * n], n);
    }
}

void applyPauliSum(Qureg inQureg, enum pauliOpType *allPauliCodes, qreal *termCoeffs, int numSumTerms,
                   Qureg outQureg) {
    API_TRY
    qv::matchingQuregTypes(inQureg, outQureg, __func__);
    qv::matchingQuregDims(inQureg, outQureg, __func__);
    qv::numPauliSumTerms(numSumTerms, __func__);
    qv::pauliCodes(allPauliCodes, numSumTerms * inQureg.numQubitsRepresented, __func__);
    qv::matchingPrecision(inQureg, outQureg, __func__);
    pauliSum(inQureg, allPauliCodes, termCoeffs, numSumTerms, outQureg);
    API_CATCH_VOID
}

void applyPauliHamil(Qureg inQureg, PauliHamil hamil, Qureg outQureg) {
    API_TRY
    qv::matchingQuregTypes(inQureg, outQureg, __func__);
    qv::matchingQuregDims(inQureg, outQureg, __func__);
    qv::pauliHamil(hamil, __func__);
    qv::matchingQuregPauliHamilDims(inQureg, hamil, __func__);
    qv::matchingPrecision(inQureg, outQureg, __func__);
    pauliSum(inQureg, hamil.pauliCodes, hamil.termCoeffs, hamil.numSumTerms, outQureg);
    API_CATCH_VOID
}

int qmxTrotterSweepsPerRep(PauliHamil hamil, int order, int isDensityMatrix, int precision) {
    if (hamil.numQubits < 1 || hamil.numSumTerms < 1 || (order != 1 && (order < 2 || order % 2))) return 0;
    Qureg shape;
    std::memset(&shape, 0, sizeof shape);
    shape.isDensityMatrix = isDensityMatrix != 0;
    shape.numQubitsRepresented = hamil.numQubits;
    std::vector<qmx::PauliExp> seq;
    symmetrizedTrotter(seq, shape, hamil, 1.0, order);
    const int nq = isDensityMatrix ? 2 * hamil.numQubits : hamil.numQubits;
    return qmx::pauliSequenceBlocks(nq, precision == QMX_PREC_FP32 ? qmx::FP32 : qmx::FP64, seq.data(),
                                    (int)seq.size());
}

void applyTrotterCircuit(Qureg qureg, PauliHamil hamil, qreal time, int order, int reps) {
    API_TRY
    qv::trotterParams(order, reps, __func__);
    qv::pauliHamil(hamil, __func__);
    qv::matchingQuregPauliHamilDims(qureg, hamil, __func__);
    if (time == 0) return;
    std::vector<qmx::PauliExp> seq;
    symmetrizedTrotter(seq, qureg, hamil, time / reps, order);
    qmx::applyPauliSequence(D(qureg), seq.data(), (int)seq.size(), reps);
    API_CATCH_VOID
}

void applyMatrix2(Qureg qureg, int targetQubit, ComplexMatrix2 u) {
    API_TRY
    qv::target(qureg, targetQubit, __func__);
    Cplx m[4];
    m2ToCplx(u, m);
    qmx::applyMatrix2(D(qureg), targetQubit, 0, 0, m);
    API_CATCH_VOID
}

void applyMatrix4(Qureg qureg, int targetQubit1, int targetQubit2, ComplexMatrix4 u) {
    API_TRY
    int t[2] = {targetQubit1, targetQubit2};
    qv::multiTargets(qureg, t, 2, __func__);
    qv::matrixFitsInNode(qureg, 2, __func__);
    std::vector<Cplx> m = matToCplx<4>(u, false);
    qmx::applyMatrixN(D(qureg), 0, t, 2, m.data());
    API_CATCH_VOID
}

void applyMatrixN(Qureg qureg, int *targs, int numTargs, ComplexMatrixN u) {
    API_TRY
    qv::multiTargets(qureg, targs, numTargs, __func__);
    qv::multiQubitMatrix(qureg, u, numTargs, __func__);
    std::vector<Cplx> m = matNToCplx(u, false);
    qmx::applyMatrixN(D(qureg), 0, targs, numTargs, m.data());
    API_CATCH_VOID
}

void applyMultiControlledMatrixN(Qureg qureg, int *ctrls, int numCtrls, int *targs, int numTargs,
                                 ComplexMatrixN u) {
    API_TRY
    qv::multiControlsMultiTargets(qureg, ctrls, numCtrls, targs, numTargs, __func__);
    qv::multiQubitMatrix(qureg, u, numTargs, __func__);
    std::vector<Cplx> m = matNToCplx(u, false);
    qmx::applyMatrixN(D(qureg), bitMask(ctrls, numCtrls), targs, numTargs, m.data());
    API_CATCH_VOID
}

void applyProjector(Qureg qureg, int qubit, int outcome) {
    API_TRY
    qv::target(qureg, qubit, __func__);
    qv::outcome(outcome, __func__);
    collapseKnown(qureg, qubit, outcome, 1.0);
    API_CATCH_VOID
}

/* ---- calculations ---------------------------------------------------------------------------- */

qreal calcTotalProb(Qureg qureg) {
    return qureg.isDensityMatrix ? qmx::densTotalProb(D(qureg)) : qmx::totalProb(D(qureg));
}

qreal calcProbOfOutcome(Qureg qureg, int measureQubit, int outcome) {
    API_TRY
    qv::target(qureg, measureQubit, __func__);
    qv::outcome(outcome, __func__);
    return probOfOutcome(qureg, measureQubit, outcome);
    API_CATCH(0)
}

void calcProbOfAllOutcomes(qreal *outcomeProbs, Qureg qureg, int *qubits, int numQubits) {
    API_TRY
    qv::multiTargets(qureg, qubits, numQubits, __func__);
    std::vector<double> p((size_t)1 << numQubits);
    qmx::probOfAllOutcomes(D(qureg), qubits, numQubits, p.data());
    for (size_t o = 0; o < p.size(); ++o) outcomeProbs[o] = (qreal)p[o];
    API_CATCH_VOID
}

Complex calcInnerProduct(Qureg bra, Qureg ket) {
    Complex z{0, 0};
    API_TRY
    qv::stateVecQureg(bra, __func__);
    qv::stateVecQureg(ket, __func__);
    qv::matchingQuregDims(bra, ket, __func__);
    Cplx r = qmx::innerProduct(D(bra), D(ket));
    z.real = r.re;
    z.imag = r.im;
    return z;
    API_CATCH(z)
}

qreal calcDensityInnerProduct(Qureg rho1, Qureg rho2) {
    API_TRY
    qv::densityMatrQureg(rho1, __func__);
    qv::densityMatrQureg(rho2, __func__);
    qv::matchingQuregDims(rho1, rho2, __func__);
    qv::matchingPrecision(rho1, rho2, __func__);
    return qmx::densInnerProduct(D(rho1), D(rho2));
    API_CATCH(0)
}

qreal calcPurity(Qureg qureg) {
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    return qmx::purity(D(qureg));
    API_CATCH(0)
}

qreal calcFidelity(Qureg qureg, Qureg pureState) {
    API_TRY
    qv::secondQuregStateVec(pureState, __func__);
    qv::matchingQuregDims(qureg, pureState, __func__);
    if (qureg.isDensityMatrix) return qmx::densFidelity(D(qureg), D(pureState));
    Cplx ip = qmx::innerProduct(D(qureg), D(pureState));
    return ip.re * ip.re + ip.im * ip.im;
    API_CATCH(0)
}

qreal calcExpecPauliProd(Qureg qureg, int *targetQubits, enum pauliOpType *pauliCodes, int numTargets,
                         Qureg workspace) {
    API_TRY
    qv::multiTargets(qureg, targetQubits, numTargets, __func__);
    qv::pauliCodes(pauliCodes, numTargets, __func__);
    qv::matchingQuregTypes(qureg, workspace, __func__);
    qv::matchingQuregDims(qureg, workspace, __func__);
    qv::matchingPrecision(qureg, workspace, __func__);
    return expecPauliProd(qureg, targetQubits, pauliCodes, numTargets, workspace);
    API_CATCH(0)
}

static double expecPauliSum(const Qureg &q, const enum pauliOpType *codes, const qreal *coeffs, int numTerms,
                            const Qureg &ws) {
    int n = q.numQubitsRepresented;
    std::vector<int> targs(n);
    for (int i = 0; i < n; ++i) targs[i] = i;
    double value = 0;
    for (int t = 0; t < numTerms; ++t) value += coeffs[t] * expecPauliProd(q, targs.data(), &codes[(size_t)t * n], n, ws);
    return value;
}

qreal calcExpecPauliSum(Qureg qureg, enum pauliOpType *allPauliCodes, qreal *termCoeffs, int numSumTerms,
                        Qureg workspace) {
    API_TRY
    qv::numPauliSumTerms(numSumTerms, __func__);
    qv::pauliCodes(allPauliCodes, numSumTerms * qureg.numQubitsRepresented, __func__);
    qv::matchingQuregTypes(qureg, workspace, __func__);
    qv::matchingQuregDims(qureg, workspace, __func__);
    qv::matchingPrecision(qureg, workspace, __func__);
    return expecPauliSum(qureg, allPauliCodes, termCoeffs, numSumTerms, workspace);
    API_CATCH(0)
}

qreal calcExpecPauliHamil(Qureg qureg, PauliHamil hamil, Qureg workspace) {
    API_TRY
    qv::matchingQuregTypes(qureg, workspace, __func__);
    qv::matchingQuregDims(qureg, workspace, __func__);
    qv::pauliHamil(hamil, __func__);
    qv::matchingQuregPauliHamilDims(qureg, hamil, __func__);
    qv::matchingPrecision(qureg, workspace, __func__);
    return expecPauliSum(qureg, hamil.pauliCodes, hamil.termCoeffs, hamil.numSumTerms, workspace);
    API_CATCH(0)
}

qreal calcHilbertSchmidtDistance(Qureg a, Qureg b) {
    API_TRY
    qv::densityMatrQureg(a, __func__);
    qv::densityMatrQureg(b, __func__);
    qv::matchingQuregDims(a, b, __func__);
    qv::matchingPrecision(a, b, __func__);
    return std::sqrt(qmx::hsDistanceSquared(D(a), D(b)));
    API_CATCH(0)
}

/* ---- measurement ------------------------------------------------------------------------------ */

qreal collapseToOutcome(Qureg qureg, int measureQubit, int outcome) {
    API_TRY
    qv::target(qureg, measureQubit, __func__);
    qv::outcome(outcome, __func__);
    double p = probOfOutcome(qureg, measureQubit, outcome);
    qv::measurementProb(p, __func__);
    collapseKnown(qureg, measureQubit, outcome, p);
    return p;
    API_CATCH(0)
}

int measureWithStats(Qureg qureg, int measureQubit, qreal *outcomeProb) {
    API_TRY
    qv::target(qureg, measureQubit, __func__);
    double p0 = probOfOutcome(qureg, measureQubit, 0);
    double prob = 0;
    int outcome = generateMeasurementOutcome(p0, &prob);
    collapseKnown(qureg, measureQubit, outcome, prob);
    *outcomeProb = (qreal)prob;
    return outcome;
    API_CATCH(0)
}

int measure(Qureg qureg, int measureQubit) {
    API_TRY
    qv::target(qureg, measureQubit, __func__);
    double discard;
    double p0 = probOfOutcome(qureg, measureQubit, 0);
    int outcome = generateMeasurementOutcome(p0, &discard);
    collapseKnown(qureg, measureQubit, outcome, discard);
    return outcome;
    API_CATCH(0)
}

/* ---- decoherence ---------------------------------------------------------------------------------- */

void mixDephasing(Qureg qureg, int targetQubit, qreal prob) {
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    qv::target(qureg, targetQubit, __func__);
    qv::oneQubitDephaseProb(prob, __func__);
    // QuEST.c:1259-1267 -> retain = 1 - 2p
    qmx::densDephase(D(qureg), targetQubit, 1 - 2 * prob);
    API_CATCH_VOID
}

void mixTwoQubitDephasing(Qureg qureg, int qubit1, int qubit2, qreal prob) {
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    qv::uniqueTargets(qureg, qubit1, qubit2, __func__);
    qv::twoQubitDephaseProb(prob, __func__);
    // QuEST.c:1269-1280 -> retain = 1 - 4p/3
    qmx::densTwoQubitDephase(D(qureg), qubit1, qubit2, 1 - (4 * prob) / 3.0);
    API_CATCH_VOID
}

void mixDepolarising(Qureg qureg, int targetQubit, qreal prob) {
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    qv::target(qureg, targetQubit, __func__);
    qv::oneQubitDepolProb(prob, __func__);
    double level = (4 * prob) / 3.0;
    if (level != 0) qmx::densDepolarise(D(qureg), targetQubit, level);
    API_CATCH_VOID
}

void mixDamping(Qureg qureg, int targetQubit, qreal prob) {
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    qv::target(qureg, targetQubit, __func__);
    qv::oneQubitDampingProb(prob, __func__);
    if (prob != 0) qmx::densDamp(D(qureg), targetQubit, prob);
    API_CATCH_VOID
}

void mixTwoQubitDepolarising(Qureg qureg, int qubit1, int qubit2, qreal prob) {
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    qv::uniqueTargets(qureg, qubit1, qubit2, __func__);
    qv::twoQubitDepolProb(prob, __func__);
    if (prob == 0) return;
    // Kraus form: sqrt(1-p) I, sqrt(p/15) P for the 15 non-identity Pauli pairs
    static const Cplx P[4][4] = {{{1, 0}, {0, 0}, {0, 0}, {1, 0}},
                                 {{0, 0}, {1, 0}, {1, 0}, {0, 0}},
                                 {{0, 0}, {0, -1}, {0, 1}, {0, 0}},
                                 {{1, 0}, {0, 0}, {0, 0}, {-1, 0}}};
    auto K = [&](int n, int r, int c) {
        int a = n / 4, b = n % 4;  // operator P_a on qubit1 (low bit), P_b on qubit2
        double s = (n == 0) ? std::sqrt(1 - prob) : std::sqrt(prob / 15.0);
        Cplx x = P[a][(r & 1) * 2 + (c & 1)], y = P[b][(r >> 1) * 2 + (c >> 1)];
        return Cplx{s * (x.re * y.re - x.im * y.im), s * (x.re * y.im + x.im * y.re)};
    };
    std::vector<Cplx> sup = krausSuperOp(16, 4, K);
    int t[2] = {qubit1, qubit2};
    applySuperOp(qureg, t, 2, sup);
    API_CATCH_VOID
}

void mixPauli(Qureg qureg, int qubit, qreal probX, qreal probY, qreal probZ) {
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    qv::target(qureg, qubit, __func__);
    qv::oneQubitPauliProbs(probX, probY, probZ, __func__);
    // QuEST_common.c:717-740
    double f[4] = {std::sqrt(1 - (probX + probY + probZ)), std::sqrt(probX), std::sqrt(probY), std::sqrt(probZ)};
    Cplx ops[4][4] = {};
    ops[0][0] = {f[0], 0};
    ops[0][3] = {f[0], 0};
    ops[1][1] = {f[1], 0};
    ops[1][2] = {f[1], 0};
    ops[2][1] = {0, -f[2]};
    ops[2][2] = {0, f[2]};
    ops[3][0] = {f[3], 0};
    ops[3][3] = {-f[3], 0};
    std::vector<Cplx> sup = krausSuperOp(4, 2, [&](int n, int r, int c) { return ops[n][r * 2 + c]; });
    applySuperOp(qureg, &qubit, 1, sup);
    API_CATCH_VOID
}

void mixDensityMatrix(Qureg combineQureg, qreal prob, Qureg otherQureg) {
    API_TRY
    qv::densityMatrQureg(combineQureg, __func__);
    qv::densityMatrQureg(otherQureg, __func__);
    qv::matchingQuregDims(combineQureg, otherQureg, __func__);
    qv::prob(prob, __func__);
    qv::matchingPrecision(combineQureg, otherQureg, __func__);
    qmx::mixDensityMatrix(D(combineQureg), prob, D(otherQureg));
    API_CATCH_VOID
}

static void kraus1(const Qureg &q, int target, const ComplexMatrix2 *ops, int numOps) {
    std::vector<Cplx> sup = krausSuperOp(numOps, 2, [&](int n, int r, int c) {
        return Cplx{ops[n].real[r][c], ops[n].imag[r][c]};
    });
    applySuperOp(q, &target, 1, sup);
}

static void kraus2(const Qureg &q, int t1, int t2, const ComplexMatrix4 *ops, int numOps) {
    std::vector<Cplx> sup = krausSuperOp(numOps, 4, [&](int n, int r, int c) {
        return Cplx{ops[n].real[r][c], ops[n].imag[r][c]};
    });
    int t[2] = {t1, t2};
    applySuperOp(q, t, 2, sup);
}

static void krausN(const Qureg &q, const int *targets, int numTargets, const ComplexMatrixN *ops, int numOps) {
    int d = 1 << ops[0].numQubits;
    std::vector<Cplx> sup = krausSuperOp(numOps, d, [&](int n, int r, int c) {
        return Cplx{ops[n].real[r][c], ops[n].imag[r][c]};
    });
    applySuperOp(q, targets, numTargets, sup);
}

void mixKrausMap(Qureg qureg, int target, ComplexMatrix2 *ops, int numOps) {
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    qv::target(qureg, target, __func__);
    qv::oneQubitKrausMap(qureg, ops, numOps, __func__);
    kraus1(qureg, target, ops, numOps);
    API_CATCH_VOID
}

void mixTwoQubitKrausMap(Qureg qureg, int target1, int target2, ComplexMatrix4 *ops, int numOps) {
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    int t[2] = {target1, target2};
    qv::multiTargets(qureg, t, 2, __func__);
    qv::twoQubitKrausMap(qureg, ops, numOps, __func__);
    kraus2(qureg, target1, target2, ops, numOps);
    API_CATCH_VOID
}

void mixMultiQubitKrausMap(Qureg qureg, int *targets, int numTargets, ComplexMatrixN *ops, int numOps) {
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    qv::multiTargets(qureg, targets, numTargets, __func__);
    qv::multiQubitKrausMap(qureg, numTargets, ops, numOps, __func__);
    krausN(qureg, targets, numTargets, ops, numOps);
    API_CATCH_VOID
}

void mixNonTPKrausMap(Qureg qureg, int target, ComplexMatrix2 *ops, int numOps) {
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    qv::target(qureg, target, __func__);
    qv::oneQubitKrausMapDims(qureg, numOps, __func__);
    kraus1(qureg, target, ops, numOps);
    API_CATCH_VOID
}

void mixNonTPTwoQubitKrausMap(Qureg qureg, int target1, int target2, ComplexMatrix4 *ops, int numOps) {
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    int t[2] = {target1, target2};
    qv::multiTargets(qureg, t, 2, __func__);
    qv::twoQubitKrausMapDims(qureg, numOps, __func__);
    kraus2(qureg, target1, target2, ops, numOps);
    API_CATCH_VOID
}

void mixNonTPMultiQubitKrausMap(Qureg qureg, int *targets, int numTargets, ComplexMatrixN *ops, int numOps) {
    API_TRY
    qv::densityMatrQureg(qureg, __func__);
    qv::multiTargets(qureg, targets, numTargets, __func__);
    qv::multiQubitKrausMapDims(qureg, numTargets, ops, numOps, __func__);
    krausN(qureg, targets, numTargets, ops, numOps);
    API_CATCH_VOID
}

/* ---- QASM: accepted, nothing recorded ------------------------------------------------------------- */

void startRecordingQASM(Qureg qureg) {
    if (qureg.qasmLog) qureg.qasmLog->isLogging = 1;
}
void stopRecordingQASM(Qureg qureg) {
    if (qureg.qasmLog) qureg.qasmLog->isLogging = 0;
}
void clearRecordedQASM(Qureg) {}
void printRecordedQASM(Qureg) {}
void writeRecordedQASMToFile(Qureg qureg, char *filename) {
    API_TRY
    (void)qureg;
    FILE *f = std::fopen(filename, "w");
    qv::fileOpened(f != nullptr, filename, __func__);
    std::fprintf(f, "// QASM recording is not provided by libquest_mi355x\n");
    std::fclose(f);
    API_CATCH_VOID
}

/* ---- extensions ---------------------------------------------------------------------------------- */

void qmxSetErrorMode(int mode) { qv::setMode(mode); }
int qmxLastError(char *buf, int len) { return qv::lastError(buf, len); }
void qmxClearError(void) { qv::clear(); }
int qmxDeviceInfo(char *buf, int len) { return qmx::deviceDescription(buf, len); }
/* ---- diagonal operators (QuEST.c:1136-1146, 1237-1244, 1557-1606) ----------------------------- */

DiagonalOp createDiagonalOp(int numQubits, QuESTEnv env) {
    (void)env;
    DiagonalOp op = emptyDiagonalOp();
    API_TRY
    qv::numQubitsInDiagOp(numQubits, __func__, env.numRanks);
    qmx::runtimeInit();
    const long long n = numQubits < 62 ? 1LL << numQubits : 0;
    op.numQubits = numQubits;
    op.numElemsPerChunk = n;
    op.numChunks = 1;
    op.chunkId = 0;
    op.real = n ? static_cast<qreal *>(std::calloc((size_t)n, sizeof(qreal))) : nullptr;
    op.imag = n ? static_cast<qreal *>(std::calloc((size_t)n, sizeof(qreal))) : nullptr;
    if (!op.real || !op.imag) {
        std::free(op.real);
        std::free(op.imag);
        op = emptyDiagonalOp();
        qv::raise(qv::message(qv::E_DIAGONAL_OP_NOT_ALLOCATED), __func__);
    }
    auto *dev = static_cast<double *>(qmxTryDevAlloc((size_t)n * 2 * sizeof(double)));
    if (!dev) {
        std::free(op.real);
        std::free(op.imag);
        op = emptyDiagonalOp();
        qv::raise(qv::message(qv::E_DIAGONAL_OP_NOT_ALLOCATED_ON_GPU), __func__);
    }
    op.deviceOperator.real = reinterpret_cast<qreal *>(dev);
    op.deviceOperator.imag = reinterpret_cast<qreal *>(dev + n);
    diagToDevice(op.deviceOperator.real, op.real, (size_t)n);
    diagToDevice(op.deviceOperator.imag, op.imag, (size_t)n);
    return op;
    API_CATCH(op)
}

void destroyDiagonalOp(DiagonalOp op, QuESTEnv env) {
    (void)env;
    API_TRY
    qv::diagOpInit(op, __func__);
    std::free(op.real);
    std::free(op.imag);
    qmx::devFree(op.deviceOperator.real);
    API_CATCH_VOID
}

void syncDiagonalOp(DiagonalOp op) {
    API_TRY
    qv::diagOpInit(op, __func__);
    diagToDevice(op.deviceOperator.real, op.real, (size_t)op.numElemsPerChunk);
    diagToDevice(op.deviceOperator.imag, op.imag, (size_t)op.numElemsPerChunk);
    API_CATCH_VOID
}

// host and device copies both take the elements (QuEST_cpu.c:4211-4249 for the host side)
static void setDiagElems(DiagonalOp op, long long start, const qreal *re, const qreal *im, long long n) {
    std::memcpy(op.real + start, re, (size_t)n * sizeof(qreal));
    std::memcpy(op.imag + start, im, (size_t)n * sizeof(qreal));
    diagToDevice(reinterpret_cast<qreal *>(devDiag(op.deviceOperator.real) + start), re, (size_t)n);
    diagToDevice(reinterpret_cast<qreal *>(devDiag(op.deviceOperator.imag) + start), im, (size_t)n);
}

void initDiagonalOp(DiagonalOp op, qreal *real, qreal *imag) {
    API_TRY
    qv::diagOpInit(op, __func__);
    setDiagElems(op, 0, real, imag, 1LL << op.numQubits);
    API_CATCH_VOID
}

void setDiagonalOpElems(DiagonalOp op, long long int startInd, qreal *real, qreal *imag, long long int numElems) {
    API_TRY
    qv::diagOpInit(op, __func__);
    qv::numElems(op, startInd, numElems, __func__);
    setDiagElems(op, startInd, real, imag, numElems);
    API_CATCH_VOID
}

static void diagFromHamil(DiagonalOp op, const PauliHamil &h) {
    std::vector<uint64_t> z(h.numSumTerms, 0);
    for (int t = 0; t < h.numSumTerms; ++t)
        for (int q = 0; q < h.numQubits; ++q)
            if (h.pauliCodes[(size_t)t * h.numQubits + q] == PAULI_Z) z[t] |= 1ull << q;
    std::vector<double> c(h.termCoeffs, h.termCoeffs + h.numSumTerms);
    qmx::diagonalFromZHamil(devDiag(op.deviceOperator.real), devDiag(op.deviceOperator.imag), op.numQubits, z.data(), c.data(),
                            h.numSumTerms);
    diagToHost(op.real, op.deviceOperator.real, (size_t)op.numElemsPerChunk);
    diagToHost(op.imag, op.deviceOperator.imag, (size_t)op.numElemsPerChunk);
}

void initDiagonalOpFromPauliHamil(DiagonalOp op, PauliHamil hamil) {
    API_TRY
    qv::hamilParams(hamil.numQubits, hamil.numSumTerms, __func__);
    qv::diagOpInit(op, __func__);
    qv::diagPauliHamil(op, hamil, __func__);
    diagFromHamil(op, hamil);
    API_CATCH_VOID
}

DiagonalOp createDiagonalOpFromPauliHamilFile(char *fn, QuESTEnv env) {
    DiagonalOp op = emptyDiagonalOp();
    PauliHamil h = createPauliHamilFromFile(fn);  // validates the file; a recorded error leaves h empty
    if (!h.pauliCodes) return op;
    API_TRY
    if (!qv::diagonalHamil(h)) {
        destroyPauliHamil(h);
        qv::raise(qv::message(qv::E_PAULI_HAMIL_NOT_DIAGONAL), __func__);
    }
    op = createDiagonalOp(h.numQubits, env);
    if (op.real) diagFromHamil(op, h);
    destroyPauliHamil(h);
    return op;
    API_CATCH(op)
}

void applyDiagonalOp(Qureg qureg, DiagonalOp op) {
    API_TRY
    qv::diagonalOp(qureg, op, __func__);
    qmx::applyDiagonal(D(qureg), devDiag(op.deviceOperator.real), devDiag(op.deviceOperator.imag));
    API_CATCH_VOID
}

Complex calcExpecDiagonalOp(Qureg qureg, DiagonalOp op) {
    Complex z{0, 0};
    API_TRY
    qv::diagonalOp(qureg, op, __func__);
    const Cplx r = qmx::expecDiagonal(D(qureg), devDiag(op.deviceOperator.real), devDiag(op.deviceOperator.imag));
    z.real = r.re;
    z.imag = r.im;
    return z;
    API_CATCH(z)
}

/* ---- phase functions and QFT (QuEST.c:732-891) -------------------------------------------------- */

// each public entry validates under its own name (the reference's __func__)
// withOverrides: called from an *Overrides entry point, whose override count
// is validated even when the index array is NULL (QuEST.c:732-891)
static void phaseFuncImpl(const Qureg &qureg, int *qubits, int numQubits, int encoding, qreal *coeffs,
                          qreal *exponents, int numTerms, long long *overrideInds, qreal *overridePhases,
                          int numOverrides, bool withOverrides, const char *caller) {
    const Doubles c(coeffs, (size_t)std::max(0, numTerms)), e(exponents, (size_t)std::max(0, numTerms));
    const Doubles ph(overridePhases, (size_t)std::max(0, numOverrides));
    qv::phaseMultiQubits(qureg, qubits, numQubits, caller);
    qv::bitEncoding(numQubits, encoding, caller);
    if (withOverrides) qv::phaseFuncOverrides(numQubits, encoding, overrideInds, numOverrides, caller);
    qv::phaseFuncTerms(numQubits, encoding, e.get(), numTerms, overrideInds, numOverrides, caller);
    qmx::PhaseFuncArgs a{};
    a.numRegs = 1;
    a.encoding = encoding;
    a.mode = 0;
    a.qubits = qubits;
    a.numQubitsPerReg = &numQubits;
    a.coeffs = c.get();
    a.exponents = e.get();
    a.numTermsPerReg = &numTerms;
    a.overrideInds = overrideInds;
    a.overridePhases = ph.get();
    a.numOverrides = withOverrides && overrideInds ? numOverrides : 0;
    phaseFuncPasses(qureg, a);
}

static void multiVarImpl(const Qureg &qureg, int *qubits, int *numQubitsPerReg, int numRegs, int encoding,
                         qreal *coeffs, qreal *exponents, int *numTermsPerReg, long long *overrideInds,
                         qreal *overridePhases, int numOverrides, bool withOverrides, const char *caller) {
    qv::qubitSubregs(qureg, qubits, numQubitsPerReg, numRegs, caller);
    qv::multiRegBitEncoding(numQubitsPerReg, numRegs, encoding, caller);
    // the term counts are read only after the checks QuEST makes before them
    size_t totalTerms = 0;
    if (numTermsPerReg)
        for (int r = 0; r < numRegs; ++r) totalTerms += (size_t)std::max(0, numTermsPerReg[r]);
    const Doubles c(coeffs, totalTerms), e(exponents, totalTerms);
    const Doubles ph(overridePhases, (size_t)std::max(0, numOverrides));
    qv::multiVarPhaseFuncTerms(numRegs, encoding, e.get(), numTermsPerReg, caller);
    if (withOverrides)
        qv::multiVarPhaseFuncOverrides(numQubitsPerReg, numRegs, encoding, overrideInds, numOverrides, caller);
    qmx::PhaseFuncArgs a{};
    a.numRegs = numRegs;
    a.encoding = encoding;
    a.mode = 0;
    a.qubits = qubits;
    a.numQubitsPerReg = numQubitsPerReg;
    a.coeffs = c.get();
    a.exponents = e.get();
    a.numTermsPerReg = numTermsPerReg;
    a.overrideInds = overrideInds;
    a.overridePhases = ph.get();
    a.numOverrides = withOverrides && overrideInds ? numOverrides : 0;
    phaseFuncPasses(qureg, a);
}

static void namedImpl(const Qureg &qureg, int *qubits, int *numQubitsPerReg, int numRegs, int encoding, int code,
                      qreal *params, int numParams, long long *overrideInds, qreal *overridePhases, int numOverrides,
                      bool withOverrides, const char *caller) {
    const Doubles pr(params, (size_t)std::max(0, numParams));
    const Doubles ph(overridePhases, (size_t)std::max(0, numOverrides));
    qv::qubitSubregs(qureg, qubits, numQubitsPerReg, numRegs, caller);
    qv::multiRegBitEncoding(numQubitsPerReg, numRegs, encoding, caller);
    qv::phaseFuncName(code, numRegs, numParams, caller);
    if (withOverrides)
        qv::multiVarPhaseFuncOverrides(numQubitsPerReg, numRegs, encoding, overrideInds, numOverrides, caller);
    qmx::PhaseFuncArgs a{};
    a.numRegs = numRegs;
    a.encoding = encoding;
    a.mode = 1;
    a.funcCode = code;
    a.qubits = qubits;
    a.numQubitsPerReg = numQubitsPerReg;
    a.params = pr.get();
    a.numParams = params ? numParams : 0;
    a.overrideInds = overrideInds;
    a.overridePhases = ph.get();
    a.numOverrides = withOverrides && overrideInds ? numOverrides : 0;
    phaseFuncPasses(qureg, a);
}

void applyPhaseFunc(Qureg qureg, int *qubits, int numQubits, enum bitEncoding encoding, qreal *coeffs,
                    qreal *exponents, int numTerms) {
    API_TRY
    phaseFuncImpl(qureg, qubits, numQubits, encoding, coeffs, exponents, numTerms, nullptr, nullptr, 0, false, __func__);
    API_CATCH_VOID
}

void applyPhaseFuncOverrides(Qureg qureg, int *qubits, int numQubits, enum bitEncoding encoding, qreal *coeffs,
                             qreal *exponents, int numTerms, long long int *overrideInds, qreal *overridePhases,
                             int numOverrides) {
    API_TRY
    phaseFuncImpl(qureg, qubits, numQubits, encoding, coeffs, exponents, numTerms, overrideInds, overridePhases,
                  numOverrides, true, __func__);
    API_CATCH_VOID
}

void applyMultiVarPhaseFunc(Qureg qureg, int *qubits, int *numQubitsPerReg, int numRegs, enum bitEncoding encoding,
                            qreal *coeffs, qreal *exponents, int *numTermsPerReg) {
    API_TRY
    multiVarImpl(qureg, qubits, numQubitsPerReg, numRegs, encoding, coeffs, exponents, numTermsPerReg, nullptr,
                 nullptr, 0, false, __func__);
    API_CATCH_VOID
}

void applyMultiVarPhaseFuncOverrides(Qureg qureg, int *qubits, int *numQubitsPerReg, int numRegs,
                                     enum bitEncoding encoding, qreal *coeffs, qreal *exponents, int *numTermsPerReg,
                                     long long int *overrideInds, qreal *overridePhases, int numOverrides) {
    API_TRY
    multiVarImpl(qureg, qubits, numQubitsPerReg, numRegs, encoding, coeffs, exponents, numTermsPerReg, overrideInds,
                 overridePhases, numOverrides, true, __func__);
    API_CATCH_VOID
}

void applyNamedPhaseFunc(Qureg qureg, int *qubits, int *numQubitsPerReg, int numRegs, enum bitEncoding encoding,
                         enum phaseFunc functionNameCode) {
    API_TRY
    namedImpl(qureg, qubits, numQubitsPerReg, numRegs, encoding, functionNameCode, nullptr, 0, nullptr, nullptr, 0,
              false, __func__);
    API_CATCH_VOID
}

void applyNamedPhaseFuncOverrides(Qureg qureg, int *qubits, int *numQubitsPerReg, int numRegs,
                                  enum bitEncoding encoding, enum phaseFunc functionNameCode,
                                  long long int *overrideInds, qreal *overridePhases, int numOverrides) {
    API_TRY
    namedImpl(qureg, qubits, numQubitsPerReg, numRegs, encoding, functionNameCode, nullptr, 0, overrideInds,
              overridePhases, numOverrides, true, __func__);
    API_CATCH_VOID
}

void applyParamNamedPhaseFunc(Qureg qureg, int *qubits, int *numQubitsPerReg, int numRegs, enum bitEncoding encoding,
                              enum phaseFunc functionNameCode, qreal *params, int numParams) {
    API_TRY
    namedImpl(qureg, qubits, numQubitsPerReg, numRegs, encoding, functionNameCode, params, numParams, nullptr, nullptr,
              0, false, __func__);
    API_CATCH_VOID
}

void applyParamNamedPhaseFuncOverrides(Qureg qureg, int *qubits, int *numQubitsPerReg, int numRegs,
                                       enum bitEncoding encoding, enum phaseFunc functionNameCode, qreal *params,
                                       int numParams, long long int *overrideInds, qreal *overridePhases,
                                       int numOverrides) {
    API_TRY
    namedImpl(qureg, qubits, numQubitsPerReg, numRegs, encoding, functionNameCode, params, numParams, overrideInds,
              overridePhases, numOverrides, true, __func__);
    API_CATCH_VOID
}

void applyQFT(Qureg qureg, int *qubits, int numQubits) {
    API_TRY
    qv::multiTargets(qureg, qubits, numQubits, __func__);
    qftOn(qureg, qubits, numQubits);
    API_CATCH_VOID
}

void applyFullQFT(Qureg qureg) {
    API_TRY
    std::vector<int> qs(qureg.numQubitsRepresented);
    for (int i = 0; i < qureg.numQubitsRepresented; ++i) qs[i] = i;
    qftOn(qureg, qs.data(), qureg.numQubitsRepresented);
    API_CATCH_VOID
}

/* ---- C-only helpers for variable-length 2D arrays (QuEST.c:1425-1434, QuEST_common.c:649-664) ---
 * Declared for C callers only (qreal re[][1 << n] decays to a pointer to rows of
 * 2^n qreals: the same ABI as the flat row-major pointer taken here). */
void initComplexMatrixN(ComplexMatrixN m, qreal *re, qreal *im) {
    API_TRY
    qv::matrixInit(m, __func__);
    const int dim = 1 << m.numQubits;
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j) {
            m.real[i][j] = re[(size_t)i * dim + j];
            m.imag[i][j] = im[(size_t)i * dim + j];
        }
    API_CATCH_VOID
}

ComplexMatrixN bindArraysToStackComplexMatrixN(int numQubits, qreal *re, qreal *im, qreal **reStorage,
                                               qreal **imStorage) {
    ComplexMatrixN m;
    m.numQubits = numQubits;
    m.real = reStorage;
    m.imag = imStorage;
    const int dim = 1 << numQubits;
    for (int i = 0; i < dim; ++i) {
        m.real[i] = re + (size_t)i * dim;
        m.imag[i] = im + (size_t)i * dim;
    }
    return m;
}

void qmxSync(void) { qmx::deviceSync(); }
void qmxDensityFusionStats(long long *sweeps, long long *operations) { qmx::densSuperOpStats(sweeps, operations); }

// QuantumLab's amplitude encoding (AmplitudeEncodeCuda.cu:498-533: angle tree,
// then an Ry and a Gray-code ladder of CNOT + Ry per level, one gate launch
// each) as one state preparation: the product of the tree's cos/sin factors
// along each basis state's path, written into qubits 0..numQubits-1.
void qmxAmplitudeEncodeReal(Qureg qureg, const qreal *vec, int numQubits) {
    API_TRY
    qv::stateVecQureg(qureg, __func__);
    if (numQubits < 1 || numQubits > 12 || numQubits > qureg.numQubitsRepresented)
        qv::raise("Invalid number of qubits to encode. Must be >0, <=12 and <=numQubits.", __func__);
    if (!vec) qv::raise("The vector is null.", __func__);
    Doubles v(vec, (size_t)1 << numQubits);
    qmx::amplitudeEncode(D(qureg), v.get(), numQubits);
    API_CATCH_VOID
}

// QuantumLab's complex-vector encoder (AmplitudeEncodeOneVector, bHasPhase
// false: AmplitudeEncodeCuda.cu:691-724, ExchangeToYZGate :452-496, FRyz
// FRGate.cu:400-460) as one state preparation: magnitude and phase trees, then
// the product of each basis state's cos/sin factors and phase terms.
void qmxAmplitudeEncodeComplex(Qureg qureg, const qreal *re, const qreal *im, int numQubits) {
    API_TRY
    qv::stateVecQureg(qureg, __func__);
    if (numQubits < 1 || numQubits > 12 || numQubits > qureg.numQubitsRepresented)
        qv::raise("Invalid number of qubits to encode. Must be >0, <=12 and <=numQubits.", __func__);
    if (!re || !im) qv::raise("The vector is null.", __func__);
    const size_t n = (size_t)1 << numQubits;
    Doubles r(re, n), i(im, n);
    qmx::amplitudeEncodeComplex(D(qureg), r.get(), i.get(), numQubits);
    API_CATCH_VOID
}

namespace {
// controls valid, distinct and apart from the target; 0 to 16 of them
void uniformControls(const Qureg &q, const int *controls, int numControls, int target, const char *f) {
    if (numControls < 0 || numControls > 16)
        qv::raise("Invalid number of control qubits for a uniformly controlled gate. Must be >=0 and <=16.", f);
    if (numControls == 0) {
        qv::target(q, target, f);
        return;
    }
    if (!controls) qv::raise("The control qubit list is null.", f);
    qv::multiControlsTarget(q, controls, numControls, target, f);
}

// ops: 4 entries (row-major) per control value; density registers take the
// conjugates on the column qubits
void uniform2(const Qureg &q, const int *controls, int numControls, int target, const std::vector<Cplx> &ops) {
    qmx::applyUniform2(D(q), controls, numControls, target, ops.data());
    if (q.isDensityMatrix) {
        const int s = q.numQubitsRepresented;
        std::vector<int> cs(controls, controls + numControls);
        for (int &c : cs) c += s;
        std::vector<Cplx> oc(ops.size());
        for (size_t e = 0; e < ops.size(); ++e) oc[e] = conjC(ops[e]);
        qmx::applyUniform2(D(q), cs.data(), numControls, target + s, oc.data());
    }
}
}  // namespace

// A uniformly controlled (multiplexed) one-qubit gate: ops[k] on the target
// where the controls read k. One sweep; QuantumLab composes these from 2^c
// gates (FRGate.cu).
void qmxUniformlyControlledUnitary(Qureg qureg, const int *controls, int numControls, int target,
                                   const ComplexMatrix2 *ops) {
    API_TRY
    uniformControls(qureg, controls, numControls, target, __func__);
    if (!ops) qv::raise("The matrix list is null.", __func__);
    const size_t count = (size_t)1 << numControls;
    std::vector<Cplx> m(4 * count);
    for (size_t k = 0; k < count; ++k) {
        qv::oneQubitUnitary(ops[k], __func__);
        m2ToCplx(ops[k], &m[4 * k]);
    }
    uniform2(qureg, controls, numControls, target, m);
    API_CATCH_VOID
}

// FRy / FRz / FRp (FRGate.cu:181-330, 460-515) in one sweep: Ry(angles[k]) as
// rotateY, Rz(angles[k]) as rotateZ, or diag(1, e^{i angles[k]}) as phaseShift,
// where the controls read k.
void qmxUniformlyControlledRotation(Qureg qureg, const int *controls, int numControls, int target,
                                    const qreal *angles, enum qmxUniformGate kind) {
    API_TRY
    uniformControls(qureg, controls, numControls, target, __func__);
    if (kind != QMX_UNIFORM_RY && kind != QMX_UNIFORM_RZ && kind != QMX_UNIFORM_PHASE)
        qv::raise("Invalid uniformly controlled gate kind. Must be QMX_UNIFORM_RY, QMX_UNIFORM_RZ or "
                  "QMX_UNIFORM_PHASE.", __func__);
    if (!angles) qv::raise("The angle list is null.", __func__);
    const size_t count = (size_t)1 << numControls;
    std::vector<Cplx> m(4 * count);
    for (size_t k = 0; k < count; ++k) {
        const double a = (double)angles[k];
        Cplx *u = &m[4 * k];
        if (kind == QMX_UNIFORM_RY) {
            const double c = std::cos(a / 2), sn = std::sin(a / 2);
            u[0] = Cplx{c, 0}, u[1] = Cplx{-sn, 0}, u[2] = Cplx{sn, 0}, u[3] = Cplx{c, 0};
        } else if (kind == QMX_UNIFORM_RZ) {
            u[0] = Cplx{std::cos(a / 2), -std::sin(a / 2)}, u[1] = Cplx{0, 0}, u[2] = Cplx{0, 0};
            u[3] = Cplx{std::cos(a / 2), std::sin(a / 2)};
        } else {
            u[0] = Cplx{1, 0}, u[1] = Cplx{0, 0}, u[2] = Cplx{0, 0}, u[3] = Cplx{std::cos(a), std::sin(a)};
        }
    }
    uniform2(qureg, controls, numControls, target, m);
    API_CATCH_VOID
}

// Repeated measurement without re-uploading the state. QuantumLab's shot loop
// (QLSimulatorMeasure.cpp:113-140) restores the register from the host before
// every shot and calls measure() on each listed qubit in order, so shot s is a
// chain of conditional outcomes, each drawn by generateMeasurementOutcome from
// QuEST's generator. Here the joint distribution of the listed qubits is
// reduced once on the device (calcProbOfAllOutcomes), and each shot replays
// that chain on the host: P(qubit j = 0 | earlier outcomes) is the partial sum
// of the entries that extend the prefix, scaled by 1 / (product of the earlier
// outcome probabilities) exactly as collapse-and-renormalise scales the state.
// The same draws are consumed, so outcomes and the generator's later state
// match the loop's (a draw could only differ if it fell within rounding of P0).
int qmxSampleShots(Qureg qureg, const int *qubits, int numQubits, int shots, int untilCount, int *outcomes) {
    API_TRY
    qv::multiTargets(qureg, qubits, numQubits, __func__);
    if (numQubits > 26) qv::raise("At most 26 qubits may be sampled jointly.", __func__);
    if (shots < 0) qv::raise("The number of shots must be >= 0.", __func__);
    if (shots > 0 && !outcomes) qv::raise("The outcome array is null.", __func__);
    const int k = numQubits;
    // level d holds, for every setting of qubits[0..d-1], the probability mass
    // of the entries that extend it (index bit j = outcome of qubits[j])
    std::vector<std::vector<double>> level(k + 1);
    level[k].resize((size_t)1 << k);
    qmx::probOfAllOutcomes(D(qureg), qubits, k, level[k].data());
    for (int d = k - 1; d >= 0; --d) {
        level[d].resize((size_t)1 << d);
        for (size_t p = 0; p < level[d].size(); ++p) level[d][p] = level[d + 1][p] + level[d + 1][p | ((size_t)1 << d)];
    }
    std::vector<long long> counts(untilCount > 0 ? ((size_t)1 << k) : 0, 0);
    int drawn = 0;
    while (drawn < shots) {
        size_t prefix = 0;
        double scale = 1.0;
        for (int j = 0; j < k; ++j) {
            double outcomeProb;
            const int o = generateMeasurementOutcome(level[j + 1][prefix] * scale, &outcomeProb);
            prefix |= (size_t)o << j;
            scale /= outcomeProb;
        }
        outcomes[drawn++] = (int)prefix;
        if (untilCount > 0 && ++counts[prefix] >= untilCount) break;
    }
    return drawn;
    API_CATCH(0)
}
void qmxTimerBegin(void) { qmx::timerBegin(); }
double qmxTimerEndMs(void) { return qmx::timerEndMs(); }

qreal qmxRenormalise(Qureg qureg, int rescale) { return qmx::normSquaredThenScale(D(qureg), rescale != 0); }

// ---- wave-function CSV files ------------------------------------------------------
// QuantumLab's format (FermionSimulation/Fermion1D.cpp:13-83): one amplitude per
// line, "%f, %f" (real, imaginary), no header. A line is read as two fields
// split on ',', blanks and tabs around them ignored; anything else is an error.

int qmxSaveStateCSV(Qureg qureg, const char *path) {
    API_TRY
    FILE *f = std::fopen(path, "w");
    qv::fileOpened(f != nullptr, path, __func__);
    const long long n = qureg.numAmpsTotal;
    constexpr long long kChunk = 1 << 20;
    std::vector<double> re((size_t)std::min(n, kChunk)), im(re.size());
    for (long long s = 0; s < n; s += kChunk) {
        const long long c = std::min(kChunk, n - s);
        qmx::downloadAmps(D(qureg), s, re.data(), im.data(), c);
        for (long long i = 0; i < c; ++i) std::fprintf(f, "%f, %f\n", (qreal)re[i], (qreal)im[i]);
    }
    std::fclose(f);
    return 1;
    API_CATCH(0)
}

namespace {

// one "re, im" line; false if it is not exactly two numeric fields
bool csvAmp(const std::string &line, double *re, double *im) {
    const size_t comma = line.find(',');
    if (comma == std::string::npos || line.find(',', comma + 1) != std::string::npos) return false;
    auto field = [](std::string t, double *out) {
        const size_t a = t.find_first_not_of(" \t\r");
        if (a == std::string::npos) return false;
        const size_t b = t.find_last_not_of(" \t\r");
        t = t.substr(a, b - a + 1);
        char *end = nullptr;
        *out = std::strtod(t.c_str(), &end);
        return end == t.c_str() + t.size();
    };
    return field(line.substr(0, comma), re) && field(line.substr(comma + 1), im);
}

}  // namespace

long long qmxLoadStateCSV(Qureg qureg, const char *path) {
    API_TRY
    FILE *f = std::fopen(path, "r");
    qv::fileOpened(f != nullptr, path, __func__);
    std::vector<double> re, im;
    std::string line;
    int ch;
    bool ok = true;
    auto takeLine = [&]() {
        double r, i;
        if (!csvAmp(line, &r, &i)) ok = false;
        re.push_back(r);
        im.push_back(i);
        line.clear();
    };
    while ((ch = std::fgetc(f)) != EOF) {
        if (ch == '\n') takeLine();
        else line.push_back((char)ch);
        if (!ok) break;
    }
    if (ok && !line.empty()) takeLine();
    std::fclose(f);
    if (!ok) qv::raise("Invalid wave-function file. Each line must be two numbers separated by a comma.", __func__);
    if ((long long)re.size() != qureg.numAmpsTotal)
        qv::raise("The wave-function file's number of amplitudes does not match the register.", __func__);
    qmx::uploadAmps(D(qureg), 0, re.data(), im.data(), (long long)re.size());
    return (long long)re.size();
    API_CATCH(0)
}

}  // extern "C"
