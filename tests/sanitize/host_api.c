/* The C API's host code under AddressSanitizer and UndefinedBehaviorSanitizer.
 *
 * One source, two builds (tests/sanitize/Makefile):
 *   host_api_cpuref  linked to the oracle's C restatement (oracle/cpuref/cpuref.c):
 *                    every gate, channel, reduction, measurement and state
 *                    initialisation the oracle implements, on a 5-qubit state
 *                    vector and a 3-qubit density register, and its generator
 *                    and swap-test closed form;
 *   host_api_lib     linked to libquest_mi355x's host code (its .cpp files built
 *                    with the sanitizers, its .hip objects as built; also at
 *                    QuEST_PREC=1 as host_api_lib_f32): what runs
 *                    without a device — circuit recording and fusion, PauliHamil
 *                    files, matrices, the planner's inputs, seeding — and the
 *                    rejection of bad input
 *                    through a returning error hook.
 * The pass condition is exit status 0 with no sanitizer report. */
#define _POSIX_C_SOURCE 200809L  /* mkdtemp */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "QuEST.h"

static void fail(const char *what, int line) {
    fprintf(stderr, "host_api: check failed at line %d: %s\n", line, what);
    exit(2);
}
#define CHECK(cond) ((cond) ? (void)0 : fail(#cond, __LINE__))

#ifdef HOST_LIB
#include "quest_mi355x.h"

/* QuEST's hook (QuEST_validation.c:240-245), returning: the library then
 * unwinds the call that failed and returns to the caller. */
static int numErrors = 0;
static char lastError[512];
void invalidQuESTInputError(const char *errMsg, const char *errFunc) {
    ++numErrors;
    snprintf(lastError, sizeof lastError, "%s: %s", errFunc, errMsg);
}
#define EXPECT_ERROR(stmt)                                  \
    do {                                                    \
        int before_ = numErrors;                            \
        stmt;                                               \
        if (numErrors != before_ + 1) fail(#stmt, __LINE__); \
    } while (0)
#define EXPECT_OK(stmt)                                     \
    do {                                                    \
        int before_ = numErrors;                            \
        stmt;                                               \
        if (numErrors != before_) {                         \
            fprintf(stderr, "host_api: %s\n", lastError);   \
            fail(#stmt, __LINE__);                          \
        }                                                   \
    } while (0)

static ComplexMatrix2 hadamardMatrix(void) {
    const qreal s = (qreal)(1 / sqrt(2.0));
    ComplexMatrix2 u;
    memset(&u, 0, sizeof u);
    u.real[0][0] = s, u.real[0][1] = s, u.real[1][0] = s, u.real[1][1] = -s;
    return u;
}

/* fusion at every block width, each step read back in full and with NULL outputs */
static void fuseAtEveryWidth(qmxCircuit *c, int n, int recorded) {
    static qreal bre[1 << 20], bim[1 << 20];
    int targs[64];
    for (int k = 1; k <= 10; ++k) {
        int steps = 0;
        EXPECT_OK(steps = qmxCircuitFuse(c, k));
        CHECK(steps > 0 && steps <= recorded);
        for (int s = 0; s < steps; ++s) {
            const int w = qmxCircuitStep(c, s, targs, bre, bim);
            CHECK(w >= 0 && w <= (k > 5 ? k : 5) && w <= n);
            for (int i = 1; i < w; ++i) CHECK(targs[i - 1] < targs[i]);
            for (int i = 0; i < w; ++i) CHECK(targs[i] >= 0 && targs[i] < n);
            CHECK(qmxCircuitStep(c, s, NULL, NULL, NULL) == w);
            if (w > 0) {  /* each block is unitary: its first column has unit norm */
                const int dim = 1 << w;
                double norm = 0;
                for (int r = 0; r < dim; ++r) norm += bre[r * dim] * bre[r * dim] + bim[r * dim] * bim[r * dim];
                CHECK(fabs(norm - 1) < (sizeof(qreal) == 8 ? 1e-12 : 1e-5));
            }
        }
        CHECK(qmxCircuitStep(c, steps, targs, bre, bim) == -1);
        CHECK(qmxCircuitNumBlocks(c) <= steps);
    }
}

/* every operation code, the typed recorders, fusion at every width, and the
 * records a circuit must refuse */
static void circuits(void) {
    const int n = 10;
    qmxCircuit *c = qmxCircuitCreate(n);
    CHECK(c != NULL);
    for (int rep = 0; rep < 3; ++rep)
        for (int op = QMX_OP_H; op <= QMX_OP_COMPOSITE; ++op) {
            const int two = (op >= QMX_OP_CX && op <= QMX_OP_CRZ) || op == QMX_OP_NOISE_TWO_QUBIT_DEPHASING ||
                            op == QMX_OP_NOISE_TWO_QUBIT_DEPOLARISING;
            int qs[2] = {(op * 3 + rep) % n, (op * 7 + rep + 1) % n};
            if (qs[0] == qs[1]) qs[1] = (qs[1] + 1) % n;
            qreal param = (qreal)(0.01 + 0.001 * op); /* angles, and probabilities every channel accepts */
            if (op == QMX_OP_CC) param = (qreal)(rep & 1);
            EXPECT_OK(qmxCircuitAddOperation(c, op, qs, two ? 2 : 1, param));
        }
    qmxNoiseAfterGate noise = {(qreal)0.01, (qreal)0.02, (qreal)0.03, (qreal)0.01, (qreal)0.02, (qreal)0.03};
    EXPECT_OK(qmxCircuitSetNoiseAfterGate(c, noise));
    Complex a = {(qreal)0.6, 0}, b = {0, (qreal)0.8};
    EXPECT_OK(qmxCircuitCompactUnitary(c, 9, a, b));
    EXPECT_OK(qmxCircuitUnitary(c, 0, hadamardMatrix()));
    EXPECT_OK(qmxCircuitRotateZ(c, 4, (qreal)0.3));
    EXPECT_OK(qmxCircuitControlledNot(c, 8, 1));
    int ctrls[3] = {2, 5, 7};
    EXPECT_OK(qmxCircuitMultiControlledUnitary(c, ctrls, 3, 9, hadamardMatrix()));
    /* a 5-target permutation (exactly unitary): index i -> (i * 5 + 3) mod 32 */
    qreal re[32 * 32], im[32 * 32];
    memset(re, 0, sizeof re);
    memset(im, 0, sizeof im);
    for (int i = 0; i < 32; ++i) re[((i * 5 + 3) % 32) * 32 + i] = 1;
    int targs5[5] = {9, 0, 4, 6, 2};
    EXPECT_OK(qmxCircuitMultiQubitUnitary(c, targs5, 5, re, im));
    const int recorded = qmxCircuitNumGates(c);
    CHECK(recorded == 3 * 27 + 6);

    /* refusals */
    int bad[2] = {n, 0};
    EXPECT_ERROR(qmxCircuitAddOperation(c, QMX_OP_H, bad, 1, 0));
    bad[0] = -1;
    EXPECT_ERROR(qmxCircuitAddOperation(c, QMX_OP_X, bad, 1, 0));
    int same[2] = {3, 3};
    EXPECT_ERROR(qmxCircuitAddOperation(c, QMX_OP_CX, same, 2, 0));
    int ok2[2] = {1, 2};
    EXPECT_ERROR(qmxCircuitAddOperation(c, QMX_OP_COMPOSITE + 1, ok2, 1, 0));
    EXPECT_ERROR(qmxCircuitAddOperation(c, -1, ok2, 1, 0));
    EXPECT_ERROR(qmxCircuitAddOperation(c, QMX_OP_CX, ok2, 1, 0));
    EXPECT_ERROR(qmxCircuitControlledNot(c, 4, 4));
    EXPECT_ERROR(qmxCircuitRotateZ(c, n, 1));
    re[0] = 2;  /* no longer unitary */
    EXPECT_ERROR(qmxCircuitMultiQubitUnitary(c, targs5, 5, re, im));
    int dupTargs[5] = {9, 0, 4, 9, 2};
    re[0] = 0;
    EXPECT_ERROR(qmxCircuitMultiQubitUnitary(c, dupTargs, 5, re, im));
    CHECK(qmxCircuitNumGates(c) == recorded);
    EXPECT_ERROR(qmxCircuitFuse(c, 0));
    EXPECT_ERROR(qmxCircuitFuse(c, 11));
    qmxCircuit *none = NULL;
    EXPECT_ERROR(none = qmxCircuitCreate(0));
    CHECK(none == NULL);

    fuseAtEveryWidth(c, n, recorded);
    qmxCircuitDestroy(c);
}

/* circuits narrower than a block (padding cannot reach five qubits) and wider
 * than 32 qubits (64-bit masks), fused at every width */
static void circuitWidths(void) {
    static const int widths[] = {1, 2, 3, 5, 6, 33, 40};
    qreal re[32 * 32], im[32 * 32];
    memset(re, 0, sizeof re);
    memset(im, 0, sizeof im);
    for (int i = 0; i < 32; ++i) re[((i * 7 + 1) % 32) * 32 + i] = 1;
    for (size_t w = 0; w < sizeof widths / sizeof widths[0]; ++w) {
        const int n = widths[w];
        qmxCircuit *c = qmxCircuitCreate(n);
        CHECK(c != NULL);
        for (int i = 0; i < 40; ++i) {
            const int t = (i * 5) % n, u = (t + 1 + i % 3) % n;
            int qs[2] = {t, u};
            EXPECT_OK(qmxCircuitAddOperation(c, QMX_OP_H, qs, 1, 0));
            EXPECT_OK(qmxCircuitRotateZ(c, t, (qreal)(0.1 * i)));
            if (n > 1 && u != t) EXPECT_OK(qmxCircuitAddOperation(c, QMX_OP_CRY, qs, 2, (qreal)0.2));
            if (n >= 4 && i % 7 == 0) {
                int ctrls[3] = {(t + 1) % n, (t + 2) % n, (t + 3) % n};
                EXPECT_OK(qmxCircuitMultiControlledUnitary(c, ctrls, 3, t, hadamardMatrix()));
            }
            if (n >= 5 && i % 11 == 0) {
                int targs5[5] = {0, n / 4, n / 2, (3 * n) / 4, n - 1};
                if (n == 5) targs5[1] = 1, targs5[2] = 2, targs5[3] = 3;
                EXPECT_OK(qmxCircuitMultiQubitUnitary(c, targs5, 5, re, im));
            }
            if (i == 20) EXPECT_OK(qmxCircuitAddOperation(c, QMX_OP_CC, qs, 1, 1));
        }
        fuseAtEveryWidth(c, n, qmxCircuitNumGates(c));
        qmxCircuitDestroy(c);
    }
}

static void writeFile(const char *path, const char *text) {
    FILE *f = fopen(path, "wb");
    CHECK(f != NULL);
    fputs(text, f);
    fclose(f);
}

/* PauliHamil files (QuEST.c:1459-1541): layouts it accepts and files it refuses */
static void hamilFiles(const char *dir) {
    static const struct {
        const char *text;
        int qubits, terms;
        double lastCoeff;
    } good[] = {
        {"0.5 1 2 3\n-0.25 0 0 3\n", 3, 2, -0.25},
        {"0.5 1 2 3\n-0.25 0 0 3", 3, 2, -0.25},            /* no final newline */
        {"1.5 3 3\n2.5 1 1\n", 2, 2, 2.5},
        {"1e-3 0\n", 1, 1, 1e-3},
        {"0.125 1 0 2 3 1 2 0 3\n", 8, 1, 0.125},
    };
    static const char *bad[] = {
        "0.5 1 4 3\n",        /* Pauli code out of range */
        "0.5 1 x 3\n",        /* code that is not a number */
        "abc 1 2 3\n",        /* coefficient that is not a number */
        "0.5 1 2 3\n0.5 1\n", /* short second term */
        "",                   /* empty file */
    };
    char path[1024];
    snprintf(path, sizeof path, "%s/h.txt", dir);
    for (size_t i = 0; i < sizeof good / sizeof good[0]; ++i) {
        writeFile(path, good[i].text);
        PauliHamil h;
        EXPECT_OK(h = createPauliHamilFromFile(path));
        CHECK(h.numQubits == good[i].qubits && h.numSumTerms == good[i].terms);
        CHECK(fabs(h.termCoeffs[h.numSumTerms - 1] - good[i].lastCoeff) < 1e-6);
        for (int t = 0; t < h.numQubits * h.numSumTerms; ++t) CHECK(h.pauliCodes[t] >= 0 && h.pauliCodes[t] <= 3);
        destroyPauliHamil(h);
    }
    for (size_t i = 0; i < sizeof bad / sizeof bad[0]; ++i) {
        writeFile(path, bad[i]);
        PauliHamil h;
        EXPECT_ERROR(h = createPauliHamilFromFile(path));
        CHECK(h.termCoeffs == NULL && h.pauliCodes == NULL);
    }
    unlink(path);
    snprintf(path, sizeof path, "%s/absent.txt", dir);
    EXPECT_ERROR(createPauliHamilFromFile(path));

    PauliHamil h;
    EXPECT_OK(h = createPauliHamil(3, 2));
    qreal coeffs[2] = {(qreal)0.5, (qreal)-1};
    enum pauliOpType codes[6] = {PAULI_X, PAULI_I, PAULI_Z, PAULI_Y, PAULI_Y, PAULI_I};
    EXPECT_OK(initPauliHamil(h, coeffs, codes));
    CHECK(h.pauliCodes[3] == PAULI_Y && h.termCoeffs[1] == -1);
    codes[2] = (enum pauliOpType)4;
    EXPECT_ERROR(initPauliHamil(h, coeffs, codes));
    destroyPauliHamil(h);
    EXPECT_ERROR(createPauliHamil(0, 1));
    EXPECT_ERROR(createPauliHamil(2, 0));
}

static void matrices(void) {
    ComplexMatrixN m;
    EXPECT_OK(m = createComplexMatrixN(3));
    qreal re[8][8], im[8][8];
    for (int i = 0; i < 64; ++i) re[i / 8][i % 8] = (qreal)i, im[i / 8][i % 8] = (qreal)-i;
    EXPECT_OK(initComplexMatrixN(m, re, im));
    CHECK(m.real[7][5] == 61 && m.imag[2][3] == -19);
    destroyComplexMatrixN(m);
    EXPECT_ERROR(createComplexMatrixN(0));
}

/* the Trotter planner's inputs from a Hamiltonian, and a code array holding a
 * value outside 0..3, which it refuses */
static void plannerFeatures(void) {
    PauliHamil h;
    EXPECT_OK(h = createPauliHamil(6, 4));
    for (int t = 0; t < 4; ++t) h.termCoeffs[t] = (qreal)(0.3 * (t + 1));
    for (int i = 0; i < 24; ++i) h.pauliCodes[i] = (enum pauliOpType)((i * 5 + 2) % 4);
    float f[16];
    CHECK(qmxTrotterFeatures(h, (qreal)0.5, f) == 1);
    for (int i = 0; i < 16; ++i) CHECK(isfinite(f[i]));
    h.pauliCodes[7] = (enum pauliOpType)7;
    CHECK(qmxTrotterFeatures(h, (qreal)0.5, f) == 0);
    destroyPauliHamil(h);
}

/* QuEST's seed array (QuEST.c:1597-1626) and the generator state it feeds */
static void seeding(void) {
    QuESTEnv env;
    memset(&env, 0, sizeof env);
    unsigned long keys[3] = {1234, 5678, 42};
    seedQuEST(&env, keys, 3);
    unsigned long *seeds = NULL;
    int num = 0;
    getQuESTSeeds(env, &seeds, &num);
    CHECK(num == 3 && seeds[2] == 42);
    seedQuEST(&env, keys, 1);
    getQuESTSeeds(env, &seeds, &num);
    CHECK(num == 1 && seeds[0] == 1234);
    destroyQuESTEnv(env);
}

int main(void) {
    const char *tmp = getenv("TMPDIR");
    char dir[900];
    snprintf(dir, sizeof dir, "%s/host_api_XXXXXX", tmp && *tmp ? tmp : "/tmp");
    CHECK(mkdtemp(dir) != NULL);
    circuits();
    circuitWidths();
    hamilFiles(dir);
    matrices();
    plannerFeatures();
    seeding();
    rmdir(dir);
    printf("host_api (library host code, qreal of %d bytes): %d rejections reported, all expected\n",
           (int)sizeof(qreal), numErrors);
    return 0;
}

#else /* the oracle */

static ComplexMatrix2 rotation(double t, double p) {
    ComplexMatrix2 u;
    memset(&u, 0, sizeof u);
    u.real[0][0] = cos(t), u.real[1][1] = cos(t) * cos(p), u.imag[1][1] = cos(t) * sin(p);
    u.real[0][1] = -sin(t), u.real[1][0] = sin(t) * cos(p), u.imag[1][0] = sin(t) * sin(p);
    /* columns: (cos t, sin t e^{ip}) and (-sin t, cos t e^{ip}) */
    return u;
}

static ComplexMatrix4 swapLike(void) {
    ComplexMatrix4 u;
    memset(&u, 0, sizeof u);
    u.real[0][0] = 1, u.real[1][2] = 1, u.real[2][1] = 1, u.imag[3][3] = 1;
    return u;
}

static ComplexMatrixN cyclic(int k) {
    ComplexMatrixN m = createComplexMatrixN(k);
    const int dim = 1 << k;
    for (int i = 0; i < dim; ++i) m.real[(i + 1) % dim][i] = 1;
    return m;
}

static void gates(Qureg q) {
    Complex a = {0.6, 0}, b = {0, 0.8};
    Vector axis = {1, 2, 2};
    int ctrls[3] = {0, 2, 4}, state[3] = {1, 0, 1}, targs[3] = {1, 3, 2};
    int all[5] = {0, 1, 2, 3, 4};
    enum pauliOpType paulis[3] = {PAULI_X, PAULI_Y, PAULI_Z};
    ComplexMatrixN m3 = cyclic(3), m2 = cyclic(2);
    const int n = q.numQubitsRepresented;
    for (int t = 0; t < n; ++t) {
        const int u = (t + 1) % n;
        hadamard(q, t);
        pauliX(q, t), pauliY(q, t), pauliZ(q, t), sGate(q, t), tGate(q, t);
        phaseShift(q, t, 0.3), controlledPhaseShift(q, u, t, 0.4), controlledPhaseFlip(q, u, t);
        rotateX(q, t, 0.5), rotateY(q, t, 0.6), rotateZ(q, t, 0.7), rotateAroundAxis(q, t, 0.8, axis);
        controlledRotateX(q, u, t, 0.1), controlledRotateY(q, u, t, 0.2), controlledRotateZ(q, u, t, 0.3);
        compactUnitary(q, t, a, b), controlledCompactUnitary(q, u, t, a, b);
        unitary(q, t, rotation(0.4, 0.9)), controlledUnitary(q, u, t, rotation(0.5, 1.1));
        controlledNot(q, u, t), controlledPauliY(q, u, t), swapGate(q, u, t);
        twoQubitUnitary(q, t, u, swapLike());
    }
    if (n >= 5) {
        multiControlledPhaseShift(q, ctrls, 3, 0.2), multiControlledPhaseFlip(q, ctrls, 3);
        multiControlledUnitary(q, ctrls, 2, 3, rotation(0.3, 0.2));
        multiStateControlledUnitary(q, ctrls, state, 3, 1, rotation(0.7, 0.1));
        multiQubitUnitary(q, targs, 3, m3);
        controlledMultiQubitUnitary(q, 0, targs, 3, m3);
        multiControlledMultiQubitUnitary(q, ctrls + 1, 2, targs, 2, m2);
        multiQubitNot(q, targs, 3), multiControlledMultiQubitNot(q, ctrls, 1, targs, 3);
        multiRotateZ(q, all, 5, 0.45), multiRotatePauli(q, targs, paulis, 3, 0.55);
    } else {
        int low[2] = {2, 0};
        multiQubitUnitary(q, low, 2, m2);
        multiRotatePauli(q, low, paulis, 2, 0.55);
    }
    destroyComplexMatrixN(m3);
    destroyComplexMatrixN(m2);
}

static void channels(Qureg rho) {
    ComplexMatrix2 kraus[2];
    memset(kraus, 0, sizeof kraus);
    kraus[0].real[0][0] = kraus[0].real[1][1] = sqrt(0.7);
    kraus[1].real[0][1] = kraus[1].real[1][0] = sqrt(0.3);
    ComplexMatrix4 kraus4[1] = {swapLike()};
    mixDamping(rho, 0, 0.1), mixDephasing(rho, 1, 0.2), mixDepolarising(rho, 2, 0.3);
    mixPauli(rho, 0, 0.1, 0.05, 0.2);
    mixTwoQubitDephasing(rho, 0, 2, 0.3), mixTwoQubitDepolarising(rho, 1, 2, 0.4);
    mixKrausMap(rho, 1, kraus, 2), mixTwoQubitKrausMap(rho, 2, 0, kraus4, 1);
}

static void reductions(Qureg q, QuESTEnv env) {
    const int n = q.numQubitsRepresented;
    CHECK(fabs(calcTotalProb(q) - 1) < 1e-12);
    for (int t = 0; t < n; ++t) CHECK(calcProbOfOutcome(q, t, 0) > -1e-12);
    qreal probs[8];
    int qs[3] = {2, 0, 1};
    calcProbOfAllOutcomes(probs, q, qs, 3);
    double sum = 0;
    for (int i = 0; i < 8; ++i) sum += probs[i];
    CHECK(fabs(sum - 1) < 1e-12);
    Qureg copy = createCloneQureg(q, env);
    if (q.isDensityMatrix) {
        CHECK(calcPurity(q) <= 1 + 1e-12);
        CHECK(fabs(calcDensityInnerProduct(q, copy) - calcPurity(q)) < 1e-12);
        CHECK(calcHilbertSchmidtDistance(q, copy) < 1e-12);
        Complex e = getDensityAmp(q, 1, 2);
        (void)e;
        mixDensityMatrix(copy, 0.25, q);
    } else {
        Complex ip = calcInnerProduct(q, copy);
        CHECK(fabs(ip.real - 1) < 1e-12 && fabs(ip.imag) < 1e-12);
        Complex amp = getAmp(q, 3);
        CHECK(amp.real == getRealAmp(q, 3) && amp.imag == getImagAmp(q, 3));
    }
    destroyQureg(copy, env);
}

static void measurements(Qureg q) {
    qreal p = 0;
    const int n = q.numQubitsRepresented;
    for (int t = 0; t < n; ++t) {
        const int out = measureWithStats(q, t, &p);
        CHECK(out == 0 || out == 1);
        CHECK(p > 0 && p <= 1 + 1e-12);
    }
    initPlusState(q);
    CHECK(fabs(collapseToOutcome(q, 0, 1) - 0.5) < 1e-12);
    (void)measure(q, 1);
}

static void initialisations(Qureg q, QuESTEnv env) {
    const long long dim = 1LL << q.numQubitsRepresented;
    initBlankState(q), initZeroState(q), initPlusState(q), initDebugState(q);
    initClassicalState(q, dim - 1);
    CHECK(fabs(calcProbOfOutcome(q, 0, 1) - 1) < 1e-12);
    Qureg pure = createQureg(q.numQubitsRepresented, env);
    initPlusState(pure);
    initPureState(q, pure);
    if (!q.isDensityMatrix) {
        qreal re[32], im[32];
        for (long long i = 0; i < dim; ++i) re[i] = (qreal)(i == 2), im[i] = 0;
        initStateFromAmps(q, re, im);
        CHECK(getRealAmp(q, 2) == 1);
        q.stateVec.real[1] = 0.5;  /* the host mirror, uploaded */
        copyStateToGPU(q);
        copyStateFromGPU(q);
        CHECK(q.stateVec.real[1] == 0.5);
        initStateFromAmps(q, re, im);
    }
    Qureg other = createCloneQureg(q, env);
    cloneQureg(other, q);
    destroyQureg(other, env);
    destroyQureg(pure, env);
    initZeroState(q);
}

static void trotter(Qureg q) {
    const int n = q.numQubitsRepresented;
    PauliHamil h = createPauliHamil(n, 3);
    const qreal coeffs[3] = {0.5, -0.3, 0.2};
    for (int t = 0; t < 3; ++t) h.termCoeffs[t] = coeffs[t];
    for (int i = 0; i < 3 * n; ++i) h.pauliCodes[i] = (enum pauliOpType)((i * 7 + 1) % 4);
    for (int order = 1; order <= 4; order *= 2) applyTrotterCircuit(q, h, 0.3, order, 2);
    destroyPauliHamil(h);
}

/* the oracle's extra entry points: its generator and the swap-test closed form
 * (ragged batch: registers of 1, 3 and 2 qubits) */
void cpurefMtInitByArray(const unsigned long *key, int len);
unsigned int cpurefMtNextU32(void);
int cpurefNumThreads(void);
void cpurefSwapP0ClosedForm(int numTests, const int *q, const double *a, const long long *aOff, const double *b,
                            const long long *bOff, double *p0);

static void extras(void) {
    unsigned long key[3] = {1, 2, 3};
    cpurefMtInitByArray(key, 3);
    unsigned int x = 0;
    for (int i = 0; i < 1000; ++i) x ^= cpurefMtNextU32();
    (void)x;
    CHECK(cpurefNumThreads() >= 1);
    const int q[3] = {1, 3, 2};
    const long long off[3] = {0, 2, 10};
    double a[14], b[14], p0[3];
    for (int i = 0; i < 14; ++i) a[i] = b[i] = 0;
    a[0] = b[1] = 1;                    /* orthogonal: P0 = 1/2 */
    for (int i = 2; i < 10; ++i) a[i] = b[i] = sqrt(1.0 / 8);  /* equal: P0 = 1 */
    a[10] = b[10] = 1;                  /* equal basis states: P0 = 1 */
    cpurefSwapP0ClosedForm(3, q, a, off, b, off, p0);
    CHECK(fabs(p0[0] - 0.5) < 1e-15 && fabs(p0[1] - 1) < 1e-12 && fabs(p0[2] - 1) < 1e-15);
}

int main(void) {
    extras();
    QuESTEnv env = createQuESTEnv();
    unsigned long keys[2] = {1234, 5678};
    seedQuEST(&env, keys, 2);
    Qureg sv = createQureg(5, env), rho = createDensityQureg(3, env);
    for (int pass = 0; pass < 2; ++pass) {
        Qureg q = pass ? rho : sv;
        initialisations(q, env);
        initPlusState(q);
        gates(q);
        if (q.isDensityMatrix) channels(q);
        trotter(q);
        reductions(q, env);
        measurements(q);
    }
    destroyQureg(sv, env);
    destroyQureg(rho, env);
    destroyQuESTEnv(env);
    printf("host_api (oracle): every oracle entry point ran clean\n");
    return 0;
}
#endif
