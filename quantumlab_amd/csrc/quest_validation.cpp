// Input validation for the QuEST C API front-end.
//
// Checks and message texts follow QuEST_validation.c:126-214 (messages) and
// :262-790 (checks); the texts are kept verbatim because callers and QuEST's
// own tests match on their substrings ("Invalid target", "unitary", ...).
// Every check runs before any state change. A failing check calls
// invalidQuESTInputError(msg, func): weak here, so a caller can replace it
// (QuEST_validation.c:240-245); if the handler returns, the API call is
// abandoned by throwing qmx::InputError, which every API entry catches.
#include "quest_validation.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace qv {

namespace {

const char *const kMessages[] = {
    /* E_SUCCESS */ "",
    /* E_INVALID_NUM_CREATE_QUBITS */ "Invalid number of qubits. Must create >0.",
    /* E_INVALID_QUBIT_INDEX */ "Invalid qubit index. Must be >=0 and <numQubits.",
    /* E_INVALID_TARGET_QUBIT */ "Invalid target qubit. Must be >=0 and <numQubits.",
    /* E_INVALID_CONTROL_QUBIT */ "Invalid control qubit. Must be >=0 and <numQubits.",
    /* E_INVALID_STATE_INDEX */ "Invalid state index. Must be >=0 and <2^numQubits.",
    /* E_INVALID_AMP_INDEX */ "Invalid amplitude index. Must be >=0 and <2^numQubits.",
    /* E_INVALID_NUM_AMPS */
    "Invalid number of amplitudes. Must be >=0 and <=2^numQubits (or for density matrices, <=2^(2 numQubits)).",
    /* E_INVALID_OFFSET_NUM_AMPS_QUREG */ "More amplitudes given than exist in the state from the given starting index.",
    /* E_TARGET_IS_CONTROL */ "Control qubit cannot equal target qubit.",
    /* E_TARGET_IN_CONTROLS */ "Control qubits cannot include target qubit.",
    /* E_CONTROL_TARGET_COLLISION */ "Control and target qubits must be disjoint.",
    /* E_QUBITS_NOT_UNIQUE */ "The qubits must be unique.",
    /* E_TARGETS_NOT_UNIQUE */ "The target qubits must be unique.",
    /* E_CONTROLS_NOT_UNIQUE */ "The control qubits should be unique.",
    /* E_INVALID_NUM_QUBITS */ "Invalid number of qubits. Must be >0 and <=numQubits.",
    /* E_INVALID_NUM_TARGETS */ "Invalid number of target qubits. Must be >0 and <=numQubits.",
    /* E_INVALID_NUM_CONTROLS */ "Invalid number of control qubits. Must be >0 and <numQubits.",
    /* E_NON_UNITARY_MATRIX */ "Matrix is not unitary.",
    /* E_NON_UNITARY_COMPLEX_PAIR */ "Compact matrix formed by given complex numbers is not unitary.",
    /* E_ZERO_VECTOR */ "Invalid axis vector. Must be non-zero.",
    /* E_COLLAPSE_STATE_ZERO_PROB */ "Can't collapse to state with zero probability.",
    /* E_INVALID_QUBIT_OUTCOME */ "Invalid measurement outcome -- must be either 0 or 1.",
    /* E_CANNOT_OPEN_FILE */ "Could not open file (%s).",
    /* E_SECOND_ARG_MUST_BE_STATEVEC */ "Second argument must be a state-vector.",
    /* E_MISMATCHING_QUREG_DIMENSIONS */ "Dimensions of the qubit registers don't match.",
    /* E_MISMATCHING_QUREG_TYPES */ "Registers must both be state-vectors or both be density matrices.",
    /* E_DEFINED_ONLY_FOR_STATEVECS */ "Operation valid only for state-vectors.",
    /* E_DEFINED_ONLY_FOR_DENSMATRS */ "Operation valid only for density matrices.",
    /* E_INVALID_PROB */ "Probabilities must be in [0, 1].",
    /* E_INVALID_ONE_QUBIT_DEPHASE_PROB */
    "The probability of a single qubit dephase error cannot exceed 1/2, which maximally mixes.",
    /* E_INVALID_TWO_QUBIT_DEPHASE_PROB */
    "The probability of a two-qubit qubit dephase error cannot exceed 3/4, which maximally mixes.",
    /* E_INVALID_ONE_QUBIT_DEPOL_PROB */
    "The probability of a single qubit depolarising error cannot exceed 3/4, which maximally mixes.",
    /* E_INVALID_TWO_QUBIT_DEPOL_PROB */
    "The probability of a two-qubit depolarising error cannot exceed 15/16, which maximally mixes.",
    /* E_INVALID_ONE_QUBIT_PAULI_PROBS */
    "The probability of any X, Y or Z error cannot exceed the probability of no error.",
    /* E_INVALID_CONTROLS_BIT_STATE */ "The state of the control qubits must be a bit sequence (0s and 1s).",
    /* E_INVALID_PAULI_CODE */
    "Invalid Pauli code. Codes must be 0 (or PAULI_I), 1 (PAULI_X), 2 (PAULI_Y) or 3 (PAULI_Z) to indicate the "
    "identity, X, Y and Z operators respectively.",
    /* E_INVALID_NUM_SUM_TERMS */ "Invalid number of terms in the Pauli sum. The number of terms must be >0.",
    /* E_CANNOT_FIT_MULTI_QUBIT_MATRIX */
    "The specified matrix targets too many qubits; the batches of amplitudes to modify cannot all fit in a "
    "single distributed node's memory allocation.",
    /* E_INVALID_UNITARY_SIZE */ "The matrix size does not match the number of target qubits.",
    /* E_COMPLEX_MATRIX_NOT_INIT */
    "The ComplexMatrixN was not successfully created (possibly insufficient memory available).",
    /* E_INVALID_NUM_ONE_QUBIT_KRAUS_OPS */ "At least 1 and at most 4 single qubit Kraus operators may be specified.",
    /* E_INVALID_NUM_TWO_QUBIT_KRAUS_OPS */ "At least 1 and at most 16 two-qubit Kraus operators may be specified.",
    /* E_INVALID_NUM_N_QUBIT_KRAUS_OPS */ "At least 1 and at most 4*N^2 of N-qubit Kraus operators may be specified.",
    /* E_INVALID_KRAUS_OPS */ "The specified Kraus map is not a completely positive, trace preserving map.",
    /* E_MISMATCHING_NUM_TARGS_KRAUS_SIZE */
    "Every Kraus operator must be of the same number of qubits as the number of targets.",
    /* E_NUM_AMPS_EXCEED_TYPE */
    "Too many qubits (max of log2(SIZE_MAX)). Cannot store the number of amplitudes per-node in the size_t type.",
    /* E_INVALID_PAULI_HAMIL_PARAMS */ "The number of qubits and terms in the PauliHamil must be strictly positive.",
    /* E_INVALID_PAULI_HAMIL_FILE_PARAMS */
    "The number of qubits and terms in the PauliHamil file (%s) must be strictly positive.",
    /* E_CANNOT_PARSE_PAULI_HAMIL_FILE_COEFF */
    "Failed to parse the next expected term coefficient in PauliHamil file (%s).",
    /* E_CANNOT_PARSE_PAULI_HAMIL_FILE_PAULI */ "Failed to parse the next expected Pauli code in PauliHamil file (%s).",
    /* E_INVALID_PAULI_HAMIL_FILE_PAULI_CODE */
    "The PauliHamil file (%s) contained an invalid pauli code (%d). Codes must be 0 (or PAULI_I), 1 (PAULI_X), 2 "
    "(PAULI_Y) or 3 (PAULI_Z) to indicate the identity, X, Y and Z operators respectively.",
    /* E_MISMATCHING_PAULI_HAMIL_QUREG_NUM_QUBITS */ "The PauliHamil must act on the same number of qubits as exist in the Qureg.",
    /* E_INVALID_TROTTER_ORDER */
    "The Trotterisation order must be 1, or an even number (for higher-order Suzuki symmetrized expansions).",
    /* E_INVALID_TROTTER_REPS */ "The number of Trotter repetitions must be >=1.",
    /* E_QUREG_NOT_ALLOCATED */ "Could not allocate memory for Qureg. Possibly insufficient memory.",
    /* E_QUREG_NOT_ALLOCATED_ON_GPU */ "Could not allocate memory for Qureg on GPU. Possibly insufficient memory.",
    /* E_MATRIX_TOO_LARGE_FOR_DEVICE */
    "The matrix targets more than 10 qubits, beyond what this MI355X backend applies as one dense block.",
    /* E_MISMATCHING_PRECISION */
    "Registers must share the same storage precision for this operation.",
    /* E_INVALID_PRECISION */ "Invalid storage precision. Must be QMX_PREC_FP64 (0) or QMX_PREC_FP32 (1).",
    /* E_INVALID_ELEM_INDEX */ "Invalid element index. Must be >=0 and <2^numQubits.",
    /* E_INVALID_NUM_ELEMS */ "Invalid number of elements. Must be >=0 and <=2^numQubits.",
    /* E_INVALID_OFFSET_NUM_ELEMS_DIAG */
    "More elements given than exist in the diagonal operator from the given starting index.",
    /* E_MISMATCHING_QUREG_DIAGONAL_OP_SIZE */
    "The qureg must represent an equal number of qubits as that in the applied diagonal operator.",
    /* E_DIAGONAL_OP_NOT_INITIALISED */
    "The diagonal operator has not been initialised through createDiagonalOperator().",
    /* E_PAULI_HAMIL_NOT_DIAGONAL */
    "The Pauli Hamiltonian contained operators other than PAULI_Z and PAULI_I, and hence cannot be expressed as a "
    "diagonal matrix.",
    /* E_MISMATCHING_PAULI_HAMIL_DIAGONAL_OP_SIZE */
    "The Pauli Hamiltonian and diagonal operator have different, incompatible dimensions.",
    /* E_DIAGONAL_OP_NOT_ALLOCATED */ "Could not allocate memory for DiagonalOp. Possibly insufficient memory.",
    /* E_DIAGONAL_OP_NOT_ALLOCATED_ON_GPU */
    "Could not allocate memory for DiagonalOp on GPU. Possibly insufficient memory.",
    /* E_INVALID_NUM_SUBREGISTERS */ "Invalid number of qubit subregisters, which must be >0 and <=100.",
    /* E_INVALID_NUM_PHASE_FUNC_TERMS */ "Invalid number of terms in the phase function specified. Must be >0.",
    /* E_INVALID_NUM_PHASE_FUNC_OVERRIDES */
    "Invalid number of phase function overrides specified. Must be >=0, and for single-variable phase functions, "
    "<=2^numQubits (the maximum unique binary values of the sub-register). Note that uniqueness of overriding "
    "indices is not checked.",
    /* E_INVALID_PHASE_FUNC_OVERRIDE_UNSIGNED_INDEX */
    "Invalid phase function override index, in the UNSIGNED encoding. Must be >=0, and <= the maximum index "
    "possible of the corresponding qubit subregister (2^numQubits-1).",
    /* E_INVALID_PHASE_FUNC_OVERRIDE_TWOS_COMPLEMENT_INDEX */
    "Invalid phase function override index, in the TWOS_COMPLEMENT encoding. Must be between (inclusive) "
    "-2^(N-1) and +2^(N-1)-1, where N is the number of qubits (including the sign qubit).",
    /* E_INVALID_PHASE_FUNC_NAME */
    "Invalid named phase function, which must be one of {NORM, SCALED_NORM, INVERSE_NORM, SCALED_INVERSE_NORM, "
    "PRODUCT, SCALED_PRODUCT, INVERSE_PRODUCT, SCALED_INVERSE_PRODUCT, DISTANCE, SCALED_DISTANCE, "
    "INVERSE_DISTANCE, SCALED_INVERSE_DISTANCE}.",
    /* E_INVALID_NUM_NAMED_PHASE_FUNC_PARAMS */
    "Invalid number of parameters passed for the given named phase function. {NORM, PRODUCT, DISTANCE} accept 0 "
    "parameters, {INVERSE_NORM, INVERSE_PRODUCT, INVERSE_DISTANCE} accept 1 parameter (the phase at the "
    "divergence), {SCALED_NORM, SCALED_INVERSE_NORM, SCALED_PRODUCT} accept 1 parameter (the scaling "
    "coefficient), {SCALED_INVERSE_PRODUCT, SCALED_DISTANCE, SCALED_INVERSE_DISTANCE} accept 2 parameters (the "
    "coefficient then divergence phase), SCALED_INVERSE_SHIFTED_NORM accepts 2 + (number of sub-registers) "
    "parameters (the coefficient, then the divergence phase, followed by the offset for each sub-register), "
    "SCALED_INVERSE_SHIFTED_DISTANCE accepts 2 + (number of sub-registers) / 2 parameters (the coefficient, then "
    "the divergence phase, followed by the offset for each pair of sub-registers).",
    /* E_INVALID_NUM_REGS_DISTANCE_PHASE_FUNC */
    "Phase functions DISTANCE, INVERSE_DISTANCE, SCALED_DISTANCE and SCALED_INVERSE_DISTANCE require a strictly "
    "even number of sub-registers.",
    /* E_INVALID_BIT_ENCODING */ "Invalid bit encoding. Must be one of {UNSIGNED, TWOS_COMPLEMENT}.",
    /* E_INVALID_NUM_QUBITS_TWOS_COMPLEMENT */
    "A sub-register contained too few qubits to employ TWOS_COMPLEMENT encoding. Must use >1 qubits (allocating "
    "one for the sign).",
    /* E_NEGATIVE_EXPONENT_WITHOUT_ZERO_OVERRIDE */
    "The phase function contained a negative exponent which would diverge at zero, but the zero index was not "
    "overriden.",
    /* E_FRACTIONAL_EXPONENT_WITHOUT_NEG_OVERRIDE */
    "The phase function contained a fractional exponent, which in TWOS_COMPLEMENT encoding, requires all negative "
    "indices are overriden. However, one or more negative indices were not overriden.",
    /* E_NEGATIVE_EXPONENT_MULTI_VAR */
    "The phase function contained an illegal negative exponent. One must instead call applyPhaseFuncOverrides() "
    "once for each register, so that the zero index of each register is overriden, independent of the indices of "
    "all other registers.",
    /* E_FRACTIONAL_EXPONENT_MULTI_VAR */
    "The phase function contained a fractional exponent, which is illegal in TWOS_COMPLEMENT encoding, since it "
    "cannot be (efficiently) checked that all negative indices were overriden. One must instead call "
    "applyPhaseFuncOverrides() once for each register, so that each register's negative indices can be "
    "overriden, independent of the indices of all other registers.",
    /* E_DISTRIB_QUREG_TOO_SMALL */
    "Too few qubits. The created qureg must have at least one amplitude per node used in distributed simulation.",
    /* E_DISTRIB_DIAG_OP_TOO_SMALL */
    "Too few qubits. The created DiagonalOp must contain at least one element per node used in distributed "
    "simulation.",
};
static_assert(sizeof(kMessages) / sizeof(kMessages[0]) == E_NUM_CODES, "message table out of sync");

std::mutex g_mu;
int g_mode = 0;
bool g_have = false;
std::string g_last;
char g_fmtBuf[1024];

}  // namespace

const char *message(ErrorCode code) { return kMessages[code]; }

void setMode(int mode) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_mode = mode;
}

int mode() { return g_mode; }

void record(const char *msg, const char *func) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_have = true;
    g_last = std::string(func ? func : "?") + ": " + (msg ? msg : "");
}

int lastError(char *buf, int len) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (buf && len > 0) {
        std::snprintf(buf, (size_t)len, "%s", g_last.c_str());
    }
    return g_have ? 1 : 0;
}

void clear() {
    std::lock_guard<std::mutex> lk(g_mu);
    g_have = false;
    g_last.clear();
}

[[noreturn]] void raise(const char *msg, const char *func) {
    invalidQuESTInputError(msg, func);
    // a replacement handler that returns: abandon the call, state untouched
    throw qmx::InputError();
}

void check(bool ok, ErrorCode code, const char *func) {
    if (!ok) raise(kMessages[code], func);
}

void checkFmt(bool ok, ErrorCode code, const char *func, const char *arg) {
    if (ok) return;
    std::snprintf(g_fmtBuf, sizeof g_fmtBuf, kMessages[code], arg);
    raise(g_fmtBuf, func);
}

void checkFmtCode(bool ok, ErrorCode code, const char *func, const char *arg, int code2) {
    if (ok) return;
    std::snprintf(g_fmtBuf, sizeof g_fmtBuf, kMessages[code], arg, code2);
    raise(g_fmtBuf, func);
}

// ---- predicates (QuEST_validation.c:255-342) ---------------------------------

static bool isComplexPairUnitary(Complex a, Complex b) {
    return std::fabs(-1 + a.real * a.real + a.imag * a.imag + b.real * b.real + b.imag * b.imag) < REAL_EPS;
}

// The sums run in qreal, as QuEST's own (QuEST_validation.c:270-343), so that a
// QuEST_PREC=1 build accepts and rejects exactly the matrices and maps QuEST's
// float build does: near REAL_EPS = 1e-5 the float rounding of the sums decides.
template <typename GetRe, typename GetIm>
static bool isUnitary(int dim, GetRe re, GetIm im) {
    for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c) {
            qreal er = 0, ei = 0;
            for (int i = 0; i < dim; ++i) {
                er += re(r, i) * re(c, i) + im(r, i) * im(c, i);
                ei += im(r, i) * re(c, i) - re(r, i) * im(c, i);
            }
            if (std::fabs(ei) > REAL_EPS || (r == c && std::fabs(er - 1) > REAL_EPS) ||
                (r != c && std::fabs(er) > REAL_EPS))
                return false;
        }
    return true;
}

bool isUnitary2(const ComplexMatrix2 &u) {
    return isUnitary(2, [&](int r, int c) { return u.real[r][c]; }, [&](int r, int c) { return u.imag[r][c]; });
}
bool isUnitary4(const ComplexMatrix4 &u) {
    return isUnitary(4, [&](int r, int c) { return u.real[r][c]; }, [&](int r, int c) { return u.imag[r][c]; });
}
bool isUnitaryN(const ComplexMatrixN &u) {
    return isUnitary(1 << u.numQubits, [&](int r, int c) { return u.real[r][c]; },
                     [&](int r, int c) { return u.imag[r][c]; });
}

template <typename GetRe, typename GetIm>
static bool isCPTP(int numOps, int dim, GetRe re, GetIm im) {
    for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c) {
            qreal er = 0, ei = 0;
            for (int n = 0; n < numOps; ++n)
                for (int k = 0; k < dim; ++k) {
                    er += re(n, k, r) * re(n, k, c) + im(n, k, r) * im(n, k, c);
                    ei += re(n, k, r) * im(n, k, c) - im(n, k, r) * re(n, k, c);
                }
            const qreal dist = std::fabs(ei) + std::fabs(er - ((r == c) ? 1 : 0));
            if (dist > REAL_EPS) return false;
        }
    return true;
}

static bool uniqueQubits(const int *q, int n) {
    long long mask = 0;
    for (int i = 0; i < n; ++i) {
        long long b = 1LL << q[i];
        if (mask & b) return false;
        mask |= b;
    }
    return true;
}

// ---- checks -------------------------------------------------------------------

// numRanks: QuESTEnv.numRanks; a register needs at least one amplitude per rank
// (QuEST_validation.c:383-393; compared in 64 bits here, so that registers of
// 31 qubits and more pass)
void numQubitsInQureg(int numQubits, const char *f, int numRanks) {
    check(numQubits > 0, E_INVALID_NUM_CREATE_QUBITS, f);
    unsigned maxQ = 0;
    for (size_t v = SIZE_MAX; v >>= 1;) ++maxQ;
    check(numQubits <= (int)maxQ, E_NUM_AMPS_EXCEED_TYPE, f);
    check(numQubits >= 62 || (1LL << numQubits) >= (long long)numRanks, E_DISTRIB_QUREG_TOO_SMALL, f);
}
void numQubitsInMatrix(int numQubits, const char *f) { check(numQubits > 0, E_INVALID_NUM_QUBITS, f); }
void stateIndex(const Qureg &q, long long ind, const char *f) {
    check(ind >= 0 && ind < (1LL << q.numQubitsRepresented), E_INVALID_STATE_INDEX, f);
}
void ampIndex(const Qureg &q, long long ind, const char *f) {
    check(ind >= 0 && ind < (1LL << q.numQubitsRepresented), E_INVALID_AMP_INDEX, f);
}
void numAmps(const Qureg &q, long long start, long long n, const char *f) {
    ampIndex(q, start, f);
    check(n >= 0 && n <= q.numAmpsTotal, E_INVALID_NUM_AMPS, f);
    check(n + start <= q.numAmpsTotal, E_INVALID_OFFSET_NUM_AMPS_QUREG, f);
}
void numDensityAmps(const Qureg &q, long long row, long long col, long long n, const char *f) {
    ampIndex(q, row, f);
    ampIndex(q, col, f);
    check(n >= 0 && n <= q.numAmpsTotal, E_INVALID_NUM_AMPS, f);
    long long start = row + col * (1LL << q.numQubitsRepresented);
    check(n + start <= q.numAmpsTotal, E_INVALID_OFFSET_NUM_AMPS_QUREG, f);
}
void target(const Qureg &q, int t, const char *f) {
    check(t >= 0 && t < q.numQubitsRepresented, E_INVALID_TARGET_QUBIT, f);
}
void control(const Qureg &q, int c, const char *f) {
    check(c >= 0 && c < q.numQubitsRepresented, E_INVALID_CONTROL_QUBIT, f);
}
void controlTarget(const Qureg &q, int c, int t, const char *f) {
    target(q, t, f);
    control(q, c, f);
    check(c != t, E_TARGET_IS_CONTROL, f);
}
void uniqueTargets(const Qureg &q, int a, int b, const char *f) {
    target(q, a, f);
    target(q, b, f);
    check(a != b, E_TARGETS_NOT_UNIQUE, f);
}
void multiTargets(const Qureg &q, const int *t, int n, const char *f) {
    check(n > 0 && n <= q.numQubitsRepresented, E_INVALID_NUM_TARGETS, f);
    for (int i = 0; i < n; ++i) target(q, t[i], f);
    check(uniqueQubits(t, n), E_TARGETS_NOT_UNIQUE, f);
}
void multiControls(const Qureg &q, const int *c, int n, const char *f) {
    check(n > 0 && n < q.numQubitsRepresented, E_INVALID_NUM_CONTROLS, f);
    for (int i = 0; i < n; ++i) control(q, c[i], f);
    check(uniqueQubits(c, n), E_CONTROLS_NOT_UNIQUE, f);
}
void multiQubits(const Qureg &q, const int *qs, int n, const char *f) {
    check(n > 0 && n <= q.numQubitsRepresented, E_INVALID_NUM_QUBITS, f);
    for (int i = 0; i < n; ++i) check(qs[i] >= 0 && qs[i] < q.numQubitsRepresented, E_INVALID_QUBIT_INDEX, f);
    check(uniqueQubits(qs, n), E_QUBITS_NOT_UNIQUE, f);
}
void multiControlsTarget(const Qureg &q, const int *c, int n, int t, const char *f) {
    target(q, t, f);
    multiControls(q, c, n, f);
    for (int i = 0; i < n; ++i) check(c[i] != t, E_TARGET_IN_CONTROLS, f);
}
void multiControlsMultiTargets(const Qureg &q, const int *c, int nc, const int *t, int nt, const char *f) {
    multiControls(q, c, nc, f);
    multiTargets(q, t, nt, f);
    long long cm = 0, tm = 0;
    for (int i = 0; i < nc; ++i) cm |= 1LL << c[i];
    for (int i = 0; i < nt; ++i) tm |= 1LL << t[i];
    check(!(int)(cm & tm), E_CONTROL_TARGET_COLLISION, f);
}
void controlState(const int *s, int n, const char *f) {
    for (int i = 0; i < n; ++i) check(s[i] == 0 || s[i] == 1, E_INVALID_CONTROLS_BIT_STATE, f);
}
void matrixFitsInNode(const Qureg &q, int numTargs, const char *f) {
    check(q.numAmpsPerChunk >= (1LL << numTargs), E_CANNOT_FIT_MULTI_QUBIT_MATRIX, f);
    check(numTargs <= 10, E_MATRIX_TOO_LARGE_FOR_DEVICE, f);
}
void oneQubitUnitary(const ComplexMatrix2 &u, const char *f) { check(isUnitary2(u), E_NON_UNITARY_MATRIX, f); }
void twoQubitUnitary(const Qureg &q, const ComplexMatrix4 &u, const char *f) {
    matrixFitsInNode(q, 2, f);
    check(isUnitary4(u), E_NON_UNITARY_MATRIX, f);
}
void matrixInit(const ComplexMatrixN &m, const char *f) {
    check(m.real != nullptr && m.imag != nullptr, E_COMPLEX_MATRIX_NOT_INIT, f);
}
void multiQubitMatrix(const Qureg &q, const ComplexMatrixN &u, int numTargs, const char *f) {
    matrixInit(u, f);
    matrixFitsInNode(q, numTargs, f);
    check(numTargs == u.numQubits, E_INVALID_UNITARY_SIZE, f);
}
void multiQubitUnitary(const Qureg &q, const ComplexMatrixN &u, int numTargs, const char *f) {
    multiQubitMatrix(q, u, numTargs, f);
    check(isUnitaryN(u), E_NON_UNITARY_MATRIX, f);
}
void unitaryComplexPair(Complex a, Complex b, const char *f) {
    check(isComplexPairUnitary(a, b), E_NON_UNITARY_COMPLEX_PAIR, f);
}
void vector(Vector v, const char *f) {
    check(std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z) > REAL_EPS, E_ZERO_VECTOR, f);
}
void stateVecQureg(const Qureg &q, const char *f) { check(!q.isDensityMatrix, E_DEFINED_ONLY_FOR_STATEVECS, f); }
void densityMatrQureg(const Qureg &q, const char *f) { check(q.isDensityMatrix, E_DEFINED_ONLY_FOR_DENSMATRS, f); }
void outcome(int o, const char *f) { check(o == 0 || o == 1, E_INVALID_QUBIT_OUTCOME, f); }
void measurementProb(double p, const char *f) { check(p > REAL_EPS, E_COLLAPSE_STATE_ZERO_PROB, f); }
void matchingQuregDims(const Qureg &a, const Qureg &b, const char *f) {
    check(a.numQubitsRepresented == b.numQubitsRepresented, E_MISMATCHING_QUREG_DIMENSIONS, f);
}
void matchingQuregTypes(const Qureg &a, const Qureg &b, const char *f) {
    check(a.isDensityMatrix == b.isDensityMatrix, E_MISMATCHING_QUREG_TYPES, f);
}
void secondQuregStateVec(const Qureg &b, const char *f) { check(!b.isDensityMatrix, E_SECOND_ARG_MUST_BE_STATEVEC, f); }
void fileOpened(bool opened, const char *fn, const char *f) { checkFmt(opened, E_CANNOT_OPEN_FILE, f, fn); }
void prob(double p, const char *f) { check(p >= 0 && p <= 1, E_INVALID_PROB, f); }
void oneQubitDephaseProb(double p, const char *f) {
    prob(p, f);
    check(p <= 1 / 2.0, E_INVALID_ONE_QUBIT_DEPHASE_PROB, f);
}
void twoQubitDephaseProb(double p, const char *f) {
    prob(p, f);
    check(p <= 3 / 4.0, E_INVALID_TWO_QUBIT_DEPHASE_PROB, f);
}
void oneQubitDepolProb(double p, const char *f) {
    prob(p, f);
    check(p <= 3 / 4.0, E_INVALID_ONE_QUBIT_DEPOL_PROB, f);
}
void oneQubitDampingProb(double p, const char *f) {
    prob(p, f);
    check(p <= 1.0, E_INVALID_ONE_QUBIT_DEPOL_PROB, f);
}
void twoQubitDepolProb(double p, const char *f) {
    prob(p, f);
    check(p <= 15 / 16.0, E_INVALID_TWO_QUBIT_DEPOL_PROB, f);
}
void oneQubitPauliProbs(double px, double py, double pz, const char *f) {
    prob(px, f);
    prob(py, f);
    prob(pz, f);
    double none = 1 - px - py - pz;
    check(px <= none, E_INVALID_ONE_QUBIT_PAULI_PROBS, f);
    check(py <= none, E_INVALID_ONE_QUBIT_PAULI_PROBS, f);
    check(pz <= none, E_INVALID_ONE_QUBIT_PAULI_PROBS, f);
}
bool validPauliCode(int code) { return code == PAULI_I || code == PAULI_X || code == PAULI_Y || code == PAULI_Z; }
// A C caller's code array may hold any int. Each element is read as its int
// storage: loading an out-of-range value through the C++ enum type is undefined,
// and would let the compiler assume the very range the checks test.
static_assert(sizeof(enum pauliOpType) == sizeof(int), "pauliOpType is stored as an int");
int pauliCodeAt(const enum pauliOpType *codes, long long i) {
    int code;
    std::memcpy(&code, codes + i, sizeof code);
    return code;
}
void pauliCodes(const enum pauliOpType *codes, int n, const char *f) {
    for (int i = 0; i < n; ++i) check(validPauliCode(pauliCodeAt(codes, i)), E_INVALID_PAULI_CODE, f);
}
void numPauliSumTerms(int n, const char *f) { check(n > 0, E_INVALID_NUM_SUM_TERMS, f); }

void oneQubitKrausMapDims(const Qureg &q, int numOps, const char *f) {
    check(numOps > 0 && numOps <= 4, E_INVALID_NUM_ONE_QUBIT_KRAUS_OPS, f);
    matrixFitsInNode(q, 2, f);
}
void oneQubitKrausMap(const Qureg &q, const ComplexMatrix2 *ops, int numOps, const char *f) {
    oneQubitKrausMapDims(q, numOps, f);
    check(isCPTP(numOps, 2, [&](int n, int r, int c) { return ops[n].real[r][c]; },
                 [&](int n, int r, int c) { return ops[n].imag[r][c]; }),
          E_INVALID_KRAUS_OPS, f);
}
void twoQubitKrausMapDims(const Qureg &q, int numOps, const char *f) {
    check(numOps > 0 && numOps <= 16, E_INVALID_NUM_TWO_QUBIT_KRAUS_OPS, f);
    matrixFitsInNode(q, 4, f);
}
void twoQubitKrausMap(const Qureg &q, const ComplexMatrix4 *ops, int numOps, const char *f) {
    twoQubitKrausMapDims(q, numOps, f);
    check(isCPTP(numOps, 4, [&](int n, int r, int c) { return ops[n].real[r][c]; },
                 [&](int n, int r, int c) { return ops[n].imag[r][c]; }),
          E_INVALID_KRAUS_OPS, f);
}
void multiQubitKrausMapDims(const Qureg &q, int numTargs, const ComplexMatrixN *ops, int numOps, const char *f) {
    int sup = 2 * numTargs;
    check(numOps > 0 && numOps <= sup * sup, E_INVALID_NUM_N_QUBIT_KRAUS_OPS, f);
    for (int n = 0; n < numOps; ++n) {
        matrixInit(ops[n], f);
        check(ops[n].numQubits == numTargs, E_MISMATCHING_NUM_TARGS_KRAUS_SIZE, f);
    }
    matrixFitsInNode(q, sup, f);
}
void multiQubitKrausMap(const Qureg &q, int numTargs, const ComplexMatrixN *ops, int numOps, const char *f) {
    multiQubitKrausMapDims(q, numTargs, ops, numOps, f);
    int dim = 1 << ops[0].numQubits;
    check(isCPTP(numOps, dim, [&](int n, int r, int c) { return ops[n].real[r][c]; },
                 [&](int n, int r, int c) { return ops[n].imag[r][c]; }),
          E_INVALID_KRAUS_OPS, f);
}
void hamilParams(int numQubits, int numTerms, const char *f) {
    check(numQubits > 0 && numTerms > 0, E_INVALID_PAULI_HAMIL_PARAMS, f);
}
void pauliHamil(const PauliHamil &h, const char *f) {
    hamilParams(h.numQubits, h.numSumTerms, f);
    pauliCodes(h.pauliCodes, h.numSumTerms * h.numQubits, f);
}
void matchingQuregPauliHamilDims(const Qureg &q, const PauliHamil &h, const char *f) {
    check(h.numQubits == q.numQubitsRepresented, E_MISMATCHING_PAULI_HAMIL_QUREG_NUM_QUBITS, f);
}
void trotterParams(int order, int reps, const char *f) {
    bool even = (order % 2) == 0;
    check(order > 0 && (even || order == 1), E_INVALID_TROTTER_ORDER, f);
    check(reps > 0, E_INVALID_TROTTER_REPS, f);
}
void matchingPrecision(const Qureg &a, const Qureg &b, const char *f) {
    check(qmx::dev(a)->prec == qmx::dev(b)->prec, E_MISMATCHING_PRECISION, f);
}
void precision(int p, const char *f) { check(p == 0 || p == 1, E_INVALID_PRECISION, f); }


// ---- diagonal operators ----------------------------------------------------------
void numQubitsInDiagOp(int numQubits, const char *f, int numRanks) {
    check(numQubits > 0, E_INVALID_NUM_CREATE_QUBITS, f);
    check(numQubits <= 63, E_NUM_AMPS_EXCEED_TYPE, f);
    check(numQubits >= 62 || (1LL << numQubits) >= (long long)numRanks, E_DISTRIB_DIAG_OP_TOO_SMALL, f);
}
void numElems(const DiagonalOp &op, long long start, long long n, const char *f) {
    const long long mx = 1LL << op.numQubits;
    check(start >= 0 && start < mx, E_INVALID_ELEM_INDEX, f);
    check(n >= 0 && n <= mx, E_INVALID_NUM_ELEMS, f);
    check(n + start <= mx, E_INVALID_OFFSET_NUM_ELEMS_DIAG, f);
}
void diagOpInit(const DiagonalOp &op, const char *f) {
    check(op.real != nullptr && op.imag != nullptr, E_DIAGONAL_OP_NOT_INITIALISED, f);
}
void diagonalOp(const Qureg &q, const DiagonalOp &op, const char *f) {
    diagOpInit(op, f);
    check(q.numQubitsRepresented == op.numQubits, E_MISMATCHING_QUREG_DIAGONAL_OP_SIZE, f);
}
bool diagonalHamil(const PauliHamil &h) {
    for (long long p = 0; p < (long long)h.numSumTerms * h.numQubits; ++p)
        if (pauliCodeAt(h.pauliCodes, p) != PAULI_I && pauliCodeAt(h.pauliCodes, p) != PAULI_Z) return false;
    return true;
}
void diagPauliHamil(const DiagonalOp &op, const PauliHamil &h, const char *f) {
    check(op.numQubits == h.numQubits, E_MISMATCHING_PAULI_HAMIL_DIAGONAL_OP_SIZE, f);
    check(diagonalHamil(h), E_PAULI_HAMIL_NOT_DIAGONAL, f);
}

// ---- phase functions -----------------------------------------------------------------
namespace {
constexpr int kMaxRegs = 100;  // MAX_NUM_REGS_APPLY_ARBITRARY_PHASE (QuEST_precision.h)
bool isFraction(double e) { return std::floor(e) != e; }
}  // namespace

void phaseMultiQubits(const Qureg &q, const int *qs, int n, const char *f) {
    check(n > 0 && n <= q.numQubitsRepresented, E_INVALID_NUM_QUBITS, f);
    for (int i = 0; i < n; ++i) check(qs[i] >= 0 && qs[i] < q.numQubitsRepresented, E_INVALID_QUBIT_INDEX, f);
    check(uniqueQubits(qs, n), E_QUBITS_NOT_UNIQUE, f);
}
void qubitSubregs(const Qureg &q, const int *qubits, const int *numQubitsPerReg, int numRegs, const char *f) {
    check(numRegs > 0 && numRegs <= kMaxRegs, E_INVALID_NUM_SUBREGISTERS, f);
    int i = 0;
    for (int r = 0; r < numRegs; ++r) {
        check(numQubitsPerReg[r] > 0 && numQubitsPerReg[r] <= q.numQubitsRepresented, E_INVALID_NUM_QUBITS, f);
        for (int k = 0; k < numQubitsPerReg[r]; ++k, ++i)
            check(qubits[i] >= 0 && qubits[i] < q.numQubitsRepresented, E_INVALID_QUBIT_INDEX, f);
    }
    check(uniqueQubits(qubits, i), E_QUBITS_NOT_UNIQUE, f);
}
void bitEncoding(int numQubits, int encoding, const char *f) {
    check(encoding == UNSIGNED || encoding == TWOS_COMPLEMENT, E_INVALID_BIT_ENCODING, f);
    if (encoding == TWOS_COMPLEMENT) check(numQubits > 1, E_INVALID_NUM_QUBITS_TWOS_COMPLEMENT, f);
}
void multiRegBitEncoding(const int *numQubitsPerReg, int numRegs, int encoding, const char *f) {
    check(encoding == UNSIGNED || encoding == TWOS_COMPLEMENT, E_INVALID_BIT_ENCODING, f);
    if (encoding == TWOS_COMPLEMENT)
        for (int r = 0; r < numRegs; ++r) check(numQubitsPerReg[r] > 1, E_INVALID_NUM_QUBITS_TWOS_COMPLEMENT, f);
}
void phaseFuncTerms(int numQubits, int encoding, const double *exponents, int numTerms,
                    const long long *overrideInds, int numOverrides, const char *f) {
    check(numTerms > 0, E_INVALID_NUM_PHASE_FUNC_TERMS, f);
    bool frac = false, neg = false;
    for (int t = 0; t < numTerms; ++t) {
        frac = frac || isFraction(exponents[t]);
        neg = neg || exponents[t] < 0;
    }
    // a negative power diverges at index 0, which must then be overridden
    if (neg) {
        bool zero = false;
        for (int v = 0; v < numOverrides && !zero; ++v) zero = overrideInds[v] == 0;
        check(zero, E_NEGATIVE_EXPONENT_WITHOUT_ZERO_OVERRIDE, f);
    }
    // a fractional power of a negative index is undefined: every negative index
    // must be overridden (checked exactly below 16 qubits, trusted above)
    if (frac && encoding == TWOS_COMPLEMENT) {
        const long long numNeg = 1LL << (numQubits - 1);
        check(numOverrides >= numNeg, E_FRACTIONAL_EXPONENT_WITHOUT_NEG_OVERRIDE, f);
        if (numQubits < 16) {
            std::vector<char> seen((size_t)numNeg, 0);
            for (int v = 0; v < numOverrides; ++v)
                if (overrideInds[v] < 0 && -1 - overrideInds[v] < numNeg) seen[(size_t)(-1 - overrideInds[v])] = 1;
            bool all = true;
            for (long long i = 0; i < numNeg && all; ++i) all = seen[(size_t)i] != 0;
            check(all, E_FRACTIONAL_EXPONENT_WITHOUT_NEG_OVERRIDE, f);
        }
    }
}
void multiVarPhaseFuncTerms(int numRegs, int encoding, const double *exponents, const int *numTermsPerReg,
                            const char *f) {
    check(numRegs > 0 && numRegs <= kMaxRegs, E_INVALID_NUM_SUBREGISTERS, f);
    int total = 0;
    for (int r = 0; r < numRegs; ++r) {
        check(numTermsPerReg[r] > 0, E_INVALID_NUM_PHASE_FUNC_TERMS, f);
        total += numTermsPerReg[r];
    }
    bool frac = false, neg = false;
    for (int t = 0; t < total; ++t) {
        frac = frac || isFraction(exponents[t]);
        neg = neg || exponents[t] < 0;
    }
    check(!neg, E_NEGATIVE_EXPONENT_MULTI_VAR, f);
    if (encoding == TWOS_COMPLEMENT) check(!frac, E_FRACTIONAL_EXPONENT_MULTI_VAR, f);
}
void phaseFuncOverrides(int numQubits, int encoding, const long long *overrideInds, int numOverrides,
                        const char *f) {
    check(numOverrides >= 0, E_INVALID_NUM_PHASE_FUNC_OVERRIDES, f);
    check(numOverrides <= (1 << numQubits), E_INVALID_NUM_PHASE_FUNC_OVERRIDES, f);
    if (encoding == UNSIGNED) {
        const long long mx = (1LL << numQubits) - 1;
        for (int v = 0; v < numOverrides; ++v)
            check(overrideInds[v] >= 0 && overrideInds[v] <= mx, E_INVALID_PHASE_FUNC_OVERRIDE_UNSIGNED_INDEX, f);
    } else if (encoding == TWOS_COMPLEMENT) {
        const long long mn = -(1LL << (numQubits - 1)), mx = (1LL << (numQubits - 1)) - 1;
        for (int v = 0; v < numOverrides; ++v)
            check(overrideInds[v] >= mn && overrideInds[v] <= mx, E_INVALID_PHASE_FUNC_OVERRIDE_TWOS_COMPLEMENT_INDEX,
                  f);
    }
}
void multiVarPhaseFuncOverrides(const int *numQubitsPerReg, int numRegs, int encoding,
                                const long long *overrideInds, int numOverrides, const char *f) {
    check(numOverrides >= 0, E_INVALID_NUM_PHASE_FUNC_OVERRIDES, f);
    int i = 0;
    for (int v = 0; v < numOverrides; ++v)
        for (int r = 0; r < numRegs; ++r, ++i) {
            if (encoding == UNSIGNED) {
                const long long mx = (1LL << numQubitsPerReg[r]) - 1;
                check(overrideInds[i] >= 0 && overrideInds[i] <= mx, E_INVALID_PHASE_FUNC_OVERRIDE_UNSIGNED_INDEX, f);
            } else if (encoding == TWOS_COMPLEMENT) {
                const long long mn = -(1LL << (numQubitsPerReg[r] - 1)), mx = (1LL << (numQubitsPerReg[r] - 1)) - 1;
                check(overrideInds[i] >= mn && overrideInds[i] <= mx,
                      E_INVALID_PHASE_FUNC_OVERRIDE_TWOS_COMPLEMENT_INDEX, f);
            }
        }
}
void phaseFuncName(int c, int numRegs, int numParams, const char *f) {
    check(c >= NORM && c <= SCALED_INVERSE_SHIFTED_DISTANCE, E_INVALID_PHASE_FUNC_NAME, f);
    int want = 0;
    switch (c) {
    case NORM: case PRODUCT: case DISTANCE: want = 0; break;
    case INVERSE_NORM: case INVERSE_PRODUCT: case INVERSE_DISTANCE:
    case SCALED_NORM: case SCALED_PRODUCT: case SCALED_DISTANCE: want = 1; break;
    case SCALED_INVERSE_NORM: case SCALED_INVERSE_PRODUCT: case SCALED_INVERSE_DISTANCE: want = 2; break;
    case SCALED_INVERSE_SHIFTED_NORM: want = 2 + numRegs; break;
    case SCALED_INVERSE_SHIFTED_DISTANCE: want = 2 + numRegs / 2; break;
    }
    check(numParams == want, E_INVALID_NUM_NAMED_PHASE_FUNC_PARAMS, f);
    if (c == DISTANCE || c == INVERSE_DISTANCE || c == SCALED_DISTANCE || c == SCALED_INVERSE_DISTANCE ||
        c == SCALED_INVERSE_SHIFTED_DISTANCE)
        check(numRegs % 2 == 0, E_INVALID_NUM_REGS_DISTANCE_PHASE_FUNC, f);
}

}  // namespace qv

// Weak default handler (QuEST_validation.c:216-245): print and exit, unless the
// host asked for recording (qmxSetErrorMode), in which case record and return;
// qv::raise then abandons the call.
extern "C" __attribute__((weak)) void invalidQuESTInputError(const char *errMsg, const char *errFunc) {
    if (qv::mode() == 1) {
        qv::record(errMsg, errFunc);
        return;
    }
    std::printf("!!!\n");
    std::printf("QuEST Error in function %s: %s\n", errFunc, errMsg);
    std::printf("!!!\n");
    std::printf("exiting..\n");
    std::fflush(stdout);
    std::exit(EXIT_FAILURE);
}
