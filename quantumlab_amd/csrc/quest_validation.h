// Validation layer of the QuEST C-API front-end (see quest_validation.cpp).
#pragma once

#include "QuEST.h"
#include "qmx_internal.h"

namespace qmx {
// thrown after a (replaced) invalidQuESTInputError returns; caught at every API entry
struct InputError {};
}  // namespace qmx

namespace qv {

enum ErrorCode {
    E_SUCCESS = 0,
    E_INVALID_NUM_CREATE_QUBITS,
    E_INVALID_QUBIT_INDEX,
    E_INVALID_TARGET_QUBIT,
    E_INVALID_CONTROL_QUBIT,
    E_INVALID_STATE_INDEX,
    E_INVALID_AMP_INDEX,
    E_INVALID_NUM_AMPS,
    E_INVALID_OFFSET_NUM_AMPS_QUREG,
    E_TARGET_IS_CONTROL,
    E_TARGET_IN_CONTROLS,
    E_CONTROL_TARGET_COLLISION,
    E_QUBITS_NOT_UNIQUE,
    E_TARGETS_NOT_UNIQUE,
    E_CONTROLS_NOT_UNIQUE,
    E_INVALID_NUM_QUBITS,
    E_INVALID_NUM_TARGETS,
    E_INVALID_NUM_CONTROLS,
    E_NON_UNITARY_MATRIX,
    E_NON_UNITARY_COMPLEX_PAIR,
    E_ZERO_VECTOR,
    E_COLLAPSE_STATE_ZERO_PROB,
    E_INVALID_QUBIT_OUTCOME,
    E_CANNOT_OPEN_FILE,
    E_SECOND_ARG_MUST_BE_STATEVEC,
    E_MISMATCHING_QUREG_DIMENSIONS,
    E_MISMATCHING_QUREG_TYPES,
    E_DEFINED_ONLY_FOR_STATEVECS,
    E_DEFINED_ONLY_FOR_DENSMATRS,
    E_INVALID_PROB,
    E_INVALID_ONE_QUBIT_DEPHASE_PROB,
    E_INVALID_TWO_QUBIT_DEPHASE_PROB,
    E_INVALID_ONE_QUBIT_DEPOL_PROB,
    E_INVALID_TWO_QUBIT_DEPOL_PROB,
    E_INVALID_ONE_QUBIT_PAULI_PROBS,
    E_INVALID_CONTROLS_BIT_STATE,
    E_INVALID_PAULI_CODE,
    E_INVALID_NUM_SUM_TERMS,
    E_CANNOT_FIT_MULTI_QUBIT_MATRIX,
    E_INVALID_UNITARY_SIZE,
    E_COMPLEX_MATRIX_NOT_INIT,
    E_INVALID_NUM_ONE_QUBIT_KRAUS_OPS,
    E_INVALID_NUM_TWO_QUBIT_KRAUS_OPS,
    E_INVALID_NUM_N_QUBIT_KRAUS_OPS,
    E_INVALID_KRAUS_OPS,
    E_MISMATCHING_NUM_TARGS_KRAUS_SIZE,
    E_NUM_AMPS_EXCEED_TYPE,
    E_INVALID_PAULI_HAMIL_PARAMS,
    E_INVALID_PAULI_HAMIL_FILE_PARAMS,
    E_CANNOT_PARSE_PAULI_HAMIL_FILE_COEFF,
    E_CANNOT_PARSE_PAULI_HAMIL_FILE_PAULI,
    E_INVALID_PAULI_HAMIL_FILE_PAULI_CODE,
    E_MISMATCHING_PAULI_HAMIL_QUREG_NUM_QUBITS,
    E_INVALID_TROTTER_ORDER,
    E_INVALID_TROTTER_REPS,
    E_QUREG_NOT_ALLOCATED,
    E_QUREG_NOT_ALLOCATED_ON_GPU,
    E_MATRIX_TOO_LARGE_FOR_DEVICE,
    E_MISMATCHING_PRECISION,
    E_INVALID_PRECISION,
    E_INVALID_ELEM_INDEX,
    E_INVALID_NUM_ELEMS,
    E_INVALID_OFFSET_NUM_ELEMS_DIAG,
    E_MISMATCHING_QUREG_DIAGONAL_OP_SIZE,
    E_DIAGONAL_OP_NOT_INITIALISED,
    E_PAULI_HAMIL_NOT_DIAGONAL,
    E_MISMATCHING_PAULI_HAMIL_DIAGONAL_OP_SIZE,
    E_DIAGONAL_OP_NOT_ALLOCATED,
    E_DIAGONAL_OP_NOT_ALLOCATED_ON_GPU,
    E_INVALID_NUM_SUBREGISTERS,
    E_INVALID_NUM_PHASE_FUNC_TERMS,
    E_INVALID_NUM_PHASE_FUNC_OVERRIDES,
    E_INVALID_PHASE_FUNC_OVERRIDE_UNSIGNED_INDEX,
    E_INVALID_PHASE_FUNC_OVERRIDE_TWOS_COMPLEMENT_INDEX,
    E_INVALID_PHASE_FUNC_NAME,
    E_INVALID_NUM_NAMED_PHASE_FUNC_PARAMS,
    E_INVALID_NUM_REGS_DISTANCE_PHASE_FUNC,
    E_INVALID_BIT_ENCODING,
    E_INVALID_NUM_QUBITS_TWOS_COMPLEMENT,
    E_NEGATIVE_EXPONENT_WITHOUT_ZERO_OVERRIDE,
    E_FRACTIONAL_EXPONENT_WITHOUT_NEG_OVERRIDE,
    E_NEGATIVE_EXPONENT_MULTI_VAR,
    E_FRACTIONAL_EXPONENT_MULTI_VAR,
    E_DISTRIB_QUREG_TOO_SMALL,
    E_DISTRIB_DIAG_OP_TOO_SMALL,
    E_NUM_CODES
};

const char *message(ErrorCode code);
void setMode(int mode);
int mode();
void record(const char *msg, const char *func);
int lastError(char *buf, int len);
void clear();
[[noreturn]] void raise(const char *msg, const char *func);
void check(bool ok, ErrorCode code, const char *func);
void checkFmt(bool ok, ErrorCode code, const char *func, const char *arg);
void checkFmtCode(bool ok, ErrorCode code, const char *func, const char *arg, int code2);
bool validPauliCode(int code);

void numQubitsInQureg(int numQubits, const char *f, int numRanks = 1);
void numQubitsInMatrix(int numQubits, const char *f);
void stateIndex(const Qureg &q, long long ind, const char *f);
void ampIndex(const Qureg &q, long long ind, const char *f);
void numAmps(const Qureg &q, long long start, long long n, const char *f);
void numDensityAmps(const Qureg &q, long long row, long long col, long long n, const char *f);
void target(const Qureg &q, int t, const char *f);
void control(const Qureg &q, int c, const char *f);
void controlTarget(const Qureg &q, int c, int t, const char *f);
void uniqueTargets(const Qureg &q, int a, int b, const char *f);
void multiTargets(const Qureg &q, const int *t, int n, const char *f);
void multiControls(const Qureg &q, const int *c, int n, const char *f);
void multiQubits(const Qureg &q, const int *qs, int n, const char *f);
void multiControlsTarget(const Qureg &q, const int *c, int n, int t, const char *f);
void multiControlsMultiTargets(const Qureg &q, const int *c, int nc, const int *t, int nt, const char *f);
void controlState(const int *s, int n, const char *f);
void matrixFitsInNode(const Qureg &q, int numTargs, const char *f);
void oneQubitUnitary(const ComplexMatrix2 &u, const char *f);
void twoQubitUnitary(const Qureg &q, const ComplexMatrix4 &u, const char *f);
void matrixInit(const ComplexMatrixN &m, const char *f);
void multiQubitMatrix(const Qureg &q, const ComplexMatrixN &u, int numTargs, const char *f);
void multiQubitUnitary(const Qureg &q, const ComplexMatrixN &u, int numTargs, const char *f);
void unitaryComplexPair(Complex a, Complex b, const char *f);
void vector(Vector v, const char *f);
void stateVecQureg(const Qureg &q, const char *f);
void densityMatrQureg(const Qureg &q, const char *f);
void outcome(int o, const char *f);
void measurementProb(double p, const char *f);
void matchingQuregDims(const Qureg &a, const Qureg &b, const char *f);
void matchingQuregTypes(const Qureg &a, const Qureg &b, const char *f);
void secondQuregStateVec(const Qureg &b, const char *f);
void fileOpened(bool opened, const char *fn, const char *f);
void prob(double p, const char *f);
void oneQubitDephaseProb(double p, const char *f);
void twoQubitDephaseProb(double p, const char *f);
void oneQubitDepolProb(double p, const char *f);
void oneQubitDampingProb(double p, const char *f);
void twoQubitDepolProb(double p, const char *f);
void oneQubitPauliProbs(double px, double py, double pz, const char *f);
// element i of a caller's Pauli-code array, read as its int storage
int pauliCodeAt(const enum pauliOpType *codes, long long i);
void pauliCodes(const enum pauliOpType *codes, int n, const char *f);
void numPauliSumTerms(int n, const char *f);
void oneQubitKrausMapDims(const Qureg &q, int numOps, const char *f);
void oneQubitKrausMap(const Qureg &q, const ComplexMatrix2 *ops, int numOps, const char *f);
void twoQubitKrausMapDims(const Qureg &q, int numOps, const char *f);
void twoQubitKrausMap(const Qureg &q, const ComplexMatrix4 *ops, int numOps, const char *f);
void multiQubitKrausMapDims(const Qureg &q, int numTargs, const ComplexMatrixN *ops, int numOps, const char *f);
void multiQubitKrausMap(const Qureg &q, int numTargs, const ComplexMatrixN *ops, int numOps, const char *f);
void hamilParams(int numQubits, int numTerms, const char *f);
void pauliHamil(const PauliHamil &h, const char *f);
void matchingQuregPauliHamilDims(const Qureg &q, const PauliHamil &h, const char *f);
void trotterParams(int order, int reps, const char *f);

// diagonal operators (QuEST_validation.c:399-409, 436-441, 773-819)
void numQubitsInDiagOp(int numQubits, const char *f, int numRanks = 1);
void numElems(const DiagonalOp &op, long long start, long long n, const char *f);
void diagOpInit(const DiagonalOp &op, const char *f);
void diagonalOp(const Qureg &q, const DiagonalOp &op, const char *f);
void diagPauliHamil(const DiagonalOp &op, const PauliHamil &h, const char *f);
bool diagonalHamil(const PauliHamil &h);   // only PAULI_I and PAULI_Z codes
// phase functions (QuEST_validation.c:487-493, 821-1050)
void phaseMultiQubits(const Qureg &q, const int *qs, int n, const char *f);
void qubitSubregs(const Qureg &q, const int *qubits, const int *numQubitsPerReg, int numRegs, const char *f);
void bitEncoding(int numQubits, int encoding, const char *f);
void multiRegBitEncoding(const int *numQubitsPerReg, int numRegs, int encoding, const char *f);
void phaseFuncTerms(int numQubits, int encoding, const double *exponents, int numTerms,
                    const long long *overrideInds, int numOverrides, const char *f);
void multiVarPhaseFuncTerms(int numRegs, int encoding, const double *exponents, const int *numTermsPerReg,
                            const char *f);
void phaseFuncOverrides(int numQubits, int encoding, const long long *overrideInds, int numOverrides,
                        const char *f);
void multiVarPhaseFuncOverrides(const int *numQubitsPerReg, int numRegs, int encoding,
                                const long long *overrideInds, int numOverrides, const char *f);
void phaseFuncName(int funcCode, int numRegs, int numParams, const char *f);
void matchingPrecision(const Qureg &a, const Qureg &b, const char *f);
void precision(int p, const char *f);

}  // namespace qv
