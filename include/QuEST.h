/* QuEST 3.4 C API, as exported by libquest_mi355x (MI355X / gfx950 HIP backend).
 *
 * This is the drop-in boundary: every declaration below keeps the name,
 * argument order, types, enum values and struct field layout of the reference
 * header Code/QuEST340/QuEST/include/QuEST.h, so QuantumLab's callers
 * (QL/Circuit/QLGate.cpp:652-803 and the QL/Simulator sources) and QuEST's own tests
 * compile and link against this library unchanged. Each declaration cites the
 * reference line it replaces.
 *
 * Semantics that callers rely on and that this library keeps:
 *  - structs are passed by value; a Qureg owns a host mirror
 *    (stateVec.real/imag, SoA, qreal) that is only meaningful after
 *    copyStateFromGPU() and only read by the device after copyStateToGPU();
 *  - invalid input calls invalidQuESTInputError(msg, func) before any state
 *    change; the default prints and exits, callers may override the weak
 *    symbol (or use qmxSetErrorMode() from quest_mi355x.h);
 *  - density matrices of n qubits are 2n-qubit vectors, element (r, c) at
 *    flat index r + c * 2^n;
 *  - work is queued on one HIP stream per process; every calc* returns a host
 *    value and therefore synchronises.
 *
 * QASM recording calls are accepted and record nothing (no QASM text output).
 */
#ifndef QUEST_H
#define QUEST_H

#define QUESTAPI

#ifndef M_PI
#define M_PI 3.1415926535897932384626433832795028841971
#endif

#include "QuEST_precision.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- data model (QuEST.h:94-400) ------------------------------------------ */

/* QuEST.h:87 */
enum phaseGateType { SIGMA_Z = 0, S_GATE = 1, T_GATE = 2 };

/* QuEST.h:94-99 — the QASM recorder; this library keeps the struct so the
 * Qureg layout is unchanged, but records nothing. */
typedef struct QASMLogger {
    char *buffer;
    int bufferSize;
    int bufferFill;
    int isLogging;
} QASMLogger;

/* QuEST.h:109-113 */
typedef struct ComplexArray {
    qreal *real;
    qreal *imag;
} ComplexArray;

/* QuEST.h:128 */
enum pauliOpType { PAULI_I = 0, PAULI_X = 1, PAULI_Y = 2, PAULI_Z = 3 };

/* QuEST.h:135-139 */
typedef struct Complex {
    qreal real;
    qreal imag;
} Complex;

/* QuEST.h:169-173 */
typedef struct ComplexMatrix2 {
    qreal real[2][2];
    qreal imag[2][2];
} ComplexMatrix2;

/* QuEST.h:207-211 */
typedef struct ComplexMatrix4 {
    qreal real[4][4];
    qreal imag[4][4];
} ComplexMatrix4;

/* QuEST.h:218-223 — row-pointer storage, real[r][c] */
typedef struct ComplexMatrixN {
    int numQubits;
    qreal **real;
    qreal **imag;
} ComplexMatrixN;

/* QuEST.h:230-233 */
typedef struct Vector {
    qreal x, y, z;
} Vector;

/* QuEST.h:263-266 */
enum phaseFunc {
    NORM = 0, SCALED_NORM = 1, INVERSE_NORM = 2, SCALED_INVERSE_NORM = 3, SCALED_INVERSE_SHIFTED_NORM = 4,
    PRODUCT = 5, SCALED_PRODUCT = 6, INVERSE_PRODUCT = 7, SCALED_INVERSE_PRODUCT = 8,
    DISTANCE = 9, SCALED_DISTANCE = 10, INVERSE_DISTANCE = 11, SCALED_INVERSE_DISTANCE = 12,
    SCALED_INVERSE_SHIFTED_DISTANCE = 13
};

/* QuEST.h:301 */
enum bitEncoding { UNSIGNED = 0, TWOS_COMPLEMENT = 1 };

/* QuEST.h:309-320 — pauliCodes[t * numQubits + q] is the operator of term t on qubit q */
typedef struct PauliHamil {
    enum pauliOpType *pauliCodes;
    qreal *termCoeffs;
    int numSumTerms;
    int numQubits;
} PauliHamil;

/* QuEST.h:329-345. Field use in this library:
 *   real, imag         host copy of the 2^numQubits diagonal (calloc'ed)
 *   deviceOperator     .real / .imag = device copies (FP64), kept in step by
 *                      initDiagonalOp, setDiagonalOpElems,
 *                      initDiagonalOpFromPauliHamil and syncDiagonalOp
 */
typedef struct DiagonalOp {
    int numQubits;
    long long int numElemsPerChunk;
    int numChunks;
    int chunkId;
    qreal *real;
    qreal *imag;
    ComplexArray deviceOperator;
} DiagonalOp;

/* QuEST.h:354-385. Field use in this library:
 *   stateVec           host mirror (SoA, qreal), lazily committed pages
 *   pairStateVec       unused (NULL): one GPU, no chunk exchange
 *   deviceStateVec     .real = device amplitude buffer (interleaved re,im in
 *                      the register's storage precision), .imag = NULL
 *   firstLevelReduction  opaque pointer to the library's per-register record
 *   secondLevelReduction unused (NULL)
 */
typedef struct Qureg {
    int isDensityMatrix;
    int numQubitsRepresented;
    int numQubitsInStateVec;
    long long int numAmpsPerChunk;
    long long int numAmpsTotal;
    int chunkId;
    int numChunks;
    ComplexArray stateVec;
    ComplexArray pairStateVec;
    ComplexArray deviceStateVec;
    qreal *firstLevelReduction, *secondLevelReduction;
    QASMLogger *qasmLog;
} Qureg;

/* QuEST.h:394-400 */
typedef struct QuESTEnv {
    int rank;
    int numRanks;
    unsigned long int *seeds;
    int numSeeds;
} QuESTEnv;

/* ---- environment (QuEST.h:1922-1983, 3469-3542) --------------------------- */
QuESTEnv QUESTAPI createQuESTEnv(void);                                   /* QuEST.h:1922 */
void QUESTAPI destroyQuESTEnv(QuESTEnv env);                              /* QuEST.h:1935 */
void QUESTAPI syncQuESTEnv(QuESTEnv env);                                 /* QuEST.h:1946 */
int syncQuESTSuccess(int successCode);                                    /* QuEST.h:1956 */
void QUESTAPI reportQuESTEnv(QuESTEnv env);                               /* QuEST.h:1964 */
void QUESTAPI getEnvironmentString(QuESTEnv env, char str[200]);          /* QuEST.h:1983 */
void QUESTAPI seedQuESTDefault(QuESTEnv *env);                            /* QuEST.h:3469 */
void QUESTAPI seedQuEST(QuESTEnv *env, unsigned long int *seedArray, int numSeeds); /* QuEST.h:3498 */
void QUESTAPI getQuESTSeeds(QuESTEnv env, unsigned long int **seeds, int *numSeeds); /* QuEST.h:3542 */

/* ---- registers and data structures ---------------------------------------- */
Qureg QUESTAPI createQureg(int numQubits, QuESTEnv env);                  /* QuEST.h:564 */
Qureg QUESTAPI createDensityQureg(int numQubits, QuESTEnv env);           /* QuEST.h:658 */
Qureg QUESTAPI createCloneQureg(Qureg qureg, QuESTEnv env);               /* QuEST.h:679 */
void QUESTAPI destroyQureg(Qureg qureg, QuESTEnv env);                    /* QuEST.h:701 */
ComplexMatrixN QUESTAPI createComplexMatrixN(int numQubits);              /* QuEST.h:756 */
void QUESTAPI destroyComplexMatrixN(ComplexMatrixN matr);                 /* QuEST.h:774 */
PauliHamil QUESTAPI createPauliHamil(int numQubits, int numSumTerms);     /* QuEST.h:839 */
void QUESTAPI destroyPauliHamil(PauliHamil hamil);                        /* QuEST.h:847 */
PauliHamil QUESTAPI createPauliHamilFromFile(char *fn);                   /* QuEST.h:895 */
void QUESTAPI initPauliHamil(PauliHamil hamil, qreal *coeffs, enum pauliOpType *codes); /* QuEST.h:934 */
void QUESTAPI reportState(Qureg qureg);                                   /* QuEST.h:1317 */
void QUESTAPI reportStateToScreen(Qureg qureg, QuESTEnv env, int reportRank); /* QuEST.h:1326 */
void QUESTAPI reportQuregParams(Qureg qureg);                             /* QuEST.h:1334 */
void QUESTAPI reportPauliHamil(PauliHamil hamil);                         /* QuEST.h:1358 */
int QUESTAPI getNumQubits(Qureg qureg);                                   /* QuEST.h:1370 */
long long int QUESTAPI getNumAmps(Qureg qureg);                           /* QuEST.h:1388 */

/* ---- initialisation (QuEST.h:1398-1630) ----------------------------------- */
void QUESTAPI initBlankState(Qureg qureg);
void QUESTAPI initZeroState(Qureg qureg);
void QUESTAPI initPlusState(Qureg qureg);
void QUESTAPI initClassicalState(Qureg qureg, long long int stateInd);
void QUESTAPI initPureState(Qureg qureg, Qureg pure);
void QUESTAPI initDebugState(Qureg qureg);
void QUESTAPI initStateFromAmps(Qureg qureg, qreal *reals, qreal *imags);
void QUESTAPI setAmps(Qureg qureg, long long int startInd, qreal *reals, qreal *imags, long long int numAmps);
void QUESTAPI setDensityAmps(Qureg qureg, long long int startRow, long long int startCol, qreal *reals, qreal *imags, long long int numAmps);
void QUESTAPI cloneQureg(Qureg targetQureg, Qureg copyQureg);             /* QuEST.h:1630 */

/* ---- host mirror (QuEST.h:2012-2215) -------------------------------------- */
void QUESTAPI copyStateToGPU(Qureg qureg);                                /* QuEST.h:2012 */
void QUESTAPI copyStateFromGPU(Qureg qureg);                              /* QuEST.h:2041 */
void QUESTAPI copySubstateToGPU(Qureg qureg, long long int startInd, long long int numAmps);   /* QuEST.h:2077 */
void QUESTAPI copySubstateFromGPU(Qureg qureg, long long int startInd, long long int numAmps); /* QuEST.h:2109 */
Complex QUESTAPI getAmp(Qureg qureg, long long int index);                /* QuEST.h:2130 */
qreal QUESTAPI getRealAmp(Qureg qureg, long long int index);              /* QuEST.h:2151 */
qreal QUESTAPI getImagAmp(Qureg qureg, long long int index);              /* QuEST.h:2172 */
qreal QUESTAPI getProbAmp(Qureg qureg, long long int index);              /* QuEST.h:2193 */
Complex QUESTAPI getDensityAmp(Qureg qureg, long long int row, long long int col); /* QuEST.h:2215 */

/* ---- unitaries (QuEST.h:1666-4945) ---------------------------------------- */
void QUESTAPI phaseShift(Qureg qureg, int targetQubit, qreal angle);                         /* :1666 */
void QUESTAPI controlledPhaseShift(Qureg qureg, int idQubit1, int idQubit2, qreal angle);    /* :1711 */
void QUESTAPI multiControlledPhaseShift(Qureg qureg, int *controlQubits, int numControlQubits, qreal angle); /* :1754 */
void QUESTAPI controlledPhaseFlip(Qureg qureg, int idQubit1, int idQubit2);                  /* :1794 */
void QUESTAPI multiControlledPhaseFlip(Qureg qureg, int *controlQubits, int numControlQubits); /* :1839 */
void QUESTAPI sGate(Qureg qureg, int targetQubit);                                           /* :1872 */
void QUESTAPI tGate(Qureg qureg, int targetQubit);                                           /* :1905 */
void QUESTAPI compactUnitary(Qureg qureg, int targetQubit, Complex alpha, Complex beta);     /* :2284 */
void QUESTAPI unitary(Qureg qureg, int targetQubit, ComplexMatrix2 u);                       /* :2325 */
void QUESTAPI rotateX(Qureg qureg, int rotQubit, qreal angle);                               /* :2363 */
void QUESTAPI rotateY(Qureg qureg, int rotQubit, qreal angle);                               /* :2402 */
void QUESTAPI rotateZ(Qureg qureg, int rotQubit, qreal angle);                               /* :2442 */
void QUESTAPI rotateAroundAxis(Qureg qureg, int rotQubit, qreal angle, Vector axis);          /* :2470 */
void QUESTAPI controlledRotateX(Qureg qureg, int controlQubit, int targetQubit, qreal angle); /* :2510 */
void QUESTAPI controlledRotateY(Qureg qureg, int controlQubit, int targetQubit, qreal angle); /* :2549 */
void QUESTAPI controlledRotateZ(Qureg qureg, int controlQubit, int targetQubit, qreal angle); /* :2589 */
void QUESTAPI controlledRotateAroundAxis(Qureg qureg, int controlQubit, int targetQubit, qreal angle, Vector axis); /* :2629 */
void QUESTAPI controlledCompactUnitary(Qureg qureg, int controlQubit, int targetQubit, Complex alpha, Complex beta); /* :2680 */
void QUESTAPI controlledUnitary(Qureg qureg, int controlQubit, int targetQubit, ComplexMatrix2 u); /* :2731 */
void QUESTAPI multiControlledUnitary(Qureg qureg, int *controlQubits, int numControlQubits, int targetQubit, ComplexMatrix2 u); /* :2795 */
void QUESTAPI pauliX(Qureg qureg, int targetQubit);                                          /* :2832 */
void QUESTAPI pauliY(Qureg qureg, int targetQubit);                                          /* :2867 */
void QUESTAPI pauliZ(Qureg qureg, int targetQubit);                                          /* :2905 */
void QUESTAPI hadamard(Qureg qureg, int targetQubit);                                        /* :2937 */
void QUESTAPI controlledNot(Qureg qureg, int controlQubit, int targetQubit);                 /* :2981 */
void QUESTAPI multiControlledMultiQubitNot(Qureg qureg, int *ctrls, int numCtrls, int *targs, int numTargs); /* :3057 */
void QUESTAPI multiQubitNot(Qureg qureg, int *targs, int numTargs);                          /* :3114 */
void QUESTAPI controlledPauliY(Qureg qureg, int controlQubit, int targetQubit);              /* :3156 */
void QUESTAPI swapGate(Qureg qureg, int qubit1, int qubit2);                                 /* :3969 */
void QUESTAPI sqrtSwapGate(Qureg qureg, int qb1, int qb2);                                   /* :4017 */
void QUESTAPI multiStateControlledUnitary(Qureg qureg, int *controlQubits, int *controlState, int numControlQubits, int targetQubit, ComplexMatrix2 u); /* :4078 */
void QUESTAPI multiRotateZ(Qureg qureg, int *qubits, int numQubits, qreal angle);            /* :4113 */
void QUESTAPI multiRotatePauli(Qureg qureg, int *targetQubits, enum pauliOpType *targetPaulis, int numTargets, qreal angle); /* :4168 */
void QUESTAPI multiControlledMultiRotateZ(Qureg qureg, int *controlQubits, int numControls, int *targetQubits, int numTargets, qreal angle); /* :4238 */
void QUESTAPI multiControlledMultiRotatePauli(Qureg qureg, int *controlQubits, int numControls, int *targetQubits, enum pauliOpType *targetPaulis, int numTargets, qreal angle); /* :4339 */
void QUESTAPI twoQubitUnitary(Qureg qureg, int targetQubit1, int targetQubit2, ComplexMatrix4 u); /* :4554 */
void QUESTAPI controlledTwoQubitUnitary(Qureg qureg, int controlQubit, int targetQubit1, int targetQubit2, ComplexMatrix4 u); /* :4621 */
void QUESTAPI multiControlledTwoQubitUnitary(Qureg qureg, int *controlQubits, int numControlQubits, int targetQubit1, int targetQubit2, ComplexMatrix4 u); /* :4700 */
void QUESTAPI multiQubitUnitary(Qureg qureg, int *targs, int numTargs, ComplexMatrixN u);    /* :4783 */
void QUESTAPI controlledMultiQubitUnitary(Qureg qureg, int ctrl, int *targs, int numTargs, ComplexMatrixN u); /* :4856 */
void QUESTAPI multiControlledMultiQubitUnitary(Qureg qureg, int *ctrls, int numCtrls, int *targs, int numTargs, ComplexMatrixN u); /* :4945 */

/* ---- non-unitary operators (QuEST.h:5267-5644) ---------------------------- */
void QUESTAPI setWeightedQureg(Complex fac1, Qureg qureg1, Complex fac2, Qureg qureg2, Complex facOut, Qureg out); /* :5267 */
void QUESTAPI applyPauliSum(Qureg inQureg, enum pauliOpType *allPauliCodes, qreal *termCoeffs, int numSumTerms, Qureg outQureg); /* :5326 */
void QUESTAPI applyPauliHamil(Qureg inQureg, PauliHamil hamil, Qureg outQureg);              /* :5370 */
void QUESTAPI applyTrotterCircuit(Qureg qureg, PauliHamil hamil, qreal time, int order, int reps); /* :5450 */
void QUESTAPI applyMatrix2(Qureg qureg, int targetQubit, ComplexMatrix2 u);                  /* :5471 */
void QUESTAPI applyMatrix4(Qureg qureg, int targetQubit1, int targetQubit2, ComplexMatrix4 u); /* :5523 */
void QUESTAPI applyMatrixN(Qureg qureg, int *targs, int numTargs, ComplexMatrixN u);         /* :5591 */
void QUESTAPI applyMultiControlledMatrixN(Qureg qureg, int *ctrls, int numCtrls, int *targs, int numTargs, ComplexMatrixN u); /* :5644 */
void QUESTAPI applyProjector(Qureg qureg, int qubit, int outcome);                           /* :6898 */

/* ---- calculations (QuEST.h:2242, 3190-3445, 3893-3925, 4390-4486, 5242) --- */
qreal QUESTAPI calcTotalProb(Qureg qureg);                                                   /* :2242 */
qreal QUESTAPI calcProbOfOutcome(Qureg qureg, int measureQubit, int outcome);                /* :3190 */
void QUESTAPI calcProbOfAllOutcomes(qreal *outcomeProbs, Qureg qureg, int *qubits, int numQubits); /* :3279 */
Complex QUESTAPI calcInnerProduct(Qureg bra, Qureg ket);                                     /* :3392 */
qreal QUESTAPI calcDensityInnerProduct(Qureg rho1, Qureg rho2);                              /* :3445 */
qreal QUESTAPI calcPurity(Qureg qureg);                                                      /* :3893 */
qreal QUESTAPI calcFidelity(Qureg qureg, Qureg pureState);                                   /* :3925 */
qreal QUESTAPI calcExpecPauliProd(Qureg qureg, int *targetQubits, enum pauliOpType *pauliCodes, int numTargets, Qureg workspace); /* :4390 */
qreal QUESTAPI calcExpecPauliSum(Qureg qureg, enum pauliOpType *allPauliCodes, qreal *termCoeffs, int numSumTerms, Qureg workspace); /* :4445 */
qreal QUESTAPI calcExpecPauliHamil(Qureg qureg, PauliHamil hamil, Qureg workspace);          /* :4486 */
qreal QUESTAPI calcHilbertSchmidtDistance(Qureg a, Qureg b);                                 /* :5242 */

/* ---- measurement (QuEST.h:3314-3365) -------------------------------------- */
qreal QUESTAPI collapseToOutcome(Qureg qureg, int measureQubit, int outcome);                /* :3314 */
int QUESTAPI measure(Qureg qureg, int measureQubit);                                         /* :3339 */
int QUESTAPI measureWithStats(Qureg qureg, int measureQubit, qreal *outcomeProb);            /* :3365 */

/* ---- decoherence (QuEST.h:3622-3865, 4991-5209) --------------------------- */
void QUESTAPI mixDephasing(Qureg qureg, int targetQubit, qreal prob);                        /* :3622 */
void QUESTAPI mixTwoQubitDephasing(Qureg qureg, int qubit1, int qubit2, qreal prob);         /* :3654 */
void QUESTAPI mixDepolarising(Qureg qureg, int targetQubit, qreal prob);                     /* :3697 */
void QUESTAPI mixDamping(Qureg qureg, int targetQubit, qreal prob);                          /* :3735 */
void QUESTAPI mixTwoQubitDepolarising(Qureg qureg, int qubit1, int qubit2, qreal prob);      /* :3802 */
void QUESTAPI mixPauli(Qureg qureg, int targetQubit, qreal probX, qreal probY, qreal probZ); /* :3843 */
void QUESTAPI mixDensityMatrix(Qureg combineQureg, qreal prob, Qureg otherQureg);            /* :3865 */
void QUESTAPI mixKrausMap(Qureg qureg, int target, ComplexMatrix2 *ops, int numOps);         /* :4991 */
void QUESTAPI mixTwoQubitKrausMap(Qureg qureg, int target1, int target2, ComplexMatrix4 *ops, int numOps); /* :5032 */
void QUESTAPI mixMultiQubitKrausMap(Qureg qureg, int *targets, int numTargets, ComplexMatrixN *ops, int numOps); /* :5084 */
void QUESTAPI mixNonTPKrausMap(Qureg qureg, int target, ComplexMatrix2 *ops, int numOps);    /* :5119 */
void QUESTAPI mixNonTPTwoQubitKrausMap(Qureg qureg, int target1, int target2, ComplexMatrix4 *ops, int numOps); /* :5158 */
void QUESTAPI mixNonTPMultiQubitKrausMap(Qureg qureg, int *targets, int numTargets, ComplexMatrixN *ops, int numOps); /* :5209 */

/* ---- diagonal operators (QuEST.h:1014-1292) -------------------------------- */
DiagonalOp QUESTAPI createDiagonalOp(int numQubits, QuESTEnv env);                                /* :1014 */
void QUESTAPI destroyDiagonalOp(DiagonalOp op, QuESTEnv env);                                      /* :1028 */
void QUESTAPI syncDiagonalOp(DiagonalOp op);                                                       /* :1048 */
void QUESTAPI initDiagonalOp(DiagonalOp op, qreal *real, qreal *imag);                              /* :1076 */
void QUESTAPI initDiagonalOpFromPauliHamil(DiagonalOp op, PauliHamil hamil);                       /* :1130 */
DiagonalOp QUESTAPI createDiagonalOpFromPauliHamilFile(char *fn, QuESTEnv env);                    /* :1174 */
void QUESTAPI setDiagonalOpElems(DiagonalOp op, long long int startInd, qreal *real, qreal *imag,
                                 long long int numElems);                                          /* :1222 */
void QUESTAPI applyDiagonalOp(Qureg qureg, DiagonalOp op);                                         /* :1260 */
Complex QUESTAPI calcExpecDiagonalOp(Qureg qureg, DiagonalOp op);                                  /* :1292 */

/* ---- phase functions and QFT (QuEST.h:5904-6874) --------------------------- */
void QUESTAPI applyPhaseFunc(Qureg qureg, int *qubits, int numQubits, enum bitEncoding encoding, qreal *coeffs,
                             qreal *exponents, int numTerms);                                      /* :5904 */
void QUESTAPI applyPhaseFuncOverrides(Qureg qureg, int *qubits, int numQubits, enum bitEncoding encoding,
                                      qreal *coeffs, qreal *exponents, int numTerms, long long int *overrideInds,
                                      qreal *overridePhases, int numOverrides);                    /* :6015 */
void QUESTAPI applyMultiVarPhaseFunc(Qureg qureg, int *qubits, int *numQubitsPerReg, int numRegs,
                                     enum bitEncoding encoding, qreal *coeffs, qreal *exponents,
                                     int *numTermsPerReg);                                         /* :6176 */
void QUESTAPI applyMultiVarPhaseFuncOverrides(Qureg qureg, int *qubits, int *numQubitsPerReg, int numRegs,
                                              enum bitEncoding encoding, qreal *coeffs, qreal *exponents,
                                              int *numTermsPerReg, long long int *overrideInds,
                                              qreal *overridePhases, int numOverrides);            /* :6258 */
void QUESTAPI applyNamedPhaseFunc(Qureg qureg, int *qubits, int *numQubitsPerReg, int numRegs,
                                  enum bitEncoding encoding, enum phaseFunc functionNameCode);     /* :6398 */
void QUESTAPI applyNamedPhaseFuncOverrides(Qureg qureg, int *qubits, int *numQubitsPerReg, int numRegs,
                                           enum bitEncoding encoding, enum phaseFunc functionNameCode,
                                           long long int *overrideInds, qreal *overridePhases,
                                           int numOverrides);                                      /* :6471 */
void QUESTAPI applyParamNamedPhaseFunc(Qureg qureg, int *qubits, int *numQubitsPerReg, int numRegs,
                                       enum bitEncoding encoding, enum phaseFunc functionNameCode, qreal *params,
                                       int numParams);                                             /* :6589 */
void QUESTAPI applyParamNamedPhaseFuncOverrides(Qureg qureg, int *qubits, int *numQubitsPerReg, int numRegs,
                                                enum bitEncoding encoding, enum phaseFunc functionNameCode,
                                                qreal *params, int numParams, long long int *overrideInds,
                                                qreal *overridePhases, int numOverrides);          /* :6664 */
void QUESTAPI applyFullQFT(Qureg qureg);                                                           /* :6758 */
void QUESTAPI applyQFT(Qureg qureg, int *qubits, int numQubits);                                   /* :6874 */

/* ---- C-only helpers for variable-length 2D arrays (QuEST.h:800, 5729-5795) -- */
#ifndef __cplusplus
#include <stddef.h> /* NULL in getStaticComplexMatrixN */
void QUESTAPI initComplexMatrixN(ComplexMatrixN m, qreal real[][1 << m.numQubits], qreal imag[][1 << m.numQubits]);
ComplexMatrixN bindArraysToStackComplexMatrixN(int numQubits, qreal re[][1 << numQubits], qreal im[][1 << numQubits],
                                               qreal **reStorage, qreal **imStorage);
#define UNPACK_ARR(...) __VA_ARGS__
#define getStaticComplexMatrixN(numQubits, re, im)                                                             \
    bindArraysToStackComplexMatrixN(numQubits, (qreal[1 << numQubits][1 << numQubits])UNPACK_ARR re,            \
                                    (qreal[1 << numQubits][1 << numQubits])UNPACK_ARR im,                      \
                                    (double *[1 << numQubits]){NULL}, (double *[1 << numQubits]){NULL})
#endif

/* ---- QASM recording (QuEST.h:3552-3591): accepted, records nothing -------- */
void QUESTAPI startRecordingQASM(Qureg qureg);
void QUESTAPI stopRecordingQASM(Qureg qureg);
void QUESTAPI clearRecordedQASM(Qureg qureg);
void QUESTAPI printRecordedQASM(Qureg qureg);
void QUESTAPI writeRecordedQASMToFile(Qureg qureg, char *filename);

/* ---- error hook (QuEST.h:5685; weak default in the library) --------------- */
void QUESTAPI invalidQuESTInputError(const char *errMsg, const char *errFunc);

#ifdef __cplusplus
}
#endif

#endif /* QUEST_H */
