/* Extensions of libquest_mi355x beyond the QuEST 3.4 API.
 *
 * None of these exist in the reference; they are additive (QuEST.h semantics
 * are unchanged) and carry the features the MI355X build adds:
 *   - per-register amplitude storage precision (FP32 bulk amplitudes with FP64
 *     reductions), chosen at creation;
 *   - an error mode for hosts that cannot override the weak
 *     invalidQuESTInputError symbol (ctypes, other FFIs);
 *   - device timing on the library's stream and device comparisons;
 *   - gate lists ("circuits") that the library can fuse into dense blocks.
 */
#ifndef QUEST_MI355X_H
#define QUEST_MI355X_H

#include "QuEST.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- storage precision ------------------------------------------------------ */
enum qmxPrecision { QMX_PREC_FP64 = 0, QMX_PREC_FP32 = 1 };

/* createQureg / createDensityQureg with the device amplitudes stored in the
 * given precision. Host-side values (mirror, getAmp, calc*) stay qreal. */
Qureg createQuregEx(int numQubits, QuESTEnv env, int precision);
Qureg createDensityQuregEx(int numQubits, QuESTEnv env, int precision);
int qmxQuregPrecision(Qureg qureg);

/* ---- errors ----------------------------------------------------------------- */
/* QMX_ERRORS_EXIT (default): QuEST behaviour, the weak invalidQuESTInputError
 * prints and exits. QMX_ERRORS_RECORD: the message is recorded, the failing
 * call returns without changing any state, and qmxLastError() reports it. */
enum qmxErrorMode { QMX_ERRORS_EXIT = 0, QMX_ERRORS_RECORD = 1 };
void qmxSetErrorMode(int mode);
/* copies "<func>: <message>" of the last recorded error into buf; returns 1 if
 * an error was recorded since the last qmxClearError(), else 0 */
int qmxLastError(char *buf, int len);
void qmxClearError(void);

/* ---- device ------------------------------------------------------------------ */
int qmxDeviceInfo(char *buf, int len);
void qmxSync(void);
/* HIP events on the library stream around the work queued in between */
void qmxTimerBegin(void);
double qmxTimerEndMs(void);
/* numbered event pairs (0 <= slot < 8192) around individual launches;
 * qmxTimerSlotMs waits for the slot's end event */
void qmxTimerSlotBegin(int slot);
void qmxTimerSlotEnd(int slot);
double qmxTimerSlotMs(int slot);
/* Density registers: operations on at most two qubits (gates, controlled
 * gates, channels, Kraus maps) compose into one superoperator that is swept
 * over the register once, when the next operation needs other qubits, another
 * register or the stream. Cumulative counts since load: sweeps made and
 * operations absorbed into them. */
void qmxDensityFusionStats(long long *sweeps, long long *operations);
/* QuantumLab's amplitude-encoding circuit for the real vector vec (2^numQubits
 * entries, encoded on qubits 0..numQubits-1, bit l <-> qubit l) applied to
 * |0...0>: the register ends in vec/|vec| (|0...0> for a zero vector), as
 * the reference's Ry tree + Gray-code ladder prepares it (angles and product
 * in FP64), in one launch instead of 2^numQubits gates. State vectors, at most
 * 12 encoded qubits; the register's other qubits are left |0>. */
void qmxAmplitudeEncodeReal(Qureg qureg, const qreal *vec, int numQubits);
/* QuantumLab's complex-vector encoder for v = re + i im (AmplitudeEncodeOneVector
 * with bHasPhase false, AmplitudeEncodeCuda.cu:691-724: the real encoder's
 * magnitude tree over |v_i| plus a phase tree over arg v_i, applied as Ry and
 * Rz(-Z) per level, ExchangeToYZGate :452-496 with FRyz) applied to |0...0>:
 * the register ends in v/|v| times exp(-i m), m the mean of the arg v_i (the
 * global phase the reference leaves to an EBO_Phase gate, which its QuEST
 * simulators do not apply). Angles and product in FP64, one launch; state
 * vectors, at most 12 encoded qubits (qubits 0..numQubits-1, bit l <-> qubit
 * l); the other qubits are left |0>. */
void qmxAmplitudeEncodeComplex(Qureg qureg, const qreal *re, const qreal *im, int numQubits);
/* A uniformly controlled (multiplexed) one-qubit gate in one sweep: wherever the
 * control qubits read k (bit b of k = qubit controls[b]), ops[k] acts on the
 * target. 2^numControls unitary matrices, 0 to 16 controls. State vectors and
 * density matrices (ops[k] on the rows, its conjugate on the columns). */
void qmxUniformlyControlledUnitary(Qureg qureg, const int *controls, int numControls, int target,
                                   const ComplexMatrix2 *ops);
enum qmxUniformGate { QMX_UNIFORM_RY = 2, QMX_UNIFORM_RZ = 3, QMX_UNIFORM_PHASE = 4 };
/* QuantumLab's uniformly controlled rotations FRy, FRz and FRp (FRGate.cu:
 * 181-330, 460-515), which it builds from 2^numControls rotations and CNOTs
 * (or X flips and multi-controlled phases), as one sweep: where the controls
 * read k, Ry(angles[k]) (rotateY's convention), Rz(angles[k]) (rotateZ's) or
 * the phase shift diag(1, e^{i angles[k]}) acts on the target. FRy(angles, n)
 * or FRz on register qubits b[0..n-1] is this call with target b[n-1] and
 * controls[j] = b[n-2-j]; FRp with controls[j] = b[j]. FRyz(Y, Z) is RY with
 * Y, then RZ with -Z. */
void qmxUniformlyControlledRotation(Qureg qureg, const int *controls, int numControls, int target,
                                    const qreal *angles, enum qmxUniformGate kind);
/* Repeated measurement of qubits[0..numQubits) in that order, as a loop that
 * restores the register and calls measure() on each qubit per shot would do
 * (QuantumLab's QLSimulatorMeasure.cpp:113-140), without touching the state:
 * one joint-probability reduction on the device, then every shot's chain of
 * outcomes drawn on the host from QuEST's generator (same draws, same order).
 * outcomes[s] has bit j set when qubits[j] measured 1. untilCount > 0 stops
 * after the first shot whose outcome has then occurred untilCount times
 * (the loop's m_iMeasureUntil mode). Returns the number of shots drawn.
 * At most 26 qubits. */
int qmxSampleShots(Qureg qureg, const int *qubits, int numQubits, int shots, int untilCount, int *outcomes);
/* calcProbOfAllOutcomes (QuEST.h, QuEST.c:1185-1192) without allocating any
 * device memory. No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc). Its scratch is a fixed 8 MiB device array per GPU
 * (2^20 doubles, held until the process ends), so it serves up to 2^19
 * outcomes (fewer when many of the qubits are below qubit 6 of a large
 * register); larger requests are rejected with QuEST's error convention. Calls
 * that overlap on one GPU share that array and race. Same results as
 * calcProbOfAllOutcomes (the same kernels). */
void qmxCalcProbOfAllOutcomesNoAlloc(qreal *outcomeProbs, Qureg qureg, int *qubits, int numQubits);
/* max |a_i - b_i| over amplitudes (any precisions, same size) */
qreal qmxMaxAbsDiff(Qureg a, Qureg b);
/* FP64 norm of the register; if rescale != 0 the register is renormalised */
qreal qmxRenormalise(Qureg qureg, int rescale);
/* QuantumLab's wave-function CSV (FermionSimulation/Fermion1D.cpp:13-83,
 * SaveWaveFunction / LoadWaveFunction): one "%f, %f" line (real, imaginary) per
 * amplitude, no header. Save writes the register's amplitudes (density
 * registers: the flat 4^n entries); returns 1. Load reads such a file into the
 * register; the line count must equal its amplitude count; returns the
 * amplitudes read, or 0 with the error reported. */
int qmxSaveStateCSV(Qureg qureg, const char *path);
long long qmxLoadStateCSV(Qureg qureg, const char *path);

/* ---- circuits and fusion ------------------------------------------------------
 * A circuit is an operation list for registers of one width, recorded once
 * (and validated then, with QuEST's messages) and applied to any register of
 * that width. It replaces QuantumLab's loop over QLGate::PerformBasicOperation
 * (QL/Circuit/QLGate.cpp:652-803, driven by QL/Simulator/QLSimulator*.cpp):
 * unfused, each operation is the QuEST call that loop makes; fused (state
 * vectors), the unitaries between collapses run as dense blocks of at most k
 * qubits, one sweep per block. */
typedef struct qmxCircuit qmxCircuit;

/* QuantumLab's basic operations: the codes of EBasicOperation
 * (QL/Circuit/QLGate.h:18-104). Two-qubit gates list the control first.
 * param: the angle (P, RX, RY, RZ, CP, CRX, CRY, CRZ), the outcome (CC) or the
 * probability (noise). Noise acts on density registers only and is ignored on
 * state vectors; PHASE and COMPOSITE are accepted and not dispatched, as in
 * PerformBasicOperation. */
enum qmxOperationCode {
    QMX_OP_H = 0, QMX_OP_X = 1, QMX_OP_Y = 2, QMX_OP_Z = 3, QMX_OP_P = 4, QMX_OP_PHASE = 5,
    QMX_OP_RX = 6, QMX_OP_RY = 7, QMX_OP_RZ = 8,
    QMX_OP_CX = 9, QMX_OP_CY = 10, QMX_OP_CZ = 11, QMX_OP_CP = 12, QMX_OP_CRX = 13, QMX_OP_CRY = 14,
    QMX_OP_CRZ = 15, QMX_OP_CC = 16,
    QMX_OP_NOISE_DAMPING = 17, QMX_OP_NOISE_DEPHASING = 18, QMX_OP_NOISE_DEPOLARISING = 19,
    QMX_OP_NOISE_MIX_PAULI_X = 20, QMX_OP_NOISE_MIX_PAULI_Y = 21, QMX_OP_NOISE_MIX_PAULI_Z = 22,
    QMX_OP_NOISE_MIX_PAULI_ALL = 23, QMX_OP_NOISE_TWO_QUBIT_DEPHASING = 24,
    QMX_OP_NOISE_TWO_QUBIT_DEPOLARISING = 25, QMX_OP_COMPOSITE = 26
};

/* one SBasicOperation (QL/Circuit/QLGate.h:106-112): m_eOperation,
 * m_lstQubits (at most two), m_fClassicalParameter */
typedef struct qmxBasicOperation {
    int operation;
    int numQubits;
    int qubits[2];
    qreal param;
} qmxBasicOperation;

/* The density simulator's noise after every operation that is neither noise
 * nor a collapse (QLSimulatorDensityMatrix.cpp:88-153; the members of its
 * parameters m_fDampingAfterGate .. m_fTwoDepolarisingAfterGate). Applied to
 * density registers only; zero disables a channel. */
typedef struct qmxNoiseAfterGate {
    qreal damping, dephasing, depolarising, mixPauli, twoQubitDephasing, twoQubitDepolarising;
} qmxNoiseAfterGate;

/* NULL (with the error reported) for an invalid width */
qmxCircuit *qmxCircuitCreate(int numQubits);
void qmxCircuitDestroy(qmxCircuit *c);
/* operations recorded, of every kind */
int qmxCircuitNumGates(const qmxCircuit *c);
void qmxCircuitAddOperation(qmxCircuit *c, int operation, const int *qubits, int numQubits, qreal param);
void qmxCircuitSetNoiseAfterGate(qmxCircuit *c, qmxNoiseAfterGate noise);
void qmxCircuitCompactUnitary(qmxCircuit *c, int target, Complex alpha, Complex beta);
void qmxCircuitUnitary(qmxCircuit *c, int target, ComplexMatrix2 u);
void qmxCircuitRotateZ(qmxCircuit *c, int target, qreal angle);
void qmxCircuitControlledNot(qmxCircuit *c, int control, int target);
void qmxCircuitMultiControlledUnitary(qmxCircuit *c, int *controls, int numControls, int target, ComplexMatrix2 u);
/* u is row-major 2^k x 2^k, re[r*dim+c]; targs[0] = least significant bit */
void qmxCircuitMultiQubitUnitary(qmxCircuit *c, int *targs, int numTargs, const qreal *re, const qreal *im);
/* group the unitaries into dense blocks of at most maxQubits (1..10) qubits,
 * collapses acting as barriers; returns the number of steps (blocks plus
 * collapses) */
int qmxCircuitFuse(qmxCircuit *c, int maxQubits);
/* dense blocks made by the last qmxCircuitFuse */
int qmxCircuitNumBlocks(const qmxCircuit *c);
/* Step `step` of the last qmxCircuitFuse (0 .. its return value - 1): for a
 * block, its qubits (ascending; targs[0] = least significant bit of the
 * matrix index) and its 2^k x 2^k matrix (row-major; any pointer may be NULL),
 * returning k; 0 for a collapse; -1 past the end. */
int qmxCircuitStep(const qmxCircuit *c, int step, int *targs, qreal *re, qreal *im);
/* Apply to a register of the circuit's width; returns the product of the
 * collapse probabilities (1 without collapses). fused != 0 on a state vector
 * uses the fused blocks (after qmxCircuitFuse), otherwise each operation is the
 * QuEST call PerformBasicOperation makes; density registers always run the
 * operations in order, each followed by the circuit's after-gate noise. */
qreal qmxCircuitApply(Qureg qureg, qmxCircuit *c, int fused);
/* The simulators' loop over an SBasicOperation list in one call: records ops
 * (validated), fuses into blocks of at most fuseQubits qubits when that is
 * non-zero and the register is a state vector, applies, and returns the
 * product of the collapse probabilities. noise may be NULL. */
qreal qmxApplyOperations(Qureg qureg, const qmxBasicOperation *ops, int numOps, const qmxNoiseAfterGate *noise,
                         int fuseQubits);

/* ---- batched swap tests (quantum k-means) ------------------------------------ */
/* A plan holds numPoints real vectors (point p has 2^pointQubits[p] entries,
 * packed back to back in `points`) and numCentroids centroids of 2^maxQubits
 * entries each. Every point is amplitude-encoded (uniformly controlled Ry
 * tree) into a q-qubit register and swap-tested against every centroid
 * truncated to the point's 2^q entries and L2-normalised, on a (2q+1)-qubit
 * register; P(ancilla = 0) = (1 + |<a|b>|^2)/2. One k-means iteration assigns
 * each point to the centroid of largest P0 and moves each centroid to the mean
 * of its points (coordinate j over the points that have it); it runs as one
 * persistent swap-test launch plus four small kernels, optionally replayed as
 * a captured HIP graph. precision: QMX_PREC_FP64 or QMX_PREC_FP32 (amplitudes;
 * sums are FP64 either way). Returns NULL on invalid sizes. */
typedef struct qmxSwapPlan qmxSwapPlan;
qmxSwapPlan *qmxSwapPlanCreate(int numPoints, const int *pointQubits, const double *points, int numCentroids,
                               int maxQubits, const double *centroids, int precision);
void qmxSwapPlanDestroy(qmxSwapPlan *plan);
/* times candidate tile shapes per register size on the device, keeps the fastest;
 * returns the number of timings made */
int qmxSwapPlanAutotune(qmxSwapPlan *plan);
int qmxSwapPlanRowsPerTile(const qmxSwapPlan *plan, int qubits);
int qmxSwapPlanNumTiles(const qmxSwapPlan *plan);
/* 32x32 amplitude blocks an iteration forms on the matrix cores (FP32 plans; 0 for FP64) */
long long qmxSwapPlanMfmaBlocks(const qmxSwapPlan *plan);
/* device time of one swap-test launch (ms), averaged over `launches` (overwrites the partials) */
double qmxSwapPlanTimeTiles(qmxSwapPlan *plan, int launches);
void qmxSwapPlanRun(qmxSwapPlan *plan, int iterations, int useGraph);
/* P(ancilla = 0) of the last iteration's tests, numPoints x numCentroids */
void qmxSwapPlanP0(qmxSwapPlan *plan, double *out);
void qmxSwapPlanAssignments(qmxSwapPlan *plan, int *out);
void qmxSwapPlanCentroids(qmxSwapPlan *plan, double *out);
/* replaces the centroids (numCentroids x 2^maxQubits) and re-encodes them;
 * tiles, tuned shapes and a captured graph are kept */
void qmxSwapPlanSetCentroids(qmxSwapPlan *plan, const double *centroids);
/* the encoded register state of point p (2^q amplitudes) */
void qmxSwapPlanPointState(qmxSwapPlan *plan, int p, double *out);

/* ---- Trotter evolution of many small registers (planner labels) ------------
 * Register j (hamils[j].numQubits <= 12 qubits, FP64, held in LDS) starts in
 * the product state prod_q rotateZ(phi_q) rotateY(theta_q)|0> (theta, phi packed
 * per register: hamils[0].numQubits values, then hamils[1]'s, ...) and is
 * evolved exactly as applyTrotterCircuit(reg, hamils[j], times[j], orders[j],
 * reps) would (QuEST.h:5450, QuEST_common.c:762-844). Return 1, or 0 on invalid
 * input (message on stderr). */
/* final states, packed 2^n amplitudes per register */
int qmxTrotterBatchEvolve(int numJobs, const PauliHamil *hamils, const qreal *times, const qreal *theta,
                          const qreal *phi, const int *orders, const int *reps, qreal *outReal, qreal *outImag);
/* infidelity[j] = 1 - |<psi_r|psi_4r>|^2 with r = reps[j] (the label judge) */
int qmxTrotterBatchInfidelity(int numJobs, const PauliHamil *hamils, const qreal *times, const qreal *theta,
                              const qreal *phi, const int *orders, const int *reps, qreal *infidelity);
/* ladders: for r = 2^k, k = 0..maxLog2, infidelity[j (maxLog2+1) + k] =
 * 1 - |<psi_r|psi_4r>|^2 until the first value <= eps (its k in passIndex[j],
 * -1 if none); entries not reached are NaN */
int qmxTrotterBatchLadder(int numJobs, const PauliHamil *hamils, const qreal *times, const qreal *theta,
                          const qreal *phi, const int *orders, int maxLog2, qreal eps, qreal *infidelity,
                          int *passIndex);

/* ---- learned Trotter planner (no counterpart in the reference) -------------
 * Picks the order (2 or 4) and the repetitions for applyTrotterCircuit that
 * reach a fidelity target, from 16 per-qubit features of the Pauli sum:
 *   0 t, 1 log2 t, 2 terms/n, 3 lambda/n (lambda = sum |c_j|), 4 log2(lambda/n),
 *   5 max |c_j|, 6 ||c||_2/sqrt(n), 7 log2(C1/n), 8 log2(C2/n),
 *   9 log2(lambda t/n), 10 log2(C1 t^2/n), 11 log2(C2 t^3/n),
 *   12 anticommuting partners per term, 13 mean Pauli weight, 14 max weight,
 *   15 share of lambda on terms of weight >= 3,
 * with C1 = sum_{j<k} ||[c_j P_j, c_k P_k]|| and
 * C2 = sum_{j<k} sum_l ||[c_l P_l, [c_j P_j, c_k P_k]]|| (identity strings
 * dropped; logs floored at -60). The network (16 -> 64 -> 64 -> 2, ReLU,
 * bfloat16-rounded weights, FP32 arithmetic) predicts the deviation
 * y_o = log2 r*_o - curve_o(n) for o = 2, 4 from a size curve (mean log2 r*_o
 * per register size, piecewise linear in log2 n, extended past its ends with
 * the end slopes; default log2(n)/(2o)); the planner takes
 * r_o = ceil(2^(y_o + curve_o(n) + margin)) and the order with fewer Pauli
 * exponentials (stages x r). Training sets the curve and the margin. */
typedef struct qmxPlanner qmxPlanner;
int qmxTrotterFeatures(PauliHamil hamil, qreal time, float *out16);
/* Pauli-exponential passes per repetition: 1 (order 1), 2, 10, 50, ... */
int qmxTrotterStages(int order);
/* state sweeps one repetition of applyTrotterCircuit(order) makes on a
 * register of this size and precision (fused rotation blocks plus rotations
 * too wide for a tile) */
int qmxTrotterSweepsPerRep(PauliHamil hamil, int order, int isDensityMatrix, int precision);
qmxPlanner *qmxPlannerCreate(unsigned seed);  /* He-uniform weights, zero biases */
void qmxPlannerDestroy(qmxPlanner *p);
int qmxPlannerNumParams(void);                 /* 5378 */
/* flat FP32 parameters: W1 (16x64 row-major), b1, W2 (64x64), b2, W3 (64x2), b3 */
void qmxPlannerGetParams(const qmxPlanner *p, float *out);
void qmxPlannerSetParams(qmxPlanner *p, const float *params);  /* also resets Adam */
/* inputs are standardised as (x - mean) / scale before the first layer */
void qmxPlannerSetNormalisation(qmxPlanner *p, const float *mean16, const float *scale16);
void qmxPlannerGetNormalisation(const qmxPlanner *p, float *mean16, float *scale16);
void qmxPlannerSetMargin(qmxPlanner *p, float log2Margin);
float qmxPlannerGetMargin(const qmxPlanner *p);
/* knots: log2 n ascending, log2reps knots x 2 (orders 2, 4); returns 1, or 0 if invalid */
int qmxPlannerSetSizeCurve(qmxPlanner *p, int knots, const float *log2n, const float *log2reps);
/* returns the knot count; either array may be NULL */
int qmxPlannerGetSizeCurve(const qmxPlanner *p, float *log2n, float *log2reps);
double qmxPlannerSizeCurve(const qmxPlanner *p, int order, int numQubits);
/* Y (N x 2) from X (N x 16) */
void qmxPlannerPredict(qmxPlanner *p, const float *X, int N, float *Y);
/* masked mean squared error sum(M (Y - T)^2) / max(1, sum M) and its gradient
 * (5378 floats) with respect to the rounded parameters; no update */
double qmxPlannerGradient(qmxPlanner *p, const float *X, const float *T, const float *M, int N, float *grad);
/* Adam (0.9, 0.999, 1e-8) over shuffled minibatches; returns the last epoch's
 * mean minibatch loss; lossPerEpoch (epochs floats) may be NULL */
double qmxPlannerTrain(qmxPlanner *p, const float *X, const float *T, const float *M, int N, int epochs, int batch,
                       float lr, unsigned seed, float *lossPerEpoch);
int qmxPlannerChoose(qmxPlanner *p, PauliHamil hamil, qreal time, int *order, int *reps);
/* applyTrotterCircuit at the planner's choice */
int qmxApplyTrotterLearned(Qureg qureg, PauliHamil hamil, qreal time, qmxPlanner *p, int *order, int *reps);
/* plain step doubling at a fixed order (the planner's fallback and judge):
 * r = 1, 2, 4, ... from the initial state until the normalised infidelity
 * 1 - |<psi_{r/2}|psi_r>|^2 / (|psi_{r/2}|^2 |psi_r|^2) <= eps;
 * qureg ends as psi_r; returns 1 if eps was met within maxReps */
int qmxApplyTrotterStepDoubling(Qureg qureg, PauliHamil hamil, qreal time, int order, qreal eps, int maxReps,
                                int *reps);
int qmxPlannerSave(const qmxPlanner *p, const char *path);
qmxPlanner *qmxPlannerLoad(const char *path);

#ifdef __cplusplus
}
#endif

#endif /* QUEST_MI355X_H */
