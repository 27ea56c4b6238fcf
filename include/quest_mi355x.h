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
/* max |a_i - b_i| over amplitudes (any precisions, same size) */
qreal qmxMaxAbsDiff(Qureg a, Qureg b);
/* FP64 norm of the register; if rescale != 0 the register is renormalised */
qreal qmxRenormalise(Qureg qureg, int rescale);

/* ---- circuits and fusion ------------------------------------------------------ */
typedef struct qmxCircuit qmxCircuit;
qmxCircuit *qmxCircuitCreate(int numQubits);
void qmxCircuitDestroy(qmxCircuit *c);
int qmxCircuitNumGates(const qmxCircuit *c);
void qmxCircuitCompactUnitary(qmxCircuit *c, int target, Complex alpha, Complex beta);
void qmxCircuitUnitary(qmxCircuit *c, int target, ComplexMatrix2 u);
void qmxCircuitRotateZ(qmxCircuit *c, int target, qreal angle);
void qmxCircuitControlledNot(qmxCircuit *c, int control, int target);
void qmxCircuitMultiControlledUnitary(qmxCircuit *c, int *controls, int numControls, int target, ComplexMatrix2 u);
/* u is row-major 2^k x 2^k, re[r*dim+c]; targs[0] = least significant bit */
void qmxCircuitMultiQubitUnitary(qmxCircuit *c, int *targs, int numTargs, const qreal *re, const qreal *im);
/* group the gate list into dense blocks of at most maxQubits qubits;
 * returns the number of blocks */
int qmxCircuitFuse(qmxCircuit *c, int maxQubits);
int qmxCircuitNumBlocks(const qmxCircuit *c);
/* apply to a state vector: fused != 0 uses the fused blocks (qmxCircuitFuse
 * must have been called), otherwise one QuEST-equivalent kernel per gate */
void qmxCircuitApply(Qureg qureg, qmxCircuit *c, int fused);

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

#ifdef __cplusplus
}
#endif

#endif /* QUEST_MI355X_H */
