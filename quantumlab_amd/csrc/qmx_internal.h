// Internal interface between the QuEST C-API front-end (quest_api.cpp) and the
// HIP backend (qmx_*.hip). Nothing here is exported.
//
// Device data layout (one register):
//   amps   2^numQubitsInStateVec complex amplitudes, interleaved (re, im),
//          stored as double2 (QMX_PREC_FP64) or float2 (QMX_PREC_FP32).
//          Interleaving makes every amplitude one 16 B (or 8 B) vector load,
//          so a 1-target gate streams two arrays instead of the four SoA
//          streams of the reference GPU backend (QuEST_gpu.cu:783-844).
//   The host mirror of the Qureg is SoA qreal, as callers expect.
#pragma once

#include <cstdint>
#include <cstddef>

#include "QuEST.h"

namespace qmx {

enum Prec : int { FP64 = 0, FP32 = 1 };

struct Cplx {
    double re, im;
};

// per-register record; Qureg.firstLevelReduction points at it
struct QuregDev {
    uint32_t magic;          // kMagic while alive
    int prec;                // Prec
    int numQubits;           // qubits in the state vector (2n for density matrices)
    int isDensity;
    int numQubitsRep;        // qubits represented
    long long numAmps;
    void *amps;              // device amplitudes
    size_t hostBytes;        // bytes of each host mirror array (mmap'd)
};
constexpr uint32_t kMagic = 0x514d5833u;  // "QMX3"

inline QuregDev *dev(const Qureg &q) { return reinterpret_cast<QuregDev *>(q.firstLevelReduction); }
inline size_t ampBytes(int prec) { return prec == FP64 ? 16 : 8; }

// maximum bit positions a kernel accepts in one index-insertion list
constexpr int kMaxBits = 64;

struct BitList {
    int n;
    int pos[kMaxBits];  // ascending
};

// ---- runtime (qmx_runtime.hip) ---------------------------------------------
void runtimeInit();                 // device check + stream; idempotent
void *stream();                     // hipStream_t of the library
void deviceSync();
void *devAlloc(size_t bytes);       // fails loudly
void devFree(void *p);
// register buffers: small ones recycled through per-size free lists (nullptr on failure)
void *poolAlloc(size_t bytes);
void poolFree(void *p, size_t bytes);
void hostToDevice(void *dst, const void *src, size_t bytes);
void deviceToHost(void *dst, const void *src, size_t bytes);
double *pinnedScratch();            // 64 doubles of pinned host memory
void checkLaunch(const char *what);
void timerBegin();
double timerEndMs();
// one deferred task, run just before the next use of the stream (launch, copy,
// synchronisation, timer event); nullptr clears it
void setPendingWork(void (*fn)());
int deviceDescription(char *buf, int len);
int numCUs();                       // compute units of the device
// Staging ring for small per-launch inputs: fill Staged::host (bytes <=
// kStageSlotBytes), then stageCommit queues its copy on `stream`; Staged::dev
// stays valid for the next kernels on that stream. No allocation, and the host
// waits only when a slot comes round again before its previous copy finished.
constexpr size_t kStageSlotBytes = 16384;
struct Staged {
    void *host;
    void *dev;
    int slot;
};
Staged stageBegin(size_t bytes);
void stageCommit(const Staged &s, size_t bytes, void *stream);
// an input of any size through one growable pinned / device pair
const void *stageLarge(const void *src, size_t bytes, void *stream);
// raise a kernel's dynamic-LDS limit to `bytes` (beyond the 64 KiB default);
// false, with a message, when the device has less
bool allowDynamicLds(const void *kernel, size_t bytes, const char *caller);

// ---- state set-up (qmx_state.hip) ------------------------------------------
void initBlank(QuregDev *q);
void initClassical(QuregDev *q, long long index);          // statevec index
void initUniform(QuregDev *q, double re);                  // every amp = re (plus states)
void initDebug(QuregDev *q);
void initSingleQubit(QuregDev *q, int qubit, int outcome);
void initPureDensity(QuregDev *rho, const QuregDev *psi);  // rho = |psi><psi|
void copyState(QuregDev *dst, const QuregDev *src);        // any precisions
// host SoA arrays (double, or float for the QuEST_PREC 1 flavour) <-> device amplitudes
void uploadAmps(QuregDev *q, long long start, const double *re, const double *im, long long count);
void uploadAmps(QuregDev *q, long long start, const float *re, const float *im, long long count);
void downloadAmps(const QuregDev *q, long long start, double *re, double *im, long long count);
void downloadAmps(const QuregDev *q, long long start, float *re, float *im, long long count);
void weightedSum(Cplx f1, const QuregDev *q1, Cplx f2, const QuregDev *q2, Cplx fo, QuregDev *out);
void mixDensityMatrix(QuregDev *combine, double otherProb, const QuregDev *other);
void scaleAll(QuregDev *q, double factor);

// ---- gates (qmx_gates.hip) -------------------------------------------------
// u = [[u00,u01],[u10,u11]] on `target`, applied where (i & ctrlMask) ^ flipMask == ctrlMask
void applyMatrix2(QuregDev *q, int target, uint64_t ctrlMask, uint64_t flipMask, const Cplx u[4]);
// amplitude i <-> i ^ targMask for every i whose ctrl bits are all 1
void applyNot(QuregDev *q, uint64_t ctrlMask, uint64_t targMask);
void applySwap(QuregDev *q, int q1, int q2);
// amp *= term where (i & mask) == mask (phase shift family)
void applyPhaseOnAllOnes(QuregDev *q, uint64_t mask, Cplx term);
// amp *= exp(-i angle/2 * (+1/-1 by parity of i & targMask)) where ctrl bits are 1
void applyMultiRotateZ(QuregDev *q, uint64_t ctrlMask, uint64_t targMask, double angle);
// dense 2^k x 2^k matrix (row-major, u[r*dim+c]) on targets (targs[0] = least
// significant bit of the row index), applied where ctrl bits are 1
void applyMatrixN(QuregDev *q, uint64_t ctrlMask, const int *targs, int k, const Cplx *u);
// uniformly controlled 2x2: matrix k (4 entries, row-major) where the controls read k
void applyUniform2(QuregDev *q, const int *controls, int numControls, int target, const Cplx *u);
// dense 5-qubit block on an FP32 register via MFMA (qmx_fused.hip); prepared
// once (device-resident matrix), applied to registers of >= 9 qubits
void *fused5Prepare(const int *targs, const Cplx *u);
void fused5Free(void *prepared);
void fused5Apply(QuregDev *q, const void *prepared);
// one application without a prepared form (staged per call, no allocation)
void fused5ApplyOnce(QuregDev *q, const int *targs, const Cplx *u);
// uncontrolled 5-qubit matrix on an FP64 register of >= 11 qubits (MFMA tiles)
void fused5ApplyOnce64(QuregDev *q, const int *targs, const Cplx *u);
// exp(-i angle/2 * P), P = X^{xMask} Z^{zMask} with Y = iXZ folded in (single sweep)
void applyPauliRotation(QuregDev *q, uint64_t ctrlMask, uint64_t xMask, uint64_t zMask,
                        int numY, double angle);
// an ordered list of rotations exp(-i angle/2 P) applied `reps` times, runs of
// rotations whose X/Y supports fit one LDS tile fused into one sweep (qmx_pauli.hip)
struct PauliExp {
    uint64_t x, z;
    int ny;
    double angle;
};
void applyPauliSequence(QuregDev *q, const PauliExp *seq, int count, int reps);
int pauliSequenceBlocks(int numQubits, int prec, const PauliExp *seq, int count);

// amplitude-encoding circuit of `vec` (2^qbits reals, qbits <= 12) applied to
// |0...0> of the register (qmx_swaptest.hip)
void amplitudeEncode(QuregDev *q, const double *vec, int qbits);
void amplitudeEncodeComplex(QuregDev *q, const double *re, const double *im, int qbits);

// dense 2^B x 2^B matrix (row-major; index bit j <-> register bit bits[j]) on
// B = 2..5 arbitrary bits, one group per thread in registers (qmx_dense.hip)
bool denseBitsSupported(int B);
void applyDense(QuregDev *q, const int *bits, int B, const Cplx *M);
// two matrices on disjoint bit sets (4 + 4 or 4 + 2 bits) in one sweep;
// registers of at least 12 qubits
bool densePairSupported(const QuregDev *q, int BA, int BB);
void applyDensePair(QuregDev *q, const int *bitsA, int BA, const Cplx *MA, const int *bitsB, int BB,
                    const Cplx *MB);

// ---- reductions (qmx_reduce.hip); all accumulate in FP64 -------------------
double probOfZero(const QuregDev *q, int qubit);                   // state vector
double totalProb(const QuregDev *q);                               // state vector
Cplx innerProduct(const QuregDev *bra, const QuregDev *ket);
double densProbOfZero(const QuregDev *q, int qubit);               // Σ Re ρ_ii, bit = 0
double densTotalProb(const QuregDev *q);
double purity(const QuregDev *q);                                  // Σ |ρ_ij|^2
double densInnerProduct(const QuregDev *a, const QuregDev *b);     // Σ Re(a* b)
double hsDistanceSquared(const QuregDev *a, const QuregDev *b);
double densFidelity(const QuregDev *rho, const QuregDev *psi);     // <psi|rho|psi>
void probOfAllOutcomes(const QuregDev *q, const int *qubits, int k, double *out);
// device doubles probOfAllOutcomesInto needs as scratch for these qubits
long long outcomeScratchDoubles(const QuregDev *q, const int *qubits, int k);
// the same histogram with a caller-provided device scratch (outcomeScratchDoubles):
// results left on the device (the first 2^k doubles of scratch), or copied to out
const double *probOfAllOutcomesDevice(const QuregDev *q, const int *qubits, int k, double *scratch);
void probOfAllOutcomesInto(const QuregDev *q, const int *qubits, int k, double *scratch, double *out);
double normSquaredThenScale(QuregDev *q, bool rescale);            // FP64 norm, optional 1/sqrt renorm
double maxAbsDiff(const QuregDev *a, const QuregDev *b);           // max_i |a_i - b_i|
// <D> of a diagonal operator given as device FP64 arrays of 2^numQubitsRep entries
Cplx expecDiagonal(const QuregDev *q, const double *opRe, const double *opIm);

// ---- diagonal operators, phase functions (qmx_phase.hip) --------------------
// amp_i *= D_i (state vector), rho_rc *= D_r (density matrix: rows only, as the
// reference's densmatr_applyDiagonalOp)
void applyDiagonal(QuregDev *q, const double *opRe, const double *opIm);
// D_i = Σ_t c_t (-1)^parity(i & zMask_t) into device arrays of 2^numQubits entries
void diagonalFromZHamil(double *opRe, double *opIm, int numQubits, const uint64_t *zMasks, const double *coeffs,
                        int numTerms);
// amp *= exp(i phase(x_1..x_R)) where x_r is the value of sub-register r (qubits
// flattened, LSB first; encoding 0 unsigned, 1 two's complement). mode 0:
// Σ_r Σ_t coeffs x_r^exponents; mode 1: the named function funcCode with params.
// Overrides replace the phase at listed index tuples. conj negates the phase.
struct PhaseFuncArgs {
    int numRegs, encoding, mode, funcCode;
    const int *qubits, *numQubitsPerReg;   // host arrays
    const double *coeffs, *exponents;      // mode 0
    const int *numTermsPerReg;             // mode 0
    const double *params;                  // mode 1
    int numParams;
    const long long *overrideInds;         // numOverrides x numRegs
    const double *overridePhases;
    int numOverrides;
};
void applyPhaseFunction(QuregDev *q, const PhaseFuncArgs &args, int qubitShift, bool conj);

// ---- non-unitary (qmx_density.hip) -----------------------------------------
void collapse(QuregDev *q, int qubit, int outcome, double prob);
void densCollapse(QuregDev *q, int qubit, int outcome, double prob);
void densDephase(QuregDev *q, int qubit, double retain);
void densTwoQubitDephase(QuregDev *q, int q1, int q2, double retain);
void densDepolarise(QuregDev *q, int qubit, double depolLevel);
void densDamp(QuregDev *q, int qubit, double damping);

// Operations on density registers that act on at most two qubits compose on
// the host into one pending superoperator; it is applied in a single pass over
// the register when the next operation needs more qubits, another register or
// the stream. S is row-major over 4^k entries; index bits 0..k-1 are the row
// bits of targets[0..k-1] and bits k..2k-1 their column bits (rho column-major,
// element (r, c) at r + c 2^n). More than two targets: applied at once.
void densSuperOp(QuregDev *q, const int *targets, int k, const Cplx *S);
// the unitary u (rows) and uc (columns, conj(u) for a unitary) as a superoperator
void densUnitary(QuregDev *q, const int *targets, int k, const Cplx *u, const Cplx *uc);
void densDropPending(const QuregDev *q);        // register being destroyed
void densSuperOpStats(long long *passes, long long *absorbed);

}  // namespace qmx
