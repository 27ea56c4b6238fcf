/* cpuref — CPU restatement of QuEST 3.4's CPU backend for the hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the oracle: tests/, smoke() and the
 * cpu_baseline leg of bench.py load it as the checker or as the timed CPU
 * baseline; the product library never links, loads or calls it.
 *
 * It exports the same QuEST C API names as libquest_mi355x (qreal = double,
 * SoA host state), so one ctypes binding drives either. Each function follows
 * the reference loop it restates (file:line below, paths relative to
 * Code/QuEST340/QuEST/src); loops are OpenMP-parallel with static schedules
 * like the reference, arithmetic is plain double with -ffp-contract=off.
 *
 * Parity of this file is pinned in tests/ against: golden vectors written by
 * the compiled reference (tests/golden/), the reference library itself where
 * it is built (oracle/_ref), dense-matrix brute force on small registers, and
 * closed forms (norm preservation, Kraus sums, Trotter exactness cases).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <unistd.h>

#include "QuEST.h"

typedef long long ll;

/* ---- bit helpers (CPU/QuEST_cpu_internal.h:27-54) ------------------------- */
static inline ll zeroBitAt(ll x, int p) {
    ll low = x & ((1LL << p) - 1);
    return ((x >> p) << (p + 1)) | low;
}
static inline int bitOf(ll x, int p) { return (int)((x >> p) & 1LL); }
static int oddParity(ll x) {
    int p = 0;
    while (x) { p ^= 1; x &= x - 1; }
    return p;
}
static ll maskOf(const int *q, int n) {
    ll m = 0;
    for (int i = 0; i < n; i++) m |= 1LL << q[i];
    return m;
}

/* ---- MT19937 (mt19937ar.c:62-155, published algorithm) -------------------- */
#define MT_N 624
#define MT_M 397
static uint32_t mtS[MT_N];
static int mtI = MT_N + 1;
static void mtSeed(uint32_t s) {
    mtS[0] = s;
    for (int i = 1; i < MT_N; i++) mtS[i] = 1812433253u * (mtS[i - 1] ^ (mtS[i - 1] >> 30)) + (uint32_t)i;
    mtI = MT_N;
}
static void mtByArray(const unsigned long *key, int len) {
    mtSeed(19650218u);
    int i = 1, j = 0;
    for (int k = (MT_N > len ? MT_N : len); k; k--) {
        mtS[i] = (mtS[i] ^ ((mtS[i - 1] ^ (mtS[i - 1] >> 30)) * 1664525u)) + (uint32_t)key[j] + (uint32_t)j;
        i++; j++;
        if (i >= MT_N) { mtS[0] = mtS[MT_N - 1]; i = 1; }
        if (j >= len) j = 0;
    }
    for (int k = MT_N - 1; k; k--) {
        mtS[i] = (mtS[i] ^ ((mtS[i - 1] ^ (mtS[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
        i++;
        if (i >= MT_N) { mtS[0] = mtS[MT_N - 1]; i = 1; }
    }
    mtS[0] = 0x80000000u;
    mtI = MT_N;
}
static uint32_t mtNext(void) {
    if (mtI >= MT_N) {
        if (mtI == MT_N + 1) mtSeed(5489u);
        for (int k = 0; k < MT_N; k++) {
            uint32_t y = (mtS[k] & 0x80000000u) | (mtS[(k + 1) % MT_N] & 0x7fffffffu);
            uint32_t v = mtS[(k + MT_M) % MT_N] ^ (y >> 1);
            mtS[k] = (y & 1u) ? (v ^ 0x9908b0dfu) : v;
        }
        mtI = 0;
    }
    uint32_t y = mtS[mtI++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}
static double mtReal1(void) { return mtNext() * (1.0 / 4294967295.0); }
/* hooks so tests can pin the raw stream */
void cpurefMtInitByArray(const unsigned long *key, int len) { mtByArray(key, len); }
unsigned int cpurefMtNextU32(void) { return mtNext(); }

/* ---- environment ------------------------------------------------------------- */
void seedQuEST(QuESTEnv *env, unsigned long *seedArray, int numSeeds) {
    free(env->seeds);
    env->seeds = malloc(numSeeds * sizeof(unsigned long));
    for (int i = 0; i < numSeeds; i++) env->seeds[i] = seedArray[i];
    env->numSeeds = numSeeds;
    mtByArray(seedArray, numSeeds);
}
void seedQuESTDefault(QuESTEnv *env) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    unsigned long key[2] = {(unsigned long)((double)tv.tv_sec * 1000 + tv.tv_usec / 1000), (unsigned long)getpid()};
    seedQuEST(env, key, 2);
}
QuESTEnv createQuESTEnv(void) {
    QuESTEnv e = {0, 1, NULL, 0};
    seedQuESTDefault(&e);
    return e;
}
void destroyQuESTEnv(QuESTEnv env) { free(env.seeds); }
void syncQuESTEnv(QuESTEnv env) { (void)env; }
void copyStateToGPU(Qureg q) { (void)q; }
void copyStateFromGPU(Qureg q) { (void)q; }
void invalidQuESTInputError(const char *errMsg, const char *errFunc) {
    fprintf(stderr, "cpuref: invalid input to %s: %s\n", errFunc, errMsg);
    exit(EXIT_FAILURE);
}

/* ---- registers (CPU/QuEST_cpu.c:1297-1337, QuEST.c:42-86) --------------------- */
static Qureg newQureg(int nsv, int dens, int nrep) {
    Qureg q;
    memset(&q, 0, sizeof q);
    ll n = 1LL << nsv;
    q.isDensityMatrix = dens;
    q.numQubitsRepresented = nrep;
    q.numQubitsInStateVec = nsv;
    q.numAmpsPerChunk = n;
    q.numAmpsTotal = n;
    q.numChunks = 1;
    q.stateVec.real = calloc((size_t)n, sizeof(qreal));
    q.stateVec.imag = calloc((size_t)n, sizeof(qreal));
    q.stateVec.real[0] = 1.0;
    return q;
}
Qureg createQureg(int n, QuESTEnv env) { (void)env; return newQureg(n, 0, n); }
Qureg createDensityQureg(int n, QuESTEnv env) { (void)env; return newQureg(2 * n, 1, n); }
void destroyQureg(Qureg q, QuESTEnv env) {
    (void)env;
    free(q.stateVec.real);
    free(q.stateVec.imag);
}
Qureg createCloneQureg(Qureg q, QuESTEnv env) {
    Qureg c = newQureg(q.numQubitsInStateVec, q.isDensityMatrix, q.numQubitsRepresented);
    (void)env;
    memcpy(c.stateVec.real, q.stateVec.real, q.numAmpsTotal * sizeof(qreal));
    memcpy(c.stateVec.imag, q.stateVec.imag, q.numAmpsTotal * sizeof(qreal));
    return c;
}
ComplexMatrixN createComplexMatrixN(int nq) {
    int d = 1 << nq;
    ComplexMatrixN m;
    m.numQubits = nq;
    m.real = malloc(d * sizeof(qreal *));
    m.imag = malloc(d * sizeof(qreal *));
    for (int r = 0; r < d; r++) { m.real[r] = calloc(d, sizeof(qreal)); m.imag[r] = calloc(d, sizeof(qreal)); }
    return m;
}
void destroyComplexMatrixN(ComplexMatrixN m) {
    for (int r = 0; r < (1 << m.numQubits); r++) { free(m.real[r]); free(m.imag[r]); }
    free(m.real);
    free(m.imag);
}
PauliHamil createPauliHamil(int nq, int nt) {
    PauliHamil h;
    h.numQubits = nq;
    h.numSumTerms = nt;
    h.termCoeffs = malloc(nt * sizeof(qreal));
    h.pauliCodes = calloc((size_t)nq * nt, sizeof(enum pauliOpType));
    return h;
}
void destroyPauliHamil(PauliHamil h) { free(h.termCoeffs); free(h.pauliCodes); }

/* ---- initialisation (CPU/QuEST_cpu.c:1455-1660, 1135-1253) ---------------------- */
void initBlankState(Qureg q) {
    ll n = q.numAmpsTotal;
#pragma omp parallel for schedule(static)
    for (ll i = 0; i < n; i++) { q.stateVec.real[i] = 0; q.stateVec.imag[i] = 0; }
}
void initZeroState(Qureg q) { initBlankState(q); q.stateVec.real[0] = 1.0; }
void initClassicalState(Qureg q, ll ind) {
    initBlankState(q);
    ll k = q.isDensityMatrix ? ((1LL << q.numQubitsRepresented) + 1) * ind : ind;
    q.stateVec.real[k] = 1.0;
}
void initPlusState(Qureg q) {
    ll n = q.numAmpsTotal;
    qreal v = q.isDensityMatrix ? 1.0 / (qreal)(1LL << q.numQubitsRepresented) : 1.0 / sqrt((qreal)n);
#pragma omp parallel for schedule(static)
    for (ll i = 0; i < n; i++) { q.stateVec.real[i] = v; q.stateVec.imag[i] = 0; }
}
void initDebugState(Qureg q) {
    ll n = q.numAmpsTotal;
#pragma omp parallel for schedule(static)
    for (ll i = 0; i < n; i++) {
        q.stateVec.real[i] = (i * 2.0) / 10.0;
        q.stateVec.imag[i] = (i * 2.0 + 1.0) / 10.0;
    }
}
void initStateFromAmps(Qureg q, qreal *re, qreal *im) {
    memcpy(q.stateVec.real, re, q.numAmpsTotal * sizeof(qreal));
    memcpy(q.stateVec.imag, im, q.numAmpsTotal * sizeof(qreal));
}
void cloneQureg(Qureg t, Qureg c) { initStateFromAmps(t, c.stateVec.real, c.stateVec.imag); }
void initPureState(Qureg q, Qureg pure) {
    if (!q.isDensityMatrix) { cloneQureg(q, pure); return; }
    ll dim = 1LL << q.numQubitsRepresented;
    const qreal *vr = pure.stateVec.real, *vi = pure.stateVec.imag;
#pragma omp parallel for schedule(static)
    for (ll c = 0; c < dim; c++)
        for (ll r = 0; r < dim; r++) {
            qreal kr = vr[r], ki = vi[r], br = vr[c], bi = -vi[c];
            q.stateVec.real[r + c * dim] = kr * br - ki * bi;
            q.stateVec.imag[r + c * dim] = kr * bi + ki * br;
        }
}
Complex getAmp(Qureg q, ll i) { Complex c = {q.stateVec.real[i], q.stateVec.imag[i]}; return c; }
qreal getRealAmp(Qureg q, ll i) { return q.stateVec.real[i]; }
qreal getImagAmp(Qureg q, ll i) { return q.stateVec.imag[i]; }

/* ---- state-vector kernels ------------------------------------------------------------ */
/* 2x2 on target where (index ^ flip) & ctrl == ctrl
 * (CPU/QuEST_cpu.c:2020-2062 unitaryLocal, :2251-2317 multiControlledUnitaryLocal) */
static void sv_matrix2(Qureg q, int t, ll ctrl, ll flip, const qreal ur[2][2], const qreal ui[2][2]) {
    ll half = 1LL << t, tasks = q.numAmpsTotal >> 1;
    qreal *re = q.stateVec.real, *im = q.stateVec.imag;
#pragma omp parallel for schedule(static)
    for (ll k = 0; k < tasks; k++) {
        ll up = zeroBitAt(k, t), lo = up + half;
        if (ctrl && (ctrl & (up ^ flip)) != ctrl) continue;
        qreal ar = re[up], ai = im[up], br = re[lo], bi = im[lo];
        re[up] = ur[0][0] * ar - ui[0][0] * ai + ur[0][1] * br - ui[0][1] * bi;
        im[up] = ur[0][0] * ai + ui[0][0] * ar + ur[0][1] * bi + ui[0][1] * br;
        re[lo] = ur[1][0] * ar - ui[1][0] * ai + ur[1][1] * br - ui[1][1] * bi;
        im[lo] = ur[1][0] * ai + ui[1][0] * ar + ur[1][1] * bi + ui[1][1] * br;
    }
}
/* compactUnitary (CPU/QuEST_cpu.c:1737-1794) — same four-term expressions */
static void sv_compact(Qureg q, int t, Complex a, Complex b) {
    ll half = 1LL << t, tasks = q.numAmpsTotal >> 1;
    qreal *re = q.stateVec.real, *im = q.stateVec.imag;
    qreal aR = a.real, aI = a.imag, bR = b.real, bI = b.imag;
#pragma omp parallel for schedule(static)
    for (ll k = 0; k < tasks; k++) {
        ll up = zeroBitAt(k, t), lo = up + half;
        qreal rU = re[up], iU = im[up], rL = re[lo], iL = im[lo];
        re[up] = aR * rU - aI * iU - bR * rL - bI * iL;
        im[up] = aR * iU + aI * rU - bR * iL + bI * rL;
        re[lo] = bR * rU - bI * iU + aR * rL + aI * iL;
        im[lo] = bR * iU + bI * rU + aR * iL - aI * rL;
    }
}
/* controlled compact unitary (CPU/QuEST_cpu.c controlledCompactUnitaryLocal) */
static void sv_ccompact(Qureg q, ll ctrl, int t, Complex a, Complex b) {
    qreal ur[2][2] = {{a.real, -b.real}, {b.real, a.real}};
    qreal ui[2][2] = {{a.imag, b.imag}, {b.imag, -a.imag}};
    if (!ctrl) { sv_compact(q, t, a, b); return; }
    sv_matrix2(q, t, ctrl, 0, ur, ui);
}
/* multiControlledMultiQubitNot (CPU/QuEST_cpu.c:2761-2810), controlledNot (:2662-2710) */
static void sv_not(Qureg q, ll ctrl, ll targ) {
    ll n = q.numAmpsTotal;
    qreal *re = q.stateVec.real, *im = q.stateVec.imag;
#pragma omp parallel for schedule(static)
    for (ll i = 0; i < n; i++) {
        if ((i & ctrl) != ctrl) continue;
        ll j = i ^ targ;
        if (j < i) continue; /* each pair once */
        qreal r = re[i], m = im[i];
        re[i] = re[j]; im[i] = im[j];
        re[j] = r; im[j] = m;
    }
}
/* swapQubitAmps (CPU/QuEST_cpu.c:3905-3950) */
static void sv_swap(Qureg q, int a, int b) {
    ll n = q.numAmpsTotal;
    qreal *re = q.stateVec.real, *im = q.stateVec.imag;
#pragma omp parallel for schedule(static)
    for (ll i = 0; i < n; i++) {
        if (bitOf(i, a) != 0 || bitOf(i, b) != 1) continue;
        ll j = i ^ (1LL << a) ^ (1LL << b);
        qreal r = re[i], m = im[i];
        re[i] = re[j]; im[i] = im[j];
        re[j] = r; im[j] = m;
    }
}
/* phase on all-ones of mask (phaseShiftByTerm, multiControlledPhaseShift/Flip) */
static void sv_phase(Qureg q, ll mask, qreal cr, qreal ci) {
    ll n = q.numAmpsTotal;
    qreal *re = q.stateVec.real, *im = q.stateVec.imag;
#pragma omp parallel for schedule(static)
    for (ll i = 0; i < n; i++) {
        if ((i & mask) != mask) continue;
        qreal r = re[i], m = im[i];
        re[i] = cr * r - ci * m;
        im[i] = ci * r + cr * m;
    }
}
/* multiRotateZ (CPU/QuEST_cpu.c:3299-3339) */
static void sv_rotz(Qureg q, ll ctrl, ll mask, qreal angle) {
    ll n = q.numAmpsTotal;
    qreal *re = q.stateVec.real, *im = q.stateVec.imag;
    qreal c = cos(angle / 2.0), s = sin(angle / 2.0);
#pragma omp parallel for schedule(static)
    for (ll i = 0; i < n; i++) {
        if ((i & ctrl) != ctrl) continue;
        int f = oddParity(mask & i) ? -1 : 1;
        qreal r = re[i], m = im[i];
        re[i] = c * r + f * s * m;
        im[i] = -f * s * r + c * m;
    }
}
/* multiControlledMultiQubitUnitary (CPU/QuEST_cpu.c:1895-2007): targs[0] is the
 * least significant bit of the matrix index; amplitudes gathered per task */
static void sv_matrixN(Qureg q, ll ctrl, const int *targs, int k, const qreal *ur, const qreal *ui) {
    int dim = 1 << k;
    int sorted[64];
    for (int i = 0; i < k; i++) sorted[i] = targs[i];
    for (int i = 1; i < k; i++)
        for (int j = i; j > 0 && sorted[j - 1] > sorted[j]; j--) { int t = sorted[j]; sorted[j] = sorted[j - 1]; sorted[j - 1] = t; }
    ll tasks = q.numAmpsTotal >> k;
    qreal *re = q.stateVec.real, *im = q.stateVec.imag;
#pragma omp parallel
    {
        ll *ind = malloc(dim * sizeof(ll));
        qreal *ar = malloc(dim * sizeof(qreal)), *ai = malloc(dim * sizeof(qreal));
#pragma omp for schedule(static)
        for (ll t = 0; t < tasks; t++) {
            ll base = t;
            for (int j = 0; j < k; j++) base = zeroBitAt(base, sorted[j]);
            if (ctrl && (ctrl & base) != ctrl) continue;
            for (int i = 0; i < dim; i++) {
                ll x = base;
                for (int j = 0; j < k; j++)
                    if ((i >> j) & 1) x ^= 1LL << targs[j];
                ind[i] = x;
                ar[i] = re[x];
                ai[i] = im[x];
            }
            for (int r = 0; r < dim; r++) {
                ll x = ind[r];
                re[x] = 0;
                im[x] = 0;
                for (int c = 0; c < dim; c++) {
                    qreal er = ur[r * dim + c], ei = ui[r * dim + c];
                    re[x] += ar[c] * er - ai[c] * ei;
                    im[x] += ar[c] * ei + ai[c] * er;
                }
            }
        }
        free(ind); free(ar); free(ai);
    }
}

/* ---- front-end with density doubling (QuEST.c:182-720) -------------------------------- */
static int shiftOf(Qureg q) { return q.numQubitsRepresented; }

void compactUnitary(Qureg q, int t, Complex a, Complex b) {
    sv_compact(q, t, a, b);
    if (q.isDensityMatrix) {
        Complex ca = {a.real, -a.imag}, cb = {b.real, -b.imag};
        sv_compact(q, t + shiftOf(q), ca, cb);
    }
}
static void m2Split(ComplexMatrix2 u, int conj, qreal ur[2][2], qreal ui[2][2]) {
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) { ur[r][c] = u.real[r][c]; ui[r][c] = conj ? -u.imag[r][c] : u.imag[r][c]; }
}
static void gate2(Qureg q, int t, ll ctrl, ll flip, ComplexMatrix2 u) {
    qreal ur[2][2], ui[2][2];
    m2Split(u, 0, ur, ui);
    sv_matrix2(q, t, ctrl, flip, ur, ui);
    if (q.isDensityMatrix) {
        int s = shiftOf(q);
        m2Split(u, 1, ur, ui);
        sv_matrix2(q, t + s, ctrl << s, flip << s, ur, ui);
    }
}
void unitary(Qureg q, int t, ComplexMatrix2 u) { gate2(q, t, 0, 0, u); }
void controlledUnitary(Qureg q, int c, int t, ComplexMatrix2 u) { gate2(q, t, 1LL << c, 0, u); }
void multiControlledUnitary(Qureg q, int *cs, int nc, int t, ComplexMatrix2 u) { gate2(q, t, maskOf(cs, nc), 0, u); }
void multiStateControlledUnitary(Qureg q, int *cs, int *st, int nc, int t, ComplexMatrix2 u) {
    ll flip = 0;
    for (int i = 0; i < nc; i++) if (st[i] == 0) flip |= 1LL << cs[i];
    gate2(q, t, maskOf(cs, nc), flip, u);
}
/* QuEST_common.c:117-124 */
static void rotPair(qreal ang, Vector ax, Complex *a, Complex *b) {
    qreal m = sqrt(ax.x * ax.x + ax.y * ax.y + ax.z * ax.z);
    a->real = cos(ang / 2.0);
    a->imag = -sin(ang / 2.0) * (ax.z / m);
    b->real = sin(ang / 2.0) * (ax.y / m);
    b->imag = -sin(ang / 2.0) * (ax.x / m);
}
static void rotAxis(Qureg q, ll ctrl, int t, qreal ang, Vector ax) {
    Complex a, b;
    rotPair(ang, ax, &a, &b);
    sv_ccompact(q, ctrl, t, a, b);
    if (q.isDensityMatrix) {
        Complex ca = {a.real, -a.imag}, cb = {b.real, -b.imag};
        int s = shiftOf(q);
        sv_ccompact(q, ctrl << s, t + s, ca, cb);
    }
}
void rotateX(Qureg q, int t, qreal a) { Vector v = {1, 0, 0}; rotAxis(q, 0, t, a, v); }
void rotateY(Qureg q, int t, qreal a) { Vector v = {0, 1, 0}; rotAxis(q, 0, t, a, v); }
void rotateZ(Qureg q, int t, qreal a) { Vector v = {0, 0, 1}; rotAxis(q, 0, t, a, v); }
void rotateAroundAxis(Qureg q, int t, qreal a, Vector v) { rotAxis(q, 0, t, a, v); }
void controlledRotateX(Qureg q, int c, int t, qreal a) { Vector v = {1, 0, 0}; rotAxis(q, 1LL << c, t, a, v); }
void controlledRotateY(Qureg q, int c, int t, qreal a) { Vector v = {0, 1, 0}; rotAxis(q, 1LL << c, t, a, v); }
void controlledRotateZ(Qureg q, int c, int t, qreal a) { Vector v = {0, 0, 1}; rotAxis(q, 1LL << c, t, a, v); }
void controlledCompactUnitary(Qureg q, int c, int t, Complex a, Complex b) {
    sv_ccompact(q, 1LL << c, t, a, b);
    if (q.isDensityMatrix) {
        Complex ca = {a.real, -a.imag}, cb = {b.real, -b.imag};
        int s = shiftOf(q);
        sv_ccompact(q, 1LL << (c + s), t + s, ca, cb);
    }
}
void hadamard(Qureg q, int t) {
    qreal r = 1 / sqrt(2.0);
    qreal ur[2][2] = {{r, r}, {r, -r}}, ui[2][2] = {{0, 0}, {0, 0}};
    sv_matrix2(q, t, 0, 0, ur, ui);
    if (q.isDensityMatrix) sv_matrix2(q, t + shiftOf(q), 0, 0, ur, ui);
}
void pauliX(Qureg q, int t) {
    sv_not(q, 0, 1LL << t);
    if (q.isDensityMatrix) sv_not(q, 0, 1LL << (t + shiftOf(q)));
}
void pauliY(Qureg q, int t) {
    qreal ur[2][2] = {{0, 0}, {0, 0}}, ui[2][2] = {{0, -1}, {1, 0}}, uc[2][2] = {{0, 1}, {-1, 0}};
    sv_matrix2(q, t, 0, 0, ur, ui);
    if (q.isDensityMatrix) sv_matrix2(q, t + shiftOf(q), 0, 0, ur, uc);
}
/* QuEST.c:569-578, QuEST_cpu.c:2965-3016: {{0,-i},{i,0}} on t where bit c is 1;
 * density registers: its conjugate on the column qubits (conjFac = -1) */
void controlledPauliY(Qureg q, int c, int t) {
    qreal ur[2][2] = {{0, 0}, {0, 0}}, ui[2][2] = {{0, -1}, {1, 0}}, uc[2][2] = {{0, 1}, {-1, 0}};
    sv_matrix2(q, t, 1LL << c, 0, ur, ui);
    if (q.isDensityMatrix) {
        int s = shiftOf(q);
        sv_matrix2(q, t + s, 1LL << (c + s), 0, ur, uc);
    }
}
static void phaseBoth(Qureg q, ll mask, qreal cr, qreal ci) {
    sv_phase(q, mask, cr, ci);
    if (q.isDensityMatrix) sv_phase(q, mask << shiftOf(q), cr, -ci);
}
void pauliZ(Qureg q, int t) { phaseBoth(q, 1LL << t, -1, 0); }
void sGate(Qureg q, int t) { phaseBoth(q, 1LL << t, 0, 1); }
void tGate(Qureg q, int t) { phaseBoth(q, 1LL << t, 1 / sqrt(2.0), 1 / sqrt(2.0)); }
void phaseShift(Qureg q, int t, qreal a) { phaseBoth(q, 1LL << t, cos(a), sin(a)); }
void controlledPhaseShift(Qureg q, int a, int b, qreal ang) { phaseBoth(q, (1LL << a) | (1LL << b), cos(ang), sin(ang)); }
void controlledPhaseFlip(Qureg q, int a, int b) { phaseBoth(q, (1LL << a) | (1LL << b), -1, 0); }
void multiControlledPhaseFlip(Qureg q, int *cs, int n) { phaseBoth(q, maskOf(cs, n), -1, 0); }
void multiControlledPhaseShift(Qureg q, int *cs, int n, qreal a) { phaseBoth(q, maskOf(cs, n), cos(a), sin(a)); }
void controlledNot(Qureg q, int c, int t) {
    sv_not(q, 1LL << c, 1LL << t);
    if (q.isDensityMatrix) { int s = shiftOf(q); sv_not(q, 1LL << (c + s), 1LL << (t + s)); }
}
void multiControlledMultiQubitNot(Qureg q, int *cs, int nc, int *ts, int nt) {
    sv_not(q, maskOf(cs, nc), maskOf(ts, nt));
    if (q.isDensityMatrix) { int s = shiftOf(q); sv_not(q, maskOf(cs, nc) << s, maskOf(ts, nt) << s); }
}
void multiQubitNot(Qureg q, int *ts, int nt) {
    sv_not(q, 0, maskOf(ts, nt));
    if (q.isDensityMatrix) sv_not(q, 0, maskOf(ts, nt) << shiftOf(q));
}
void swapGate(Qureg q, int a, int b) {
    sv_swap(q, a, b);
    if (q.isDensityMatrix) sv_swap(q, a + shiftOf(q), b + shiftOf(q));
}
static void matNBoth(Qureg q, ll ctrl, int *targs, int k, const qreal *ur, const qreal *ui) {
    sv_matrixN(q, ctrl, targs, k, ur, ui);
    if (q.isDensityMatrix) {
        int s = shiftOf(q), dim = 1 << k, t2[64];
        qreal *uc = malloc((size_t)dim * dim * sizeof(qreal));
        for (int i = 0; i < dim * dim; i++) uc[i] = -ui[i];
        for (int j = 0; j < k; j++) t2[j] = targs[j] + s;
        sv_matrixN(q, ctrl << s, t2, k, ur, uc);
        free(uc);
    }
}
static void flatN(ComplexMatrixN u, qreal *ur, qreal *ui) {
    int d = 1 << u.numQubits;
    for (int r = 0; r < d; r++)
        for (int c = 0; c < d; c++) { ur[r * d + c] = u.real[r][c]; ui[r * d + c] = u.imag[r][c]; }
}
void multiControlledMultiQubitUnitary(Qureg q, int *cs, int nc, int *ts, int nt, ComplexMatrixN u) {
    int d = 1 << nt;
    qreal *ur = malloc((size_t)d * d * sizeof(qreal)), *ui = malloc((size_t)d * d * sizeof(qreal));
    flatN(u, ur, ui);
    matNBoth(q, maskOf(cs, nc), ts, nt, ur, ui);
    free(ur); free(ui);
}
void multiQubitUnitary(Qureg q, int *ts, int nt, ComplexMatrixN u) { multiControlledMultiQubitUnitary(q, NULL, 0, ts, nt, u); }
void controlledMultiQubitUnitary(Qureg q, int c, int *ts, int nt, ComplexMatrixN u) {
    int cs[1] = {c};
    multiControlledMultiQubitUnitary(q, cs, 1, ts, nt, u);
}
void twoQubitUnitary(Qureg q, int a, int b, ComplexMatrix4 u) {
    qreal ur[16], ui[16];
    for (int r = 0; r < 4; r++) for (int c = 0; c < 4; c++) { ur[r * 4 + c] = u.real[r][c]; ui[r * 4 + c] = u.imag[r][c]; }
    int t[2] = {a, b};
    matNBoth(q, 0, t, 2, ur, ui);
}
void multiRotateZ(Qureg q, int *qs, int n, qreal a) {
    sv_rotz(q, 0, maskOf(qs, n), a);
    if (q.isDensityMatrix) sv_rotz(q, 0, maskOf(qs, n) << shiftOf(q), -a);
}
/* QuEST_common.c:411-448: basis change to Z, multiRotateZ, undo */
static void sv_rotPauli(Qureg q, int *targs, const enum pauliOpType *ps, int n, qreal angle, int conj) {
    qreal f = 1 / sqrt(2.0);
    Complex rxA = {f, 0}, rxB = {0, conj ? f : -f};
    Complex ryA = {f, 0}, ryB = {-f, 0};
    ll mask = maskOf(targs, n);
    for (int t = 0; t < n; t++) {
        if (ps[t] == PAULI_I) mask -= 1LL << targs[t];
        if (ps[t] == PAULI_X) sv_compact(q, targs[t], ryA, ryB);
        if (ps[t] == PAULI_Y) sv_compact(q, targs[t], rxA, rxB);
    }
    if (mask != 0) sv_rotz(q, 0, mask, conj ? -angle : angle);
    rxB.imag *= -1;
    ryB.real *= -1;
    for (int t = 0; t < n; t++) {
        if (ps[t] == PAULI_X) sv_compact(q, targs[t], ryA, ryB);
        if (ps[t] == PAULI_Y) sv_compact(q, targs[t], rxA, rxB);
    }
}
void multiRotatePauli(Qureg q, int *targs, enum pauliOpType *ps, int n, qreal angle) {
    sv_rotPauli(q, targs, ps, n, angle, 0);
    if (q.isDensityMatrix) {
        int t2[64];
        for (int i = 0; i < n; i++) t2[i] = targs[i] + shiftOf(q);
        sv_rotPauli(q, t2, ps, n, angle, 1);
    }
}
/* QuEST_common.c:762-844 */
static void expHamil(Qureg q, PauliHamil h, qreal fac, int reverse) {
    int vt[64], dt[64];
    for (int i = 0; i < h.numQubits; i++) { vt[i] = i; dt[i] = i + h.numQubits; }
    for (int i = 0; i < h.numSumTerms; i++) {
        int t = reverse ? h.numSumTerms - 1 - i : i;
        qreal angle = 2 * fac * h.termCoeffs[t];
        sv_rotPauli(q, vt, &h.pauliCodes[t * h.numQubits], h.numQubits, angle, 0);
        if (q.isDensityMatrix) sv_rotPauli(q, dt, &h.pauliCodes[t * h.numQubits], h.numQubits, angle, 1);
    }
}
static void symTrotter(Qureg q, PauliHamil h, qreal time, int order) {
    if (order == 1) {
        expHamil(q, h, time, 0);
    } else if (order == 2) {
        expHamil(q, h, time / 2., 0);
        expHamil(q, h, time / 2., 1);
    } else {
        qreal p = 1. / (4 - pow(4, 1. / (order - 1)));
        int lo = order - 2;
        symTrotter(q, h, p * time, lo);
        symTrotter(q, h, p * time, lo);
        symTrotter(q, h, (1 - 4 * p) * time, lo);
        symTrotter(q, h, p * time, lo);
        symTrotter(q, h, p * time, lo);
    }
}
void applyTrotterCircuit(Qureg q, PauliHamil h, qreal time, int order, int reps) {
    if (time == 0) return;
    for (int r = 0; r < reps; r++) symTrotter(q, h, time / reps, order);
}

/* ---- calculations (CPU/QuEST_cpu_local.c:121-170, CPU/QuEST_cpu.c:879-1129, 3385-3487) ---- */
qreal calcTotalProb(Qureg q) {
    /* Kahan summation in the reference's bracket order */
    qreal tot = 0, y, t, c = 0;
    if (q.isDensityMatrix) {
        ll d = 1LL << q.numQubitsRepresented;
        for (ll col = 0; col < d; col++) {
            y = q.stateVec.real[col * (d + 1)] - c;
            t = tot + y;
            c = (t - tot) - y;
            tot = t;
        }
        return tot;
    }
    for (ll i = 0; i < q.numAmpsTotal; i++) {
        y = q.stateVec.real[i] * q.stateVec.real[i] - c;
        t = tot + y; c = (t - tot) - y; tot = t;
        y = q.stateVec.imag[i] * q.stateVec.imag[i] - c;
        t = tot + y; c = (t - tot) - y; tot = t;
    }
    return tot;
}
static qreal probZero(Qureg q, int m) {
    qreal tot = 0;
    if (q.isDensityMatrix) {
        ll d = 1LL << q.numQubitsRepresented;
#pragma omp parallel for schedule(static) reduction(+ : tot)
        for (ll r = 0; r < d; r++)
            if (bitOf(r, m) == 0) tot += q.stateVec.real[r * (d + 1)];
        return tot;
    }
    ll tasks = q.numAmpsTotal >> 1;
#pragma omp parallel for schedule(static) reduction(+ : tot)
    for (ll k = 0; k < tasks; k++) {
        ll i = zeroBitAt(k, m);
        tot += q.stateVec.real[i] * q.stateVec.real[i] + q.stateVec.imag[i] * q.stateVec.imag[i];
    }
    return tot;
}
qreal calcProbOfOutcome(Qureg q, int m, int o) {
    qreal p = probZero(q, m);
    return o == 1 ? 1.0 - p : p;
}
void calcProbOfAllOutcomes(qreal *out, Qureg q, int *qs, int n) {
    int no = 1 << n;
    for (int o = 0; o < no; o++) out[o] = 0;
    ll d = q.isDensityMatrix ? (1LL << q.numQubitsRepresented) : q.numAmpsTotal;
    for (ll i = 0; i < d; i++) {
        int o = 0;
        for (int j = 0; j < n; j++) o |= bitOf(i, qs[j]) << j;
        if (q.isDensityMatrix) out[o] += q.stateVec.real[i * (d + 1)];
        else out[o] += q.stateVec.real[i] * q.stateVec.real[i] + q.stateVec.imag[i] * q.stateVec.imag[i];
    }
}
Complex calcInnerProduct(Qureg bra, Qureg ket) {
    qreal sr = 0, si = 0;
    ll n = bra.numAmpsTotal;
#pragma omp parallel for schedule(static) reduction(+ : sr, si)
    for (ll i = 0; i < n; i++) {
        qreal br = bra.stateVec.real[i], bi = bra.stateVec.imag[i], kr = ket.stateVec.real[i], ki = ket.stateVec.imag[i];
        sr += br * kr + bi * ki;
        si += br * ki - bi * kr;
    }
    Complex c = {sr, si};
    return c;
}
qreal calcPurity(Qureg q) {
    qreal t = 0;
    ll n = q.numAmpsTotal;
#pragma omp parallel for schedule(static) reduction(+ : t)
    for (ll i = 0; i < n; i++) t += q.stateVec.real[i] * q.stateVec.real[i] + q.stateVec.imag[i] * q.stateVec.imag[i];
    return t;
}
qreal calcDensityInnerProduct(Qureg a, Qureg b) {
    qreal t = 0;
    ll n = a.numAmpsTotal;
#pragma omp parallel for schedule(static) reduction(+ : t)
    for (ll i = 0; i < n; i++) t += a.stateVec.real[i] * b.stateVec.real[i] + a.stateVec.imag[i] * b.stateVec.imag[i];
    return t;
}
qreal calcHilbertSchmidtDistance(Qureg a, Qureg b) {
    qreal t = 0;
    ll n = a.numAmpsTotal;
#pragma omp parallel for schedule(static) reduction(+ : t)
    for (ll i = 0; i < n; i++) {
        qreal dr = a.stateVec.real[i] - b.stateVec.real[i], di = a.stateVec.imag[i] - b.stateVec.imag[i];
        t += dr * dr + di * di;
    }
    return sqrt(t);
}
/* densmatr_calcFidelityLocal (CPU/QuEST_cpu.c:1008-1087): row sums then conj(psi_r) */
qreal calcFidelity(Qureg q, Qureg pure) {
    if (!q.isDensityMatrix) {
        Complex ip = calcInnerProduct(q, pure);
        return ip.real * ip.real + ip.imag * ip.imag;
    }
    ll dim = pure.numAmpsTotal;
    const qreal *vr = pure.stateVec.real, *vi = pure.stateVec.imag, *dr = q.stateVec.real, *di = q.stateVec.imag;
    qreal tot = 0;
#pragma omp parallel for schedule(static) reduction(+ : tot)
    for (ll r = 0; r < dim; r++) {
        qreal sr = 0, si = 0;
        for (ll c = 0; c < dim; c++) {
            qreal er = dr[r + dim * c], ei = di[r + dim * c];
            sr += er * vr[c] - ei * vi[c];
            si += er * vi[c] + ei * vr[c];
        }
        tot += sr * vr[r] - si * (-vi[r]);
    }
    return tot;
}

/* ---- measurement (QuEST_common.c:165-180, CPU/QuEST_cpu.c:3750-3800, 798-877) ---------------- */
static void collapseKnown(Qureg q, int m, int o, qreal p) {
    ll n = q.numAmpsTotal;
    qreal *re = q.stateVec.real, *im = q.stateVec.imag;
    if (q.isDensityMatrix) {
        int cb = m + shiftOf(q);
#pragma omp parallel for schedule(static)
        for (ll i = 0; i < n; i++) {
            if (bitOf(i, m) == o && bitOf(i, cb) == o) { re[i] /= p; im[i] /= p; }
            else { re[i] = 0; im[i] = 0; }
        }
        return;
    }
    qreal renorm = 1 / sqrt(p);
#pragma omp parallel for schedule(static)
    for (ll i = 0; i < n; i++) {
        if (bitOf(i, m) == o) { re[i] *= renorm; im[i] *= renorm; }
        else { re[i] = 0; im[i] = 0; }
    }
}
static int drawOutcome(qreal p0, qreal *outP) {
    int o;
    if (p0 < REAL_EPS) o = 1;
    else if (1 - p0 < REAL_EPS) o = 0;
    else o = (mtReal1() > p0);
    *outP = (o == 0) ? p0 : 1 - p0;
    return o;
}
qreal collapseToOutcome(Qureg q, int m, int o) {
    qreal p = calcProbOfOutcome(q, m, o);
    collapseKnown(q, m, o, p);
    return p;
}
int measureWithStats(Qureg q, int m, qreal *p) {
    int o = drawOutcome(probZero(q, m), p);
    collapseKnown(q, m, o, *p);
    return o;
}
int measure(Qureg q, int m) { qreal p; return measureWithStats(q, m, &p); }

/* ---- decoherence (CPU/QuEST_cpu.c:61-235, QuEST.c:1259-1298) --------------------------------- */
void mixDephasing(Qureg q, int t, qreal prob) {
    qreal retain = 1 - 2 * prob;
    ll inner = 1LL << t, outer = 1LL << (t + shiftOf(q)), tot = inner | outer, n = q.numAmpsTotal;
#pragma omp parallel for schedule(static)
    for (ll i = 0; i < n; i++) {
        ll pat = i & tot;
        if (pat == inner || pat == outer) { q.stateVec.real[i] *= retain; q.stateVec.imag[i] *= retain; }
    }
}
void mixTwoQubitDephasing(Qureg q, int a, int b, qreal prob) {
    qreal retain = 1 - (4 * prob) / 3.0;
    int s = shiftOf(q);
    ll n = q.numAmpsTotal;
#pragma omp parallel for schedule(static)
    for (ll i = 0; i < n; i++)
        if (bitOf(i, a) != bitOf(i, a + s) || bitOf(i, b) != bitOf(i, b + s)) {
            q.stateVec.real[i] *= retain;
            q.stateVec.imag[i] *= retain;
        }
}
void mixDepolarising(Qureg q, int t, qreal prob) {
    qreal lvl = (4 * prob) / 3.0, retain = 1 - lvl;
    if (lvl == 0) return;
    ll inner = 1LL << t, outer = 1LL << (t + shiftOf(q)), tot = inner | outer, n = q.numAmpsTotal;
    qreal *re = q.stateVec.real, *im = q.stateVec.imag;
#pragma omp parallel for schedule(static)
    for (ll i = 0; i < n; i++) {
        ll pat = i & tot;
        if (pat == inner || pat == outer) { re[i] *= retain; im[i] *= retain; }
        else if ((i & tot) == 0) {
            ll j = i | tot;
            qreal ar = (re[i] + re[j]) / 2, ai = (im[i] + im[j]) / 2;
            re[i] = retain * re[i] + lvl * ar; im[i] = retain * im[i] + lvl * ai;
            re[j] = retain * re[j] + lvl * ar; im[j] = retain * im[j] + lvl * ai;
        }
    }
}
void mixDamping(Qureg q, int t, qreal prob) {
    if (prob == 0) return;
    qreal retain = 1 - prob, deph = sqrt(retain);
    ll inner = 1LL << t, outer = 1LL << (t + shiftOf(q)), tot = inner | outer, n = q.numAmpsTotal;
    qreal *re = q.stateVec.real, *im = q.stateVec.imag;
#pragma omp parallel for schedule(static)
    for (ll i = 0; i < n; i++) {
        ll pat = i & tot;
        if (pat == inner || pat == outer) { re[i] *= deph; im[i] *= deph; }
        else if ((i & tot) == 0) {
            ll j = i | tot;
            re[i] = re[i] + prob * re[j]; im[i] = im[i] + prob * im[j];
            re[j] = retain * re[j]; im[j] = retain * im[j];
        }
    }
}
/* Kraus superoperator Σ conj(K) ⊗ K on (t, t+n) (QuEST_common.c:590-640) */
void mixKrausMap(Qureg q, int t, ComplexMatrix2 *ops, int numOps) {
    qreal ur[16] = {0}, ui[16] = {0};
    for (int n = 0; n < numOps; n++)
        for (int i = 0; i < 2; i++) for (int j = 0; j < 2; j++) for (int k = 0; k < 2; k++) for (int l = 0; l < 2; l++) {
            int r = i * 2 + k, c = j * 2 + l;
            ur[r * 4 + c] += ops[n].real[i][j] * ops[n].real[k][l] + ops[n].imag[i][j] * ops[n].imag[k][l];
            ui[r * 4 + c] += ops[n].real[i][j] * ops[n].imag[k][l] - ops[n].imag[i][j] * ops[n].real[k][l];
        }
    int targs[2] = {t, t + shiftOf(q)};
    sv_matrixN(q, 0, targs, 2, ur, ui);
}
/* two-qubit Kraus superoperator on (t1, t2, t1+n, t2+n) (QuEST_common.c:679-700,
 * QuEST.c:1333-1342); superoperator index = row index + 4 * column index */
void mixTwoQubitKrausMap(Qureg q, int t1, int t2, ComplexMatrix4 *ops, int numOps) {
    static qreal ur[256], ui[256];
    memset(ur, 0, sizeof ur);
    memset(ui, 0, sizeof ui);
    for (int n = 0; n < numOps; n++)
        for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) for (int k = 0; k < 4; k++) for (int l = 0; l < 4; l++) {
            int r = i * 4 + k, c = j * 4 + l;
            ur[r * 16 + c] += ops[n].real[i][j] * ops[n].real[k][l] + ops[n].imag[i][j] * ops[n].imag[k][l];
            ui[r * 16 + c] += ops[n].real[i][j] * ops[n].imag[k][l] - ops[n].imag[i][j] * ops[n].real[k][l];
        }
    int targs[4] = {t1, t2, t1 + shiftOf(q), t2 + shiftOf(q)};
    sv_matrixN(q, 0, targs, 4, ur, ui);
}
/* rho -> (1-p) rho + p/15 sum of the 15 non-identity Pauli pairs (QuEST.c:1300-1310);
 * restated as its 16 Kraus operators sqrt(1-p) I, sqrt(p/15) s_a (x) s_b */
void mixTwoQubitDepolarising(Qureg q, int t1, int t2, qreal prob) {
    static const qreal pr[4][2][2] = {{{1, 0}, {0, 1}}, {{0, 1}, {1, 0}}, {{0, 0}, {0, 0}}, {{1, 0}, {0, -1}}};
    static const qreal pi[4][2][2] = {{{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}, {{0, -1}, {1, 0}}, {{0, 0}, {0, 0}}};
    ComplexMatrix4 ops[16];
    memset(ops, 0, sizeof ops);
    for (int a = 0; a < 4; a++)      /* Pauli on t1 (index bit 0) */
        for (int b = 0; b < 4; b++) { /* Pauli on t2 (index bit 1) */
            qreal f = (a == 0 && b == 0) ? sqrt(1 - prob) : sqrt(prob / 15);
            ComplexMatrix4 *m = &ops[b * 4 + a];
            for (int r = 0; r < 4; r++) for (int c = 0; c < 4; c++) {
                qreal ar = pr[a][r & 1][c & 1], ai = pi[a][r & 1][c & 1];
                qreal br = pr[b][r >> 1][c >> 1], bi = pi[b][r >> 1][c >> 1];
                m->real[r][c] = f * (ar * br - ai * bi);
                m->imag[r][c] = f * (ar * bi + ai * br);
            }
        }
    mixTwoQubitKrausMap(q, t1, t2, ops, 16);
}
/* QuEST.c:955-963: element (row, col) of the flat column-major matrix */
Complex getDensityAmp(Qureg q, ll row, ll col) {
    ll i = row + (col << q.numQubitsRepresented);
    Complex z = {q.stateVec.real[i], q.stateVec.imag[i]};
    return z;
}
void mixPauli(Qureg q, int t, qreal px, qreal py, qreal pz) {
    ComplexMatrix2 ops[4];
    memset(ops, 0, sizeof ops);
    qreal f[4] = {sqrt(1 - (px + py + pz)), sqrt(px), sqrt(py), sqrt(pz)};
    ops[0].real[0][0] = f[0]; ops[0].real[1][1] = f[0];
    ops[1].real[0][1] = f[1]; ops[1].real[1][0] = f[1];
    ops[2].imag[0][1] = -f[2]; ops[2].imag[1][0] = f[2];
    ops[3].real[0][0] = f[3]; ops[3].real[1][1] = -f[3];
    mixKrausMap(q, t, ops, 4);
}
void mixDensityMatrix(Qureg c, qreal p, Qureg o) {
    ll n = c.numAmpsTotal;
#pragma omp parallel for schedule(static)
    for (ll i = 0; i < n; i++) {
        c.stateVec.real[i] *= 1 - p;
        c.stateVec.imag[i] *= 1 - p;
        c.stateVec.real[i] += p * o.stateVec.real[i];
        c.stateVec.imag[i] += p * o.stateVec.imag[i];
    }
}

/* threads the OpenMP loops of this library use (reported with the CPU baseline) */
int cpurefNumThreads(void) { return omp_get_max_threads(); }
