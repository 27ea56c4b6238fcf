// Operation lists and the fusion planner.
//
// A qmxCircuit records the operations of one register width: QuantumLab's
// basic operations (EBasicOperation, QL/Circuit/QLGate.h:18-104, same codes)
// and the QuEST gate calls of the older recorders (compactUnitary, unitary,
// rotateZ, controlledNot, multiControlledUnitary, multiQubitUnitary). Every
// operation is validated when it is recorded, with QuEST's messages.
//
// Applied unfused, each operation is the QuEST call that
// QLGate::PerformBasicOperation makes for it (QL/Circuit/QLGate.cpp:652-803):
// noise operations act on density registers only, a controlled collapse
// (EBO_CC) returns its outcome's probability, and the product of those
// probabilities is returned, as the simulators accumulate it
// (QL/Simulator/QLSimulatorVector.cpp:53-59). On density registers every
// operation that is neither noise nor a collapse is followed by the density
// simulator's after-gate noise (QLSimulatorDensityMatrix.cpp:88-153).
//
// Applied fused (state vectors), the planner below has grouped the unitaries
// between collapses into dense blocks of at most k qubits; each block is one
// sweep (qmx::applyMatrixN, or the FP32 MFMA kernel for 5-qubit blocks)
// instead of one sweep per gate, and each collapse is a barrier no block
// crosses. On density registers, fused application is the unfused sequence:
// the library already composes every operation on at most two qubits into
// superoperator sweeps there.
//
// Planner (greedy, dependency-aware): for a gate on qubit set Q, let B be the
// last block after the latest barrier that touches Q. The gate may join B, or
// any later block (those touch none of Q, so the gate commutes past them),
// provided the union of qubits stays <= k. Among those it joins the one with
// the largest overlap, else opens a new block at the end. Block matrices are
// products of the expanded gate matrices, formed on the host in FP64.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#include "QuEST.h"
#include "qmx_internal.h"
#include "quest_mi355x.h"
#include "quest_validation.h"

using qmx::Cplx;

namespace {

enum GateKind { G_NONE = -1, G_MATRIX2 = 0, G_NOT = 1, G_MATRIXN = 2 };

struct Gate {
    int kind = G_NONE;
    int target = -1;            // G_MATRIX2 / G_NOT
    uint64_t ctrlMask = 0;      // all kinds
    std::vector<int> targs;     // G_MATRIXN (targs[0] = LSB)
    std::vector<Cplx> u;        // 4 entries (MATRIX2) or dim*dim (MATRIXN), row-major
    uint64_t qubitMask() const {
        uint64_t m = ctrlMask;
        if (kind == G_MATRIXN)
            for (int t : targs) m |= 1ull << t;
        else
            m |= 1ull << target;
        return m;
    }
};

// codes of the older recorders (QuEST calls), after QuantumLab's codes
enum : int {
    OP_COMPACT_UNITARY = 100,
    OP_UNITARY,
    OP_ROTATE_Z,
    OP_MULTI_CONTROLLED_UNITARY,
    OP_MULTI_QUBIT_UNITARY
};

struct Op {
    int code = 0;
    std::vector<int> qubits;    // as QuantumLab lists them (controls first); targets of a multi-qubit unitary
    int target = -1;            // older recorders: the 2x2 target
    double param = 0;           // angle, probability or outcome
    Complex alpha{}, beta{};    // OP_COMPACT_UNITARY
    ComplexMatrix2 m2{};        // OP_UNITARY, OP_MULTI_CONTROLLED_UNITARY
    ComplexMatrixN mN{};        // OP_MULTI_QUBIT_UNITARY (owned)
    Gate gate;                  // matrix form for fusion; G_NONE: not a unitary
    bool barrier = false;       // a collapse: no block crosses it
};

struct Block {
    uint64_t mask = 0;          // qubits of the block
    std::vector<int> gates;     // indices into the operation list, in order
    std::vector<int> targs;     // ascending qubits (targs[0] = LSB of the block matrix)
    std::vector<Cplx> u;        // dense block matrix, row-major
    void *prepared5 = nullptr;  // device-resident form for the FP32 MFMA kernel
    int barrierOp = -1;         // >= 0: this step is the collapse ops[barrierOp]
};

Cplx cmul(Cplx a, Cplx b) { return Cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
Cplx cadd(Cplx a, Cplx b) { return Cplx{a.re + b.re, a.im + b.im}; }

// apply gate g to a k-qubit column vector v (local bit j <-> qubit qs[j])
void applyLocal(const Gate &g, const std::vector<int> &qs, std::vector<Cplx> &v) {
    const int k = (int)qs.size();
    auto local = [&](int q) {
        for (int j = 0; j < k; ++j)
            if (qs[j] == q) return j;
        return -1;
    };
    uint64_t lctrl = 0;
    for (int q = 0; q < 64; ++q)
        if ((g.ctrlMask >> q) & 1ull) lctrl |= 1ull << local(q);
    const size_t dim = v.size();
    if (g.kind == G_MATRIX2 || g.kind == G_NOT) {
        const int t = local(g.target);
        const size_t bit = 1ull << t;
        for (size_t i = 0; i < dim; ++i) {
            if (i & bit) continue;
            if ((i & lctrl) != lctrl) continue;
            Cplx a = v[i], b = v[i | bit];
            if (g.kind == G_NOT) {
                v[i] = b;
                v[i | bit] = a;
            } else {
                v[i] = cadd(cmul(g.u[0], a), cmul(g.u[1], b));
                v[i | bit] = cadd(cmul(g.u[2], a), cmul(g.u[3], b));
            }
        }
        return;
    }
    const int gk = (int)g.targs.size();
    const int gd = 1 << gk;
    std::vector<int> lt(gk);
    uint64_t tmask = 0;
    for (int j = 0; j < gk; ++j) {
        lt[j] = local(g.targs[j]);
        tmask |= 1ull << lt[j];
    }
    std::vector<Cplx> in(gd), out(gd);
    std::vector<size_t> idx(gd);
    for (size_t base = 0; base < dim; ++base) {
        if (base & tmask) continue;
        if ((base & lctrl) != lctrl) continue;
        for (int c = 0; c < gd; ++c) {
            size_t i = base;
            for (int j = 0; j < gk; ++j)
                if ((c >> j) & 1) i |= 1ull << lt[j];
            idx[c] = i;
            in[c] = v[i];
        }
        for (int r = 0; r < gd; ++r) {
            Cplx s{0, 0};
            for (int c = 0; c < gd; ++c) s = cadd(s, cmul(g.u[(size_t)r * gd + c], in[c]));
            out[r] = s;
        }
        for (int r = 0; r < gd; ++r) v[idx[r]] = out[r];
    }
}

// 2x2 gate forms of QuantumLab's single-qubit operations, as the QuEST call
// each one maps to applies it (rotations: QuEST_common.c:117-124, compact form
// [[a, -conj(b)], [b, conj(a)]])
void setMatrix2(Gate &g, int target, uint64_t ctrl, Cplx u00, Cplx u01, Cplx u10, Cplx u11) {
    g.kind = G_MATRIX2;
    g.target = target;
    g.ctrlMask = ctrl;
    g.u = {u00, u01, u10, u11};
}

void setCompact(Gate &g, int target, uint64_t ctrl, Cplx a, Cplx b) {
    setMatrix2(g, target, ctrl, a, Cplx{-b.re, b.im}, b, Cplx{a.re, -a.im});
}

// the rotation about axis (x, y, z) by `angle` as a compact pair
void setRotation(Gate &g, int target, uint64_t ctrl, double angle, double x, double y, double z) {
    const double c = std::cos(angle / 2.0), s = std::sin(angle / 2.0);
    setCompact(g, target, ctrl, Cplx{c, -s * z}, Cplx{s * y, -s * x});
}

bool basicGate(Op &o) {
    const int *q = o.qubits.data();
    const double t = o.param;
    const double r = 1.0 / std::sqrt(2.0);
    Gate &g = o.gate;
    switch (o.code) {
        case QMX_OP_H: setMatrix2(g, q[0], 0, {r, 0}, {r, 0}, {r, 0}, {-r, 0}); return true;
        case QMX_OP_X:
            g.kind = G_NOT;
            g.target = q[0];
            return true;
        case QMX_OP_Y: setMatrix2(g, q[0], 0, {0, 0}, {0, -1}, {0, 1}, {0, 0}); return true;
        case QMX_OP_Z: setMatrix2(g, q[0], 0, {1, 0}, {0, 0}, {0, 0}, {-1, 0}); return true;
        case QMX_OP_P: setMatrix2(g, q[0], 0, {1, 0}, {0, 0}, {0, 0}, {std::cos(t), std::sin(t)}); return true;
        case QMX_OP_RX: setRotation(g, q[0], 0, t, 1, 0, 0); return true;
        case QMX_OP_RY: setRotation(g, q[0], 0, t, 0, 1, 0); return true;
        case QMX_OP_RZ: setRotation(g, q[0], 0, t, 0, 0, 1); return true;
        case QMX_OP_CX:
            g.kind = G_NOT;
            g.target = q[1];
            g.ctrlMask = 1ull << q[0];
            return true;
        case QMX_OP_CY: setMatrix2(g, q[1], 1ull << q[0], {0, 0}, {0, -1}, {0, 1}, {0, 0}); return true;
        case QMX_OP_CZ: setMatrix2(g, q[1], 1ull << q[0], {1, 0}, {0, 0}, {0, 0}, {-1, 0}); return true;
        case QMX_OP_CP:
            setMatrix2(g, q[1], 1ull << q[0], {1, 0}, {0, 0}, {0, 0}, {std::cos(t), std::sin(t)});
            return true;
        case QMX_OP_CRX: setRotation(g, q[1], 1ull << q[0], t, 1, 0, 0); return true;
        case QMX_OP_CRY: setRotation(g, q[1], 1ull << q[0], t, 0, 1, 0); return true;
        case QMX_OP_CRZ: setRotation(g, q[1], 1ull << q[0], t, 0, 0, 1); return true;
        default: return false;
    }
}

// qubits an operation code takes (0: not a recordable QuantumLab code)
int arity(int code) {
    switch (code) {
        case QMX_OP_H: case QMX_OP_X: case QMX_OP_Y: case QMX_OP_Z: case QMX_OP_P: case QMX_OP_RX:
        case QMX_OP_RY: case QMX_OP_RZ: case QMX_OP_CC: case QMX_OP_NOISE_DAMPING: case QMX_OP_NOISE_DEPHASING:
        case QMX_OP_NOISE_DEPOLARISING: case QMX_OP_NOISE_MIX_PAULI_X: case QMX_OP_NOISE_MIX_PAULI_Y:
        case QMX_OP_NOISE_MIX_PAULI_Z: case QMX_OP_NOISE_MIX_PAULI_ALL:
            return 1;
        case QMX_OP_CX: case QMX_OP_CY: case QMX_OP_CZ: case QMX_OP_CP: case QMX_OP_CRX: case QMX_OP_CRY:
        case QMX_OP_CRZ: case QMX_OP_NOISE_TWO_QUBIT_DEPHASING: case QMX_OP_NOISE_TWO_QUBIT_DEPOLARISING:
            return 2;
        default: return 0;
    }
}

bool isNoise(int code) { return code >= QMX_OP_NOISE_DAMPING && code <= QMX_OP_NOISE_TWO_QUBIT_DEPOLARISING; }

}  // namespace

struct qmxCircuit {
    int numQubits;
    Qureg shape;                // register of this width, for the recorders' validation
    std::vector<Op> ops;
    std::vector<Block> blocks;
    int fusedQubits = 0;
    qmxNoiseAfterGate noise{};
    void releaseBlocks() {
        for (Block &b : blocks) qmx::fused5Free(b.prepared5);
        blocks.clear();
    }
    ~qmxCircuit() {
        releaseBlocks();
        for (Op &o : ops)
            if (o.code == OP_MULTI_QUBIT_UNITARY) destroyComplexMatrixN(o.mN);
    }
};

namespace {

void buildBlockMatrix(const qmxCircuit &c, Block &b) {
    b.targs.clear();
    for (int q = 0; q < 64; ++q)
        if ((b.mask >> q) & 1ull) b.targs.push_back(q);
    const size_t dim = 1ull << b.targs.size();
    b.u.assign(dim * dim, Cplx{0, 0});
    std::vector<Cplx> col(dim);
    for (size_t cidx = 0; cidx < dim; ++cidx) {
        std::fill(col.begin(), col.end(), Cplx{0, 0});
        col[cidx] = Cplx{1, 0};
        for (int gi : b.gates) applyLocal(c.ops[gi].gate, b.targs, col);
        for (size_t r = 0; r < dim; ++r) b.u[r * dim + cidx] = col[r];
    }
}

// the QuEST call QLGate::PerformBasicOperation makes (QLGate.cpp:652-803);
// returns the collapse probability, else 1
double runOp(const Qureg &q, Op &o) {
    const int *qs = o.qubits.data();
    const qreal t = (qreal)o.param;
    const bool dm = q.isDensityMatrix != 0;
    switch (o.code) {
        case QMX_OP_H: hadamard(q, qs[0]); break;
        case QMX_OP_X: pauliX(q, qs[0]); break;
        case QMX_OP_Y: pauliY(q, qs[0]); break;
        case QMX_OP_Z: pauliZ(q, qs[0]); break;
        case QMX_OP_P: phaseShift(q, qs[0], t); break;
        case QMX_OP_RX: rotateX(q, qs[0], t); break;
        case QMX_OP_RY: rotateY(q, qs[0], t); break;
        case QMX_OP_RZ: rotateZ(q, qs[0], t); break;
        case QMX_OP_CX: controlledNot(q, qs[0], qs[1]); break;
        case QMX_OP_CY: controlledPauliY(q, qs[0], qs[1]); break;
        case QMX_OP_CZ: controlledPhaseFlip(q, qs[0], qs[1]); break;
        case QMX_OP_CP: controlledPhaseShift(q, qs[0], qs[1], t); break;
        case QMX_OP_CRX: controlledRotateX(q, qs[0], qs[1], t); break;
        case QMX_OP_CRY: controlledRotateY(q, qs[0], qs[1], t); break;
        case QMX_OP_CRZ: controlledRotateZ(q, qs[0], qs[1], t); break;
        case QMX_OP_CC: return collapseToOutcome(q, qs[0], (int)o.param);
        // noise: density registers only; a state vector ignores it, as the reference does
        case QMX_OP_NOISE_DAMPING: if (dm) mixDamping(q, qs[0], t); break;
        case QMX_OP_NOISE_DEPHASING: if (dm) mixDephasing(q, qs[0], t); break;
        case QMX_OP_NOISE_DEPOLARISING: if (dm) mixDepolarising(q, qs[0], t); break;
        case QMX_OP_NOISE_MIX_PAULI_X: if (dm) mixPauli(q, qs[0], t, 0, 0); break;
        case QMX_OP_NOISE_MIX_PAULI_Y: if (dm) mixPauli(q, qs[0], 0, t, 0); break;
        case QMX_OP_NOISE_MIX_PAULI_Z: if (dm) mixPauli(q, qs[0], 0, 0, t); break;
        case QMX_OP_NOISE_MIX_PAULI_ALL: if (dm) mixPauli(q, qs[0], t, t, t); break;
        case QMX_OP_NOISE_TWO_QUBIT_DEPHASING: if (dm) mixTwoQubitDephasing(q, qs[0], qs[1], t); break;
        case QMX_OP_NOISE_TWO_QUBIT_DEPOLARISING: if (dm) mixTwoQubitDepolarising(q, qs[0], qs[1], t); break;
        case OP_COMPACT_UNITARY: compactUnitary(q, o.target, o.alpha, o.beta); break;
        case OP_UNITARY: unitary(q, o.target, o.m2); break;
        case OP_ROTATE_Z: rotateZ(q, o.target, t); break;
        case OP_MULTI_CONTROLLED_UNITARY:
            multiControlledUnitary(q, o.qubits.data(), (int)o.qubits.size(), o.target, o.m2);
            break;
        case OP_MULTI_QUBIT_UNITARY: multiQubitUnitary(q, o.qubits.data(), (int)o.qubits.size(), o.mN); break;
        default:
            // EBO_Phase and EBO_Composite: PerformBasicOperation dispatches neither
            std::printf("not supported type!\n");
            break;
    }
    return 1.0;
}

// the density simulator's noise after every operation that is neither noise
// nor a collapse, on the qubits that operation lists (QLSimulatorDensityMatrix.cpp:88-153)
void noiseAfterGate(const Qureg &q, const Op &o, const qmxNoiseAfterGate &nz) {
    if (isNoise(o.code) || o.code == QMX_OP_CC) return;
    const qreal tiny = std::numeric_limits<qreal>::min();
    std::vector<int> qs = o.qubits;
    if (o.target >= 0) qs.push_back(o.target);
    const int n = (int)qs.size();
    if (nz.damping > tiny)
        for (int t : qs) mixDamping(q, t, nz.damping);
    if (nz.dephasing > tiny) {
        if (nz.twoQubitDephasing < tiny) {
            for (int t : qs) mixDephasing(q, t, nz.dephasing);
        } else if (n == 1) {
            mixDephasing(q, qs[0], nz.dephasing);
        }
    }
    if (nz.depolarising > tiny) {
        if (nz.twoQubitDepolarising < tiny) {
            for (int t : qs) mixDepolarising(q, t, nz.depolarising);
        } else if (n == 1) {
            mixDepolarising(q, qs[0], nz.depolarising);
        }
    }
    if (nz.mixPauli > tiny)
        for (int t : qs) mixPauli(q, t, nz.mixPauli, nz.mixPauli, nz.mixPauli);
    if (nz.twoQubitDepolarising > tiny && n == 2) mixTwoQubitDepolarising(q, qs[0], qs[1], nz.twoQubitDepolarising);
    if (nz.twoQubitDephasing > tiny && n == 2) mixTwoQubitDephasing(q, qs[0], qs[1], nz.twoQubitDephasing);
}

bool hasNoise(const qmxNoiseAfterGate &nz) {
    return nz.damping > 0 || nz.dephasing > 0 || nz.depolarising > 0 || nz.mixPauli > 0 ||
           nz.twoQubitDephasing > 0 || nz.twoQubitDepolarising > 0;
}

void recordBasic(qmxCircuit *c, int code, const int *qubits, int numQubits, double param, const char *f) {
    const int need = arity(code);
    if (code == QMX_OP_PHASE || code == QMX_OP_COMPOSITE) {
        if (numQubits < 0 || numQubits > 2) qv::raise("An operation lists at most two qubits.", f);
        for (int i = 0; i < numQubits; ++i) qv::target(c->shape, qubits[i], f);
        Op o;  // accepted and not dispatched, as in PerformBasicOperation
        o.code = code;
        o.qubits.assign(qubits, qubits + std::max(0, numQubits));
        o.param = param;
        c->ops.push_back(std::move(o));
        return;
    }
    if (need == 0) qv::raise("Invalid operation code. Must be one of QuantumLab's basic operations (0-26).", f);
    if (numQubits != need)
        qv::raise(need == 1 ? "This operation takes exactly one qubit." : "This operation takes exactly two qubits.",
                  f);
    const Qureg &s = c->shape;
    if (need == 1) {
        qv::target(s, qubits[0], f);
    } else if (isNoise(code)) {
        qv::uniqueTargets(s, qubits[0], qubits[1], f);
    } else {
        qv::controlTarget(s, qubits[0], qubits[1], f);
    }
    switch (code) {
        case QMX_OP_CC: qv::outcome((int)param, f); break;
        case QMX_OP_NOISE_DAMPING: qv::oneQubitDampingProb(param, f); break;
        case QMX_OP_NOISE_DEPHASING: qv::oneQubitDephaseProb(param, f); break;
        case QMX_OP_NOISE_DEPOLARISING: qv::oneQubitDepolProb(param, f); break;
        case QMX_OP_NOISE_MIX_PAULI_X: qv::oneQubitPauliProbs(param, 0, 0, f); break;
        case QMX_OP_NOISE_MIX_PAULI_Y: qv::oneQubitPauliProbs(0, param, 0, f); break;
        case QMX_OP_NOISE_MIX_PAULI_Z: qv::oneQubitPauliProbs(0, 0, param, f); break;
        case QMX_OP_NOISE_MIX_PAULI_ALL: qv::oneQubitPauliProbs(param, param, param, f); break;
        case QMX_OP_NOISE_TWO_QUBIT_DEPHASING: qv::twoQubitDephaseProb(param, f); break;
        case QMX_OP_NOISE_TWO_QUBIT_DEPOLARISING: qv::twoQubitDepolProb(param, f); break;
        default: break;
    }
    Op o;
    o.code = code;
    o.qubits.assign(qubits, qubits + numQubits);
    o.param = param;
    o.barrier = code == QMX_OP_CC;
    basicGate(o);
    c->ops.push_back(std::move(o));
}

void noiseParams(const qmxNoiseAfterGate &nz, const char *f) {
    qv::oneQubitDampingProb(nz.damping, f);
    qv::oneQubitDephaseProb(nz.dephasing, f);
    qv::oneQubitDepolProb(nz.depolarising, f);
    qv::oneQubitPauliProbs(nz.mixPauli, nz.mixPauli, nz.mixPauli, f);
    qv::twoQubitDephaseProb(nz.twoQubitDephasing, f);
    qv::twoQubitDepolProb(nz.twoQubitDepolarising, f);
}

void checkRegister(const qmxCircuit *c, const Qureg &q, const char *f) {
    qv::check(qmx::dev(q) != nullptr && qmx::dev(q)->magic == qmx::kMagic, qv::E_QUREG_NOT_ALLOCATED, f);
    if (q.numQubitsRepresented != c->numQubits)
        qv::raise("The circuit and the register have different numbers of qubits.", f);
}

double applyCircuit(const Qureg &q, qmxCircuit *c, int fused) {
    double prob = 1.0;
    const bool noisy = q.isDensityMatrix && hasNoise(c->noise);
    if (!fused || c->blocks.empty() || q.isDensityMatrix) {
        for (Op &o : c->ops) {
            prob *= runOp(q, o);
            if (noisy) noiseAfterGate(q, o, c->noise);
        }
        return prob;
    }
    qmx::QuregDev *d = qmx::dev(q);
    for (Block &b : c->blocks) {
        if (b.barrierOp >= 0) {
            prob *= runOp(q, c->ops[b.barrierOp]);
            continue;
        }
        if (b.gates.size() == 1) {
            runOp(q, c->ops[b.gates[0]]);  // a lone gate keeps its cheaper kernel
            continue;
        }
        if (d->prec == qmx::FP32 && b.targs.size() == 5 && d->numQubits >= 9) {
            if (!b.prepared5) b.prepared5 = qmx::fused5Prepare(b.targs.data(), b.u.data());
            qmx::fused5Apply(d, b.prepared5);
            continue;
        }
        qmx::applyMatrixN(d, 0, b.targs.data(), (int)b.targs.size(), b.u.data());
    }
    return prob;
}

// Fuse the gates of one segment (no barrier inside) into blocks of at most
// maxQubits qubits. Gates on disjoint qubits commute, so any order that keeps
// each qubit's gates in sequence is the same circuit. Each step takes the
// window of maxQubits consecutive qubits (or a ready gate's own qubits when
// they span more) that absorbs the most gates: from the window's qubits, the
// next pending gate of a qubit joins while it is next on all of its qubits
// and the block stays within maxQubits. On the 32-qubit brick periods this
// makes 18-19 steps per period where placing each gate, in list order, into the
// last block it fits made 26.
void fuseSegment(qmxCircuit *c, const std::vector<int> &seg, int maxQubits) {
    if (seg.empty()) return;
    std::vector<std::vector<int>> per(64);
    std::vector<uint64_t> maskOf(c->ops.size(), 0);
    for (int gi : seg) {
        maskOf[gi] = c->ops[gi].gate.qubitMask();
        for (uint64_t m = maskOf[gi]; m; m &= m - 1) per[__builtin_ctzll(m)].push_back(gi);
    }
    std::vector<size_t> ptr(64, 0);
    auto nextOn = [&](const std::vector<size_t> &p, int x) { return p[x] < per[x].size() ? per[x][p[x]] : -1; };
    auto readyIn = [&](const std::vector<size_t> &p, int gi) {
        for (uint64_t m = maskOf[gi]; m; m &= m - 1)
            if (nextOn(p, __builtin_ctzll(m)) != gi) return false;
        return true;
    };
    const uint64_t all = c->numQubits >= 64 ? ~0ull : (1ull << c->numQubits) - 1ull;
    size_t remaining = seg.size();
    std::vector<uint64_t> seeds;
    while (remaining) {
        seeds.clear();
        int firstReady = -1;
        for (int x = 0; x < 64; ++x) {
            const int gi = nextOn(ptr, x);
            if (gi < 0 || !readyIn(ptr, gi)) continue;
            if (firstReady < 0 || gi < firstReady) firstReady = gi;
            const uint64_t m = maskOf[gi];
            const int lo = __builtin_ctzll(m), hi = 63 - __builtin_clzll(m);
            if (hi - lo + 1 <= maxQubits) {
                for (int k = std::max(0, hi - maxQubits + 1); k <= lo; ++k) {
                    const uint64_t w = (maxQubits >= 64 ? ~0ull : ((1ull << maxQubits) - 1ull) << k) & all;
                    seeds.push_back(w);
                }
            } else if (__builtin_popcountll(m) <= maxQubits) {
                seeds.push_back(m);
            }
        }
        std::sort(seeds.begin(), seeds.end());
        seeds.erase(std::unique(seeds.begin(), seeds.end()), seeds.end());
        Block best;
        std::vector<size_t> bestPtr;
        for (uint64_t w : seeds) {
            std::vector<size_t> p = ptr;
            Block b;
            for (bool progress = true; progress;) {
                progress = false;
                for (uint64_t wm = w; wm; wm &= wm - 1) {
                    const int x = __builtin_ctzll(wm);
                    for (int gi = nextOn(p, x); gi >= 0; gi = nextOn(p, x)) {
                        if (!readyIn(p, gi) || __builtin_popcountll(b.mask | maskOf[gi]) > maxQubits) break;
                        b.mask |= maskOf[gi];
                        b.gates.push_back(gi);
                        for (uint64_t m = maskOf[gi]; m; m &= m - 1) ++p[__builtin_ctzll(m)];
                        progress = true;
                    }
                }
            }
            if (b.gates.size() > best.gates.size()) {
                best = std::move(b);
                bestPtr = std::move(p);
            }
        }
        if (best.gates.empty()) {
            // a ready gate on more than maxQubits qubits: a block of its own
            best.mask = maskOf[firstReady];
            best.gates.push_back(firstReady);
            bestPtr = ptr;
            for (uint64_t m = best.mask; m; m &= m - 1) ++bestPtr[__builtin_ctzll(m)];
        }
        remaining -= best.gates.size();
        ptr = std::move(bestPtr);
        c->blocks.push_back(std::move(best));
    }
}

int fuseCircuit(qmxCircuit *c, int maxQubits) {
    c->releaseBlocks();
    c->fusedQubits = maxQubits;
    std::vector<int> seg;  // gates since the latest barrier
    for (int gi = 0; gi < (int)c->ops.size(); ++gi) {
        const Op &o = c->ops[gi];
        if (o.barrier) {
            fuseSegment(c, seg, maxQubits);
            seg.clear();
            Block nb;
            nb.barrierOp = gi;
            c->blocks.push_back(std::move(nb));
            continue;
        }
        if (o.gate.kind == G_NONE) continue;  // noise or unsupported: nothing on a state vector
        seg.push_back(gi);
    }
    fuseSegment(c, seg, maxQubits);
    // The MFMA kernel applies exactly-5-qubit blocks: pad smaller multi-gate
    // blocks with identity on extra qubits, taken from the top so the low
    // qubits stay non-targets (consecutive groups = consecutive addresses).
    if (maxQubits == 5 && c->numQubits >= 9) {
        for (Block &b : c->blocks) {
            if (b.barrierOp >= 0 || b.gates.size() < 2) continue;
            for (int qb = c->numQubits - 1; qb >= 0 && __builtin_popcountll(b.mask) < 5; --qb)
                b.mask |= 1ull << qb;
        }
    }
    for (Block &b : c->blocks)
        if (b.barrierOp < 0) buildBlockMatrix(*c, b);
    return (int)c->blocks.size();
}

}  // namespace

#define CIRCUIT_TRY try {
#define CIRCUIT_CATCH(ret)              \
    }                                   \
    catch (const qmx::InputError &) {   \
        return ret;                     \
    }

extern "C" {

qmxCircuit *qmxCircuitCreate(int numQubits) {
    CIRCUIT_TRY
    qv::numQubitsInQureg(numQubits, __func__);
    auto *c = new qmxCircuit;
    c->numQubits = numQubits;
    std::memset(&c->shape, 0, sizeof c->shape);
    c->shape.numQubitsRepresented = numQubits;
    c->shape.numQubitsInStateVec = numQubits;
    c->shape.numAmpsPerChunk = c->shape.numAmpsTotal = numQubits < 63 ? 1LL << numQubits : 0;
    return c;
    CIRCUIT_CATCH(nullptr)
}

void qmxCircuitDestroy(qmxCircuit *c) { delete c; }

int qmxCircuitNumGates(const qmxCircuit *c) { return (int)c->ops.size(); }

void qmxCircuitAddOperation(qmxCircuit *c, int operation, const int *qubits, int numQubits, qreal param) {
    CIRCUIT_TRY
    recordBasic(c, operation, qubits, numQubits, (double)param, __func__);
    CIRCUIT_CATCH()
}

void qmxCircuitSetNoiseAfterGate(qmxCircuit *c, qmxNoiseAfterGate noise) {
    CIRCUIT_TRY
    noiseParams(noise, __func__);
    c->noise = noise;
    CIRCUIT_CATCH()
}

void qmxCircuitCompactUnitary(qmxCircuit *c, int target, Complex a, Complex b) {
    CIRCUIT_TRY
    qv::target(c->shape, target, __func__);
    qv::unitaryComplexPair(a, b, __func__);
    Op o;
    o.code = OP_COMPACT_UNITARY;
    o.target = target;
    o.alpha = a;
    o.beta = b;
    setCompact(o.gate, target, 0, Cplx{a.real, a.imag}, Cplx{b.real, b.imag});
    c->ops.push_back(std::move(o));
    CIRCUIT_CATCH()
}

void qmxCircuitUnitary(qmxCircuit *c, int target, ComplexMatrix2 u) {
    CIRCUIT_TRY
    qv::target(c->shape, target, __func__);
    qv::oneQubitUnitary(u, __func__);
    Op o;
    o.code = OP_UNITARY;
    o.target = target;
    o.m2 = u;
    setMatrix2(o.gate, target, 0, Cplx{u.real[0][0], u.imag[0][0]}, Cplx{u.real[0][1], u.imag[0][1]},
               Cplx{u.real[1][0], u.imag[1][0]}, Cplx{u.real[1][1], u.imag[1][1]});
    c->ops.push_back(std::move(o));
    CIRCUIT_CATCH()
}

// rotateZ as QuEST applies it: compactUnitary with alpha = e^{-iθ/2}, beta = 0
void qmxCircuitRotateZ(qmxCircuit *c, int target, qreal angle) {
    CIRCUIT_TRY
    qv::target(c->shape, target, __func__);
    Op o;
    o.code = OP_ROTATE_Z;
    o.target = target;
    o.param = angle;
    setRotation(o.gate, target, 0, angle, 0, 0, 1);
    c->ops.push_back(std::move(o));
    CIRCUIT_CATCH()
}

void qmxCircuitControlledNot(qmxCircuit *c, int control, int target) {
    CIRCUIT_TRY
    const int qs[2] = {control, target};
    recordBasic(c, QMX_OP_CX, qs, 2, 0.0, __func__);
    CIRCUIT_CATCH()
}

void qmxCircuitMultiControlledUnitary(qmxCircuit *c, int *controls, int numControls, int target, ComplexMatrix2 u) {
    CIRCUIT_TRY
    qv::multiControlsTarget(c->shape, controls, numControls, target, __func__);
    qv::oneQubitUnitary(u, __func__);
    Op o;
    o.code = OP_MULTI_CONTROLLED_UNITARY;
    o.qubits.assign(controls, controls + numControls);
    o.target = target;
    o.m2 = u;
    uint64_t m = 0;
    for (int i = 0; i < numControls; ++i) m |= 1ull << controls[i];
    setMatrix2(o.gate, target, m, Cplx{u.real[0][0], u.imag[0][0]}, Cplx{u.real[0][1], u.imag[0][1]},
               Cplx{u.real[1][0], u.imag[1][0]}, Cplx{u.real[1][1], u.imag[1][1]});
    c->ops.push_back(std::move(o));
    CIRCUIT_CATCH()
}

void qmxCircuitMultiQubitUnitary(qmxCircuit *c, int *targs, int numTargs, const qreal *re, const qreal *im) {
    CIRCUIT_TRY
    qv::multiTargets(c->shape, targs, numTargs, __func__);
    qv::check(numTargs <= 10, qv::E_INVALID_NUM_TARGETS, __func__);
    Op o;
    o.code = OP_MULTI_QUBIT_UNITARY;
    o.qubits.assign(targs, targs + numTargs);
    const size_t dim = 1ull << numTargs;
    o.mN = createComplexMatrixN(numTargs);
    for (size_t r = 0; r < dim; ++r)
        for (size_t col = 0; col < dim; ++col) {
            o.mN.real[r][col] = re[r * dim + col];
            o.mN.imag[r][col] = im[r * dim + col];
        }
    try {
        qv::multiQubitUnitary(c->shape, o.mN, numTargs, __func__);
    } catch (...) {  // also a caller's throwing error hook
        destroyComplexMatrixN(o.mN);
        throw;
    }
    o.gate.kind = G_MATRIXN;
    o.gate.targs = o.qubits;
    o.gate.u.resize(dim * dim);
    for (size_t e = 0; e < dim * dim; ++e) o.gate.u[e] = Cplx{re[e], im[e]};
    c->ops.push_back(std::move(o));
    CIRCUIT_CATCH()
}

int qmxCircuitFuse(qmxCircuit *c, int maxQubits) {
    CIRCUIT_TRY
    if (maxQubits < 1 || maxQubits > 10)
        qv::raise("Invalid block size. Fused blocks span at least 1 and at most 10 qubits.", __func__);
    return fuseCircuit(c, maxQubits);
    CIRCUIT_CATCH(0)
}

int qmxCircuitStep(const qmxCircuit *c, int step, int *targs, qreal *re, qreal *im) {
    if (step < 0 || step >= (int)c->blocks.size()) return -1;
    const Block &b = c->blocks[step];
    if (b.barrierOp >= 0) return 0;
    const int k = (int)b.targs.size();
    const size_t dim = 1ull << k;
    if (targs)
        for (int j = 0; j < k; ++j) targs[j] = b.targs[j];
    if (re && im)
        for (size_t e = 0; e < dim * dim; ++e) {
            re[e] = (qreal)b.u[e].re;
            im[e] = (qreal)b.u[e].im;
        }
    return k;
}

int qmxCircuitNumBlocks(const qmxCircuit *c) {
    int n = 0;
    for (const Block &b : c->blocks) n += b.barrierOp < 0;
    return n;
}

qreal qmxCircuitApply(Qureg qureg, qmxCircuit *c, int fused) {
    CIRCUIT_TRY
    checkRegister(c, qureg, __func__);
    return (qreal)applyCircuit(qureg, c, fused);
    CIRCUIT_CATCH(0)
}

qreal qmxApplyOperations(Qureg qureg, const qmxBasicOperation *ops, int numOps, const qmxNoiseAfterGate *noise,
                         int fuseQubits) {
    CIRCUIT_TRY
    qmxCircuit c;
    c.numQubits = qureg.numQubitsRepresented;
    std::memset(&c.shape, 0, sizeof c.shape);
    c.shape.numQubitsRepresented = c.numQubits;
    c.shape.numQubitsInStateVec = c.numQubits;
    c.shape.numAmpsPerChunk = c.shape.numAmpsTotal = c.numQubits < 63 ? 1LL << c.numQubits : 0;
    checkRegister(&c, qureg, __func__);
    if (numOps < 0) qv::raise("Invalid number of operations. Must be >=0.", __func__);
    if (noise) {
        noiseParams(*noise, __func__);
        c.noise = *noise;
    }
    for (int i = 0; i < numOps; ++i)
        recordBasic(&c, ops[i].operation, ops[i].qubits, ops[i].numQubits, (double)ops[i].param, __func__);
    if (fuseQubits != 0) {
        if (fuseQubits < 1 || fuseQubits > 10)
            qv::raise("Invalid block size. Fused blocks span at least 1 and at most 10 qubits.", __func__);
        if (!qureg.isDensityMatrix) fuseCircuit(&c, fuseQubits);
    }
    return (qreal)applyCircuit(qureg, &c, fuseQubits != 0);
    CIRCUIT_CATCH(0)
}

qreal qmxMaxAbsDiff(Qureg a, Qureg b) { return qmx::maxAbsDiff(qmx::dev(a), qmx::dev(b)); }

}  // extern "C"
