// Gate lists and the fusion planner.
//
// A qmxCircuit records the gates of one brick-circuit-like workload (the ops
// QLGate::PerformBasicOperation lowers to, QL/Circuit/QLGate.cpp:652-803, plus
// multi-qubit unitaries). Applied unfused, each gate is one kernel exactly as
// the QuEST API call would launch it. Applied fused, the planner below has
// grouped the gates into dense blocks of at most k qubits; each block is one
// sweep (qmx::applyMatrixN) instead of one sweep per gate.
//
// Planner (greedy, dependency-aware): for a gate on qubit set Q, let B be the
// last block that touches Q. The gate may join B, or any later block (those
// touch none of Q, so the gate commutes past them), provided the union of
// qubits stays <= k. Among those it joins the one with the largest overlap,
// else opens a new block at the end. Block matrices are products of the
// expanded gate matrices, formed on the host in FP64.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "QuEST.h"
#include "qmx_internal.h"
#include "quest_mi355x.h"

using qmx::Cplx;

namespace {

enum GateKind { G_MATRIX2 = 0, G_NOT = 1, G_MATRIXN = 2 };

struct Gate {
    int kind;
    int target;                 // G_MATRIX2 / G_NOT
    uint64_t ctrlMask;          // all kinds
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

struct Block {
    uint64_t mask = 0;          // qubits of the block
    std::vector<int> gates;     // indices into the gate list, in order
    std::vector<int> targs;     // ascending qubits (targs[0] = LSB of the block matrix)
    std::vector<Cplx> u;        // dense block matrix, row-major
    void *prepared5 = nullptr;  // device-resident form for the FP32 MFMA kernel
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

}  // namespace

struct qmxCircuit {
    int numQubits;
    std::vector<Gate> gates;
    std::vector<Block> blocks;
    int fusedQubits = 0;
    void releaseBlocks() {
        for (Block &b : blocks) qmx::fused5Free(b.prepared5);
        blocks.clear();
    }
    ~qmxCircuit() { releaseBlocks(); }
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
        for (int gi : b.gates) applyLocal(c.gates[gi], b.targs, col);
        for (size_t r = 0; r < dim; ++r) b.u[r * dim + cidx] = col[r];
    }
}

void applyGate(qmx::QuregDev *d, const Gate &g) {
    switch (g.kind) {
        case G_MATRIX2: qmx::applyMatrix2(d, g.target, g.ctrlMask, 0, g.u.data()); break;
        case G_NOT: qmx::applyNot(d, g.ctrlMask, 1ull << g.target); break;
        default: qmx::applyMatrixN(d, g.ctrlMask, g.targs.data(), (int)g.targs.size(), g.u.data()); break;
    }
}

void addMatrix2(qmxCircuit *c, int target, uint64_t ctrl, Cplx u00, Cplx u01, Cplx u10, Cplx u11) {
    Gate g;
    g.kind = G_MATRIX2;
    g.target = target;
    g.ctrlMask = ctrl;
    g.u = {u00, u01, u10, u11};
    c->gates.push_back(std::move(g));
}

}  // namespace

extern "C" {

qmxCircuit *qmxCircuitCreate(int numQubits) {
    auto *c = new qmxCircuit;
    c->numQubits = numQubits;
    return c;
}

void qmxCircuitDestroy(qmxCircuit *c) { delete c; }

int qmxCircuitNumGates(const qmxCircuit *c) { return (int)c->gates.size(); }

void qmxCircuitCompactUnitary(qmxCircuit *c, int target, Complex a, Complex b) {
    addMatrix2(c, target, 0, Cplx{a.real, a.imag}, Cplx{-b.real, b.imag}, Cplx{b.real, b.imag},
               Cplx{a.real, -a.imag});
}

void qmxCircuitUnitary(qmxCircuit *c, int target, ComplexMatrix2 u) {
    addMatrix2(c, target, 0, Cplx{u.real[0][0], u.imag[0][0]}, Cplx{u.real[0][1], u.imag[0][1]},
               Cplx{u.real[1][0], u.imag[1][0]}, Cplx{u.real[1][1], u.imag[1][1]});
}

// rotateZ as QuEST applies it: compactUnitary with alpha = e^{-iθ/2}, beta = 0
void qmxCircuitRotateZ(qmxCircuit *c, int target, qreal angle) {
    Complex a{std::cos(angle / 2.0), -std::sin(angle / 2.0)}, b{0, 0};
    qmxCircuitCompactUnitary(c, target, a, b);
}

void qmxCircuitControlledNot(qmxCircuit *c, int control, int target) {
    Gate g;
    g.kind = G_NOT;
    g.target = target;
    g.ctrlMask = 1ull << control;
    c->gates.push_back(std::move(g));
}

void qmxCircuitMultiControlledUnitary(qmxCircuit *c, int *controls, int numControls, int target, ComplexMatrix2 u) {
    uint64_t m = 0;
    for (int i = 0; i < numControls; ++i) m |= 1ull << controls[i];
    addMatrix2(c, target, m, Cplx{u.real[0][0], u.imag[0][0]}, Cplx{u.real[0][1], u.imag[0][1]},
               Cplx{u.real[1][0], u.imag[1][0]}, Cplx{u.real[1][1], u.imag[1][1]});
}

void qmxCircuitMultiQubitUnitary(qmxCircuit *c, int *targs, int numTargs, const qreal *re, const qreal *im) {
    Gate g;
    g.kind = G_MATRIXN;
    g.target = -1;
    g.ctrlMask = 0;
    g.targs.assign(targs, targs + numTargs);
    size_t dim = 1ull << numTargs;
    g.u.resize(dim * dim);
    for (size_t e = 0; e < dim * dim; ++e) g.u[e] = Cplx{re[e], im[e]};
    c->gates.push_back(std::move(g));
}

int qmxCircuitFuse(qmxCircuit *c, int maxQubits) {
    c->releaseBlocks();
    c->fusedQubits = maxQubits;
    for (int gi = 0; gi < (int)c->gates.size(); ++gi) {
        const uint64_t q = c->gates[gi].qubitMask();
        int last = -1;
        for (int b = (int)c->blocks.size() - 1; b >= 0; --b)
            if (c->blocks[b].mask & q) {
                last = b;
                break;
            }
        int best = -1, bestOverlap = -1;
        for (int b = std::max(0, last); b < (int)c->blocks.size(); ++b) {
            uint64_t u = c->blocks[b].mask | q;
            if (__builtin_popcountll(u) > maxQubits) continue;
            int ov = __builtin_popcountll(c->blocks[b].mask & q);
            if (ov > bestOverlap) {
                best = b;
                bestOverlap = ov;
            }
        }
        if (best < 0 || (last >= 0 && best < last)) {
            Block nb;
            nb.mask = q;
            nb.gates.push_back(gi);
            c->blocks.push_back(std::move(nb));
        } else {
            c->blocks[best].mask |= q;
            c->blocks[best].gates.push_back(gi);
        }
    }
    // The MFMA kernel applies exactly-5-qubit blocks: pad smaller multi-gate
    // blocks with identity on extra qubits, taken from the top so the low
    // qubits stay non-targets (consecutive groups = consecutive addresses).
    if (maxQubits == 5 && c->numQubits >= 9) {
        for (Block &b : c->blocks) {
            if (b.gates.size() < 2) continue;
            for (int qb = c->numQubits - 1; qb >= 0 && __builtin_popcountll(b.mask) < 5; --qb)
                b.mask |= 1ull << qb;
        }
    }
    for (Block &b : c->blocks) buildBlockMatrix(*c, b);
    return (int)c->blocks.size();
}

int qmxCircuitNumBlocks(const qmxCircuit *c) { return (int)c->blocks.size(); }

void qmxCircuitApply(Qureg qureg, qmxCircuit *c, int fused) {
    qmx::QuregDev *d = qmx::dev(qureg);
    if (!fused || c->blocks.empty()) {
        for (const Gate &g : c->gates) applyGate(d, g);
        return;
    }
    for (Block &b : c->blocks) {
        if (b.gates.size() == 1) {
            applyGate(d, c->gates[b.gates[0]]);  // a lone gate keeps its cheaper kernel
            continue;
        }
        if (d->prec == qmx::FP32 && b.targs.size() == 5 && d->numQubits >= 9) {
            if (!b.prepared5) b.prepared5 = qmx::fused5Prepare(b.targs.data(), b.u.data());
            qmx::fused5Apply(d, b.prepared5);
            continue;
        }
        qmx::applyMatrixN(d, 0, b.targs.data(), (int)b.targs.size(), b.u.data());
    }
}

qreal qmxMaxAbsDiff(Qureg a, Qureg b) { return qmx::maxAbsDiff(qmx::dev(a), qmx::dev(b)); }

}  // extern "C"
