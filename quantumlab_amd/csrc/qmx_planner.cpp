// Learned Trotter planner, host side: features of a Pauli sum, the decision
// (order, repetitions), the plain step-doubling method it falls back to, and
// the planner's text file format. The network itself runs on the device
// (qmx_mlp.hip).
//
// Target the network learns. For order o the infidelity of o's formula at r
// repetitions behaves as 1 - F ~ A_o r^(-2o), and A_o of a local Hamiltonian
// grows with the register size n (roughly as n for a sum of independent local
// errors), so the cheapest log2 r*_o = log2(A_o / eps) / (2o) drifts slowly
// with log2(n). The planner keeps that drift as a size curve — the mean
// log2 r*_o per register size, piecewise linear in log2(n), measured on the
// training sizes and at the deployment size — and the network, fed per-qubit
// (intensive) features, predicts the deviation y_o = log2 r*_o - curve_o(n).
// That lets labels computed on small registers serve large ones.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "qmx_planner.h"
#include "quest_mi355x.h"

namespace {

constexpr float kLogFloor = -60.f;

struct Term {
    uint64_t x, z;
    int weight;
    double c;
};

bool anticommute(uint64_t x1, uint64_t z1, uint64_t x2, uint64_t z2) {
    return ((__builtin_popcountll(x1 & z2) + __builtin_popcountll(z1 & x2)) & 1) != 0;
}

float log2f_floor(double v) { return v > 0 ? std::max(kLogFloor, (float)std::log2(v)) : kLogFloor; }

int stagesOf(int order) {
    if (order == 1) return 1;
    int s = 2;
    for (int o = 4; o <= order; o += 2) s *= 5;
    return s;
}

}  // namespace

extern "C" {

// QuEST-style PauliHamil -> the planner's 16 inputs (see the header for the list)
int qmxTrotterFeatures(PauliHamil h, qreal time, float *out) {
    if (h.numQubits < 1 || h.numQubits > 64 || h.numSumTerms < 1 || !h.pauliCodes || !h.termCoeffs || !out) {
        std::fprintf(stderr, "libquest_mi355x: qmxTrotterFeatures: invalid Hamiltonian\n");
        return 0;
    }
    const int n = h.numQubits;
    std::vector<Term> terms;
    for (int t = 0; t < h.numSumTerms; ++t) {
        Term tm{0, 0, 0, h.termCoeffs[t]};
        for (int q = 0; q < n; ++q) {
            int c;  // the int storage: a C caller may have written any value
            std::memcpy(&c, h.pauliCodes + (size_t)t * n + q, sizeof c);
            if (c < PAULI_I || c > PAULI_Z) {
                std::fprintf(stderr, "libquest_mi355x: qmxTrotterFeatures: invalid Pauli code %d\n", c);
                return 0;
            }
            const uint64_t b = 1ull << q;
            if (c == PAULI_X) tm.x |= b;
            if (c == PAULI_Y) { tm.x |= b; tm.z |= b; }
            if (c == PAULI_Z) tm.z |= b;
            tm.weight += c != PAULI_I;
        }
        if (tm.x | tm.z) terms.push_back(tm);  // identity strings are no-ops
    }
    const int T = (int)terms.size();
    double lam = 0, mx = 0, l2 = 0, lamHigh = 0, wsum = 0;
    int wmax = 0;
    for (const Term &t : terms) {
        const double a = std::fabs(t.c);
        lam += a;
        mx = std::max(mx, a);
        l2 += t.c * t.c;
        if (t.weight >= 3) lamHigh += a;
        wsum += t.weight;
        wmax = std::max(wmax, t.weight);
    }
    // C1 = sum_{j<k} ||[H_j, H_k]|| = sum over anticommuting pairs of 2 |c_j c_k|;
    // C2 = sum_{j<k} sum_l ||[H_l, [H_j, H_k]]||: [P_j, P_k] = 2 P_j P_k, and P_l
    // fails to commute with P_j P_k exactly when it anticommutes with it
    double C1 = 0, C2 = 0;
    long long anti = 0;
    for (int j = 0; j < T; ++j)
        for (int k = j + 1; k < T; ++k) {
            if (!anticommute(terms[j].x, terms[j].z, terms[k].x, terms[k].z)) continue;
            ++anti;
            const double cjk = std::fabs(terms[j].c * terms[k].c);
            C1 += 2 * cjk;
            const uint64_t px = terms[j].x ^ terms[k].x, pz = terms[j].z ^ terms[k].z;
            for (int l = 0; l < T; ++l)
                if (anticommute(terms[l].x, terms[l].z, px, pz)) C2 += 4 * cjk * std::fabs(terms[l].c);
        }
    const double t = time;
    out[0] = (float)t;
    out[1] = log2f_floor(t);
    out[2] = (float)T / n;
    out[3] = (float)(lam / n);
    out[4] = log2f_floor(lam / n);
    out[5] = (float)mx;
    out[6] = (float)std::sqrt(l2 / n);
    out[7] = log2f_floor(C1 / n);
    out[8] = log2f_floor(C2 / n);
    out[9] = log2f_floor(lam * t / n);
    out[10] = log2f_floor(C1 * t * t / n);
    out[11] = log2f_floor(C2 * t * t * t / n);
    out[12] = T ? (float)(2.0 * anti / T) : 0.f;
    out[13] = T ? (float)(wsum / T) : 0.f;
    out[14] = (float)wmax;
    out[15] = lam > 0 ? (float)(lamHigh / lam) : 0.f;
    return 1;
}

int qmxTrotterStages(int order) { return (order == 1 || (order >= 2 && order % 2 == 0)) ? stagesOf(order) : 0; }

// decision from the network's two outputs: repetitions per order and the
// cheaper of orders 2 and 4 by Pauli exponentials (stages x repetitions)
int qmxPlannerChoose(qmxPlanner *p, PauliHamil hamil, qreal time, int *order, int *reps) {
    float f[qmx::kFeatures], y[qmx::kOutputs];
    if (!qmxTrotterFeatures(hamil, time, f)) return 0;
    qmxPlannerPredict(p, f, 1, y);
    const int orders[2] = {2, 4};
    long long bestCost = -1;
    for (int o = 0; o < 2; ++o) {
        const double l2r = y[o] + qmx::sizeCurve(p->host, o, std::log2((double)hamil.numQubits)) + p->host.margin;
        const double r = std::ceil(std::exp2(std::min(30.0, l2r)) - 1e-9);
        const int ri = (int)std::max(1.0, r);
        const long long cost = (long long)stagesOf(orders[o]) * ri;
        if (bestCost < 0 || cost < bestCost) {
            bestCost = cost;
            *order = orders[o];
            *reps = ri;
        }
    }
    return 1;
}

// plain step doubling: r = 1, 2, 4, ...; each rung evolves a copy of the
// initial state; stop when 1 - |<psi_{r/2}|psi_r>|^2 / (|psi_{r/2}|^2 |psi_r|^2)
// <= eps and keep psi_r.
// Returns 1 when eps was met (qureg holds psi_r), 0 if maxReps ran out
// (qureg holds the finest rung).
int qmxApplyTrotterStepDoubling(Qureg qureg, PauliHamil hamil, qreal time, int order, qreal eps, int maxReps,
                                int *repsOut) {
    if (qureg.isDensityMatrix || maxReps < 2 || !(eps > 0)) {
        std::fprintf(stderr, "libquest_mi355x: qmxApplyTrotterStepDoubling: state vectors only, maxReps >= 2, eps > 0\n");
        return 0;
    }
    QuESTEnv env{};
    const int prec = qmxQuregPrecision(qureg);
    Qureg init = createQuregEx(qureg.numQubitsRepresented, env, prec);
    Qureg prev = createQuregEx(qureg.numQubitsRepresented, env, prec);
    cloneQureg(init, qureg);
    applyTrotterCircuit(qureg, hamil, time, order, 1);
    int r = 1, ok = 0;
    while (r < maxReps) {
        cloneQureg(prev, qureg);
        r *= 2;
        cloneQureg(qureg, init);
        applyTrotterCircuit(qureg, hamil, time, order, r);
        // normalised: FP32 registers drift in norm by ~1e-5 over long circuits
        const Complex ip = calcInnerProduct(prev, qureg);
        const double f = (ip.real * ip.real + ip.imag * ip.imag) / (calcTotalProb(prev) * calcTotalProb(qureg));
        if (1.0 - f <= eps) {
            ok = 1;
            break;
        }
    }
    destroyQureg(init, env);
    destroyQureg(prev, env);
    if (repsOut) *repsOut = r;
    return ok;
}

// one applyTrotterCircuit at the planner's (order, repetitions)
int qmxApplyTrotterLearned(Qureg qureg, PauliHamil hamil, qreal time, qmxPlanner *p, int *orderOut, int *repsOut) {
    int order = 2, reps = 1;
    if (!qmxPlannerChoose(p, hamil, time, &order, &reps)) return 0;
    applyTrotterCircuit(qureg, hamil, time, order, reps);
    if (orderOut) *orderOut = order;
    if (repsOut) *repsOut = reps;
    return 1;
}

int qmxPlannerSetSizeCurve(qmxPlanner *p, int knots, const float *log2n, const float *log2reps) {
    if (knots < 2 || knots > qmx::kMaxCurve) {
        std::fprintf(stderr, "libquest_mi355x: qmxPlannerSetSizeCurve: 2..%d knots\n", qmx::kMaxCurve);
        return 0;
    }
    for (int k = 1; k < knots; ++k)
        if (!(log2n[k] > log2n[k - 1])) {
            std::fprintf(stderr, "libquest_mi355x: qmxPlannerSetSizeCurve: knots must ascend\n");
            return 0;
        }
    p->host.knots = knots;
    for (int k = 0; k < knots; ++k) {
        p->host.knotLog2n[k] = log2n[k];
        for (int o = 0; o < qmx::kOutputs; ++o) p->host.knotLog2r[k][o] = log2reps[k * qmx::kOutputs + o];
    }
    return 1;
}

int qmxPlannerGetSizeCurve(const qmxPlanner *p, float *log2n, float *log2reps) {
    for (int k = 0; k < p->host.knots; ++k) {
        if (log2n) log2n[k] = p->host.knotLog2n[k];
        for (int o = 0; o < qmx::kOutputs && log2reps; ++o) log2reps[k * qmx::kOutputs + o] = p->host.knotLog2r[k][o];
    }
    return p->host.knots;
}

double qmxPlannerSizeCurve(const qmxPlanner *p, int order, int numQubits) {
    return qmx::sizeCurve(p->host, order == 4 ? 1 : 0, std::log2((double)numQubits));
}

// text format: a header line, then the parameters, the normalisation, the
// margin and the size curve (knot count, then log2 n and two log2 r per knot)
int qmxPlannerSave(const qmxPlanner *p, const char *path) {
    std::vector<float> w(qmx::kNumParams);
    qmxPlannerGetParams(p, w.data());
    FILE *f = std::fopen(path, "w");
    if (!f) return 0;
    std::fprintf(f, "qmx-trotter-planner 1 %d %d\n", qmx::kNumParams, qmx::kFeatures);
    for (float v : w) std::fprintf(f, "%.9g\n", v);
    for (int i = 0; i < qmx::kFeatures; ++i) std::fprintf(f, "%.9g %.9g\n", p->host.mean[i], p->host.scale[i]);
    std::fprintf(f, "%.9g\n", p->host.margin);
    std::fprintf(f, "%d\n", p->host.knots);
    for (int k = 0; k < p->host.knots; ++k)
        std::fprintf(f, "%.9g %.9g %.9g\n", p->host.knotLog2n[k], p->host.knotLog2r[k][0], p->host.knotLog2r[k][1]);
    return std::fclose(f) == 0;
}

qmxPlanner *qmxPlannerLoad(const char *path) {
    FILE *f = std::fopen(path, "r");
    if (!f) return nullptr;
    int ver = 0, np = 0, nf = 0;
    char tag[64] = {0};
    bool ok = std::fscanf(f, "%63s %d %d %d", tag, &ver, &np, &nf) == 4 && std::string(tag) == "qmx-trotter-planner" &&
              ver == 1 && np == qmx::kNumParams && nf == qmx::kFeatures;
    std::vector<float> w(qmx::kNumParams);
    float mean[qmx::kFeatures], scale[qmx::kFeatures], margin = 0;
    int knots = 0;
    float kx[qmx::kMaxCurve], ky[qmx::kMaxCurve * qmx::kOutputs];
    for (int i = 0; ok && i < qmx::kNumParams; ++i) ok = std::fscanf(f, "%f", &w[i]) == 1;
    for (int i = 0; ok && i < qmx::kFeatures; ++i) ok = std::fscanf(f, "%f %f", &mean[i], &scale[i]) == 2;
    ok = ok && std::fscanf(f, "%f", &margin) == 1 && std::fscanf(f, "%d", &knots) == 1 && knots >= 2 &&
         knots <= qmx::kMaxCurve;
    for (int k = 0; ok && k < knots; ++k) ok = std::fscanf(f, "%f %f %f", &kx[k], &ky[2 * k], &ky[2 * k + 1]) == 3;
    std::fclose(f);
    if (!ok) {
        std::fprintf(stderr, "libquest_mi355x: qmxPlannerLoad: %s is not a planner file\n", path);
        return nullptr;
    }
    qmxPlanner *p = qmxPlannerCreate(0);
    qmxPlannerSetParams(p, w.data());
    qmxPlannerSetNormalisation(p, mean, scale);
    qmxPlannerSetMargin(p, margin);
    if (!qmxPlannerSetSizeCurve(p, knots, kx, ky)) {
        qmxPlannerDestroy(p);
        return nullptr;
    }
    return p;
}

}  // extern "C"
