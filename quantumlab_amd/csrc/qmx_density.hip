// Measurement collapse and density-matrix superoperators.
//
// A density matrix of n qubits is a 2n-qubit vector; element (r, c) sits at
// r + c*2^n, so "row bit q" is bit q and "column bit q" is bit q+n. A unitary
// U on the register maps vec(rho) by conj(U) (x) U; a one-qubit channel maps
// the quad {00, 01, 10, 11} of (row bit, column bit). The reference applies a
// gate as two sweeps (rows, then conjugated columns; QuEST.c:410-421)
// and each channel as its own sweep, a dephasing sweep first for depolarising
// and damping (QuEST_gpu.cu:3027-3078). Here every operation on at most two
// qubits composes on the host into one 4^k x 4^k superoperator (k <= 2), and
// the register is swept once per composed run: a brick-circuit gate with its
// three channels per touched qubit is one sweep instead of five to eight.
#include <algorithm>
#include <cmath>
#include <vector>

#include "qmx_device.h"

namespace qmx {

namespace {

constexpr int kPer = 4;

// statevec collapse (QuEST_cpu.c:3750-3800): keep half × 1/sqrt(p), zero the other
template <typename T>
__global__ void __launch_bounds__(kBlock) kCollapse(Amp<T> *__restrict__ a, long long work, int qubit,
                                                    int outcome, double renorm) {
    const uint64_t bit = 1ull << qubit;
    long long w = flatBlockLinear() * kBlock * kPer + threadIdx.x;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        long long wj = w + (long long)j * kBlock;
        if (wj < work) {
            uint64_t i0 = insertZero((uint64_t)wj, qubit);
            uint64_t keep = outcome ? (i0 | bit) : i0, kill = outcome ? i0 : (i0 | bit);
            Amp<T> v = a[keep];
            a[keep] = Amp<T>{(T)(v.x * renorm), (T)(v.y * renorm)};
            a[kill] = Amp<T>{(T)0, (T)0};
        }
    }
}

// density collapse (QuEST_cpu.c:798-877): keep |o><o| block / p, zero the rest
template <typename T>
__global__ void __launch_bounds__(kBlock) kDensCollapse(Amp<T> *__restrict__ a, long long n, int rowBit,
                                                        int colBit, int outcome, double prob) {
    long long i = flatBlockLinear() * kBlock * kPer + threadIdx.x;
#pragma unroll
    for (int j = 0; j < kPer; ++j, i += kBlock) {
        if (i < n) {
            int r = (int)((i >> rowBit) & 1), c = (int)((i >> colBit) & 1);
            if (r == outcome && c == outcome) {
                Amp<T> v = a[i];
                a[i] = Amp<T>{(T)(v.x / prob), (T)(v.y / prob)};
            } else {
                a[i] = Amp<T>{(T)0, (T)0};
            }
        }
    }
}

// two-qubit dephasing (QuEST_cpu.c:97-136): scale where either qubit's row and
// column bits differ
template <typename T>
__global__ void __launch_bounds__(kBlock) kDephase2(Amp<T> *__restrict__ a, long long n, int q1, int q2,
                                                    int shift, double retain) {
    long long i = flatBlockLinear() * kBlock * kPer + threadIdx.x;
#pragma unroll
    for (int j = 0; j < kPer; ++j, i += kBlock) {
        if (i < n) {
            bool m1 = ((i >> q1) & 1) != ((i >> (q1 + shift)) & 1);
            bool m2 = ((i >> q2) & 1) != ((i >> (q2 + shift)) & 1);
            if (m1 || m2) {
                Amp<T> v = a[i];
                a[i] = Amp<T>{(T)(retain * v.x), (T)(retain * v.y)};
            }
        }
    }
}

}  // namespace

void collapse(QuregDev *q, int qubit, int outcome, double prob) {
    const long long work = q->numAmps >> 1;
    const double renorm = 1.0 / sqrt(prob);
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kCollapse<T>, dim3(gridFor(work, kPer)), dim3(kBlock), 0, hstream(),
                           (Amp<T> *)q->amps, work, qubit, outcome, renorm);
    });
    checkLaunch("collapse");
}

void densCollapse(QuregDev *q, int qubit, int outcome, double prob) {
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kDensCollapse<T>, dim3(gridFor(q->numAmps, kPer)), dim3(kBlock), 0, hstream(),
                           (Amp<T> *)q->amps, q->numAmps, qubit, qubit + q->numQubitsRep, outcome, prob);
    });
    checkLaunch("densCollapse");
}

void densTwoQubitDephase(QuregDev *q, int q1, int q2, double retain) {
    QMX_DISPATCH(q->prec, T, {
        hipLaunchKernelGGL(kDephase2<T>, dim3(gridFor(q->numAmps, kPer)), dim3(kBlock), 0, hstream(),
                           (Amp<T> *)q->amps, q->numAmps, q1, q2, q->numQubitsRep, retain);
    });
    checkLaunch("densTwoQubitDephase");
}

// ---- pending superoperator -------------------------------------------------
namespace {

constexpr int kPendMax = 2;                // qubits one pending superoperator spans
constexpr int kMaxEntries = 256;           // (4^2)^2

// Pending work of one register: superoperators on disjoint qubit sets of at
// most two qubits each, in creation order. Blocks on disjoint qubits commute,
// so an operation may join the one block it touches while others stay pending.
struct Block {
    int k = 0;
    int t[kPendMax] = {0, 0};
    Cplx S[kMaxEntries];
};
struct Pending {
    QuregDev *q = nullptr;
    std::vector<Block> blocks;
};
Pending g_pend;
long long g_passes = 0, g_absorbed = 0;

void launchSuperOp(QuregDev *q, const int *t, int k, const Cplx *S) {
    // index bits 0..k-1: the targets' row bits; k..2k-1: their column bits
    int bits[2 * kPendMax];
    for (int i = 0; i < k; ++i) {
        bits[i] = t[i];
        bits[k + i] = t[i] + q->numQubitsRep;
    }
    applyDense(q, bits, 2 * k, S);
}

void sweepBlock(QuregDev *q, const Block &b) {
    ++g_passes;
    launchSuperOp(q, b.t, b.k, b.S);
}

// blocks act on disjoint qubits and commute: sweep them two at a time, both
// superoperators in one pass (applyDensePair), when the register allows
void sweepBlocks(QuregDev *q, const std::vector<const Block *> &bs) {
    size_t i = 0;
    for (; i + 1 < bs.size(); i += 2) {
        const Block &a = *bs[i], &b = *bs[i + 1];
        if (!densePairSupported(q, 2 * a.k, 2 * b.k)) break;
        int ba[2 * kPendMax], bb[2 * kPendMax];
        for (int j = 0; j < a.k; ++j) {
            ba[j] = a.t[j];
            ba[a.k + j] = a.t[j] + q->numQubitsRep;
        }
        for (int j = 0; j < b.k; ++j) {
            bb[j] = b.t[j];
            bb[b.k + j] = b.t[j] + q->numQubitsRep;
        }
        ++g_passes;
        applyDensePair(q, ba, 2 * a.k, a.S, bb, 2 * b.k, b.S);
    }
    for (; i < bs.size(); ++i) sweepBlock(q, *bs[i]);
}

void flushPending() {
    setPendingWork(nullptr);
    if (!g_pend.q) return;
    QuregDev *q = g_pend.q;
    std::vector<Block> blocks;
    blocks.swap(g_pend.blocks);
    g_pend.q = nullptr;
    std::vector<const Block *> bs;
    for (const Block &b : blocks) bs.push_back(&b);
    sweepBlocks(q, bs);
}

// sub-operator s on subT (ks qubits, all in T) written into the 4^K x 4^K
// space of T: identity on the other qubits' row and column bits
void embedInto(const int *subT, int ks, const Cplx *s, const int *T, int K, Cplx *out) {
    const int ds = 1 << (2 * ks), dim = 1 << (2 * K);
    int p[kPendMax];
    unsigned subMask = 0;
    for (int j = 0; j < ks; ++j) {
        p[j] = (int)(std::find(T, T + K, subT[j]) - T);
        subMask |= (1u << p[j]) | (1u << (K + p[j]));
    }
    auto gather = [&](int X) {
        int r = 0;
        for (int j = 0; j < ks; ++j) {
            r |= ((X >> p[j]) & 1) << j;
            r |= ((X >> (K + p[j])) & 1) << (ks + j);
        }
        return r;
    };
    for (int R = 0; R < dim; ++R)
        for (int C = 0; C < dim; ++C)
            out[R * dim + C] = ((R ^ C) & ~subMask) ? Cplx{0, 0} : s[gather(R) * ds + gather(C)];
}

void channelSuperOp(QuregDev *q, int qubit, double f, double d00, double d01, double d10, double d11) {
    // entry e = row bit + 2 * column bit: e = 0, 3 diagonal, e = 1, 2 off-diagonal
    Cplx S[16] = {};
    S[0] = {d00, 0};
    S[3] = {d01, 0};
    S[12] = {d10, 0};
    S[15] = {d11, 0};
    S[5] = {f, 0};
    S[10] = {f, 0};
    densSuperOp(q, &qubit, 1, S);
}

}  // namespace

void densSuperOp(QuregDev *q, const int *targets, int k, const Cplx *S) {
    if (k > kPendMax) {
        flushPending();
        std::vector<int> all(2 * k);
        for (int i = 0; i < k; ++i) {
            all[i] = targets[i];
            all[k + i] = targets[i] + q->numQubitsRep;
        }
        applyMatrixN(q, 0, all.data(), 2 * k, S);
        return;
    }
    // sweeps below launch directly: the stream must not run the pending work
    // (these very blocks) on their behalf; the hook is set again on return
    setPendingWork(nullptr);
    if (g_pend.q && g_pend.q != q) flushPending();
    g_pend.q = q;
    std::vector<Block> &bl = g_pend.blocks;
    auto touches = [&](const Block &b) {
        for (int i = 0; i < b.k; ++i)
            for (int j = 0; j < k; ++j)
                if (b.t[i] == targets[j]) return true;
        return false;
    };
    int hit = -1, hits = 0;
    for (int i = 0; i < (int)bl.size(); ++i)
        if (touches(bl[i])) {
            hit = i;
            ++hits;
        }
    int into = -1;
    if (hits == 1) {
        int u = bl[hit].k;
        for (int j = 0; j < k; ++j)
            if (std::find(bl[hit].t, bl[hit].t + bl[hit].k, targets[j]) == bl[hit].t + bl[hit].k) ++u;
        if (u <= kPendMax) into = hit;
    } else if (hits == 0 && k == 1) {
        // pair a lone qubit with the latest one-qubit block: one sweep for both
        for (int i = (int)bl.size() - 1; i >= 0 && into < 0; --i)
            if (bl[i].k == 1) into = i;
    }
    if (into < 0 && hits > 0) {
        // the blocks this operation would bridge are swept first; an odd one
        // out takes the oldest other pending block along in the same pass (in
        // a brick layer that block is the next gate's to sweep)
        std::vector<char> sweep(bl.size(), 0);
        int bridged = 0;
        for (size_t i = 0; i < bl.size(); ++i)
            if (touches(bl[i])) {
                sweep[i] = 1;
                ++bridged;
            }
        if (bridged % 2 == 1 && densePairSupported(q, 4, 4))
            for (size_t i = 0; i < bl.size(); ++i)
                if (!sweep[i]) {
                    sweep[i] = 1;
                    break;
                }
        std::vector<const Block *> swept;
        for (size_t i = 0; i < bl.size(); ++i)
            if (sweep[i]) swept.push_back(&bl[i]);
        sweepBlocks(q, swept);
        std::vector<Block> keep;
        for (size_t i = 0; i < bl.size(); ++i)
            if (!sweep[i]) keep.push_back(bl[i]);
        bl.swap(keep);
    }
    if (into < 0) {
        bl.emplace_back();
        Block &b = bl.back();
        b.k = k;
        for (int j = 0; j < k; ++j) b.t[j] = targets[j];
        const int dim = 1 << (2 * k);
        for (int e = 0; e < dim * dim; ++e) b.S[e] = S[e];
    } else {
        Block &b = bl[into];
        int U[2 * kPendMax], K = 0;
        for (int i = 0; i < b.k; ++i) U[K++] = b.t[i];
        for (int j = 0; j < k; ++j)
            if (std::find(U, U + K, targets[j]) == U + K) U[K++] = targets[j];
        const int dim = 1 << (2 * K);
        std::vector<Cplx> cur((size_t)dim * dim), add((size_t)dim * dim);
        embedInto(b.t, b.k, b.S, U, K, cur.data());
        embedInto(targets, k, S, U, K, add.data());
        // later operation on the left
        for (int r = 0; r < dim; ++r)
            for (int c = 0; c < dim; ++c) {
                double re = 0, im = 0;
                for (int m = 0; m < dim; ++m) {
                    const Cplx x = add[(size_t)r * dim + m], y = cur[(size_t)m * dim + c];
                    re += x.re * y.re - x.im * y.im;
                    im += x.re * y.im + x.im * y.re;
                }
                b.S[r * dim + c] = Cplx{re, im};
            }
        b.k = K;
        for (int i = 0; i < K; ++i) b.t[i] = U[i];
    }
    ++g_absorbed;
    setPendingWork(&flushPending);
}

void densUnitary(QuregDev *q, const int *targets, int k, const Cplx *u, const Cplx *uc) {
    // vec(U rho U^dagger) = (conj(U) (x) U) vec(rho): rows take u, columns uc
    const int d = 1 << k, dim = d * d;
    std::vector<Cplx> S((size_t)dim * dim);
    for (int ci = 0; ci < d; ++ci)
        for (int ri = 0; ri < d; ++ri)
            for (int cj = 0; cj < d; ++cj)
                for (int rj = 0; rj < d; ++rj) {
                    const Cplx a = u[ri * d + rj], b = uc[ci * d + cj];
                    S[(size_t)(ci * d + ri) * dim + (cj * d + rj)] =
                        Cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
                }
    densSuperOp(q, targets, k, S.data());
}

void densDropPending(const QuregDev *q) {
    if (g_pend.q == q) {
        g_pend.q = nullptr;
        g_pend.blocks.clear();
        setPendingWork(nullptr);
    }
}

void densSuperOpStats(long long *passes, long long *absorbed) {
    if (passes) *passes = g_passes;
    if (absorbed) *absorbed = g_absorbed;
}

// one-qubit channels as (off-diagonal factor f, diagonal map D) on the
// (row bit, column bit) quad (QuEST_cpu.c:61-95, 138-235)
void densDephase(QuregDev *q, int qubit, double retain) { channelSuperOp(q, qubit, retain, 1, 0, 0, 1); }

void densDepolarise(QuregDev *q, int qubit, double depolLevel) {
    // diag -> (1-d) x + d (x00 + x11)/2
    const double h = depolLevel / 2;
    channelSuperOp(q, qubit, 1 - depolLevel, 1 - h, h, h, 1 - h);
}

void densDamp(QuregDev *q, int qubit, double damping) {
    // |0><0| += g |1><1|, |1><1| *= 1-g, off-diagonals *= sqrt(1-g)
    channelSuperOp(q, qubit, std::sqrt(1 - damping), 1, damping, 0, 1 - damping);
}

}  // namespace qmx
