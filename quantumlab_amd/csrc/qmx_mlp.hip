// The Trotter planner's network: 16 -> 64 -> 64 -> 2, ReLU hidden layers,
// weights and biases rounded to bfloat16, FP32 arithmetic and accumulation.
//
// One workgroup (4 waves) takes 64 samples through the whole network in LDS:
// each wave owns 16 sample rows, and every layer product — forward, and for
// training the backward products dW = H^T dZ and dZ = (dZ' W^T) * relu' — is a
// set of 16 x 16 tiles on v_mfma_f32_16x16x4f32 (A: lane l holds A[l%16][l/16];
// B: lane l holds B[l/16][l%16]; D: lane l holds D[4(l/16)+v][l%16]). A
// workgroup writes its gradient to its own slot; one reduction kernel sums the
// slots in a fixed order and applies Adam, so training is deterministic.
// Loss: mean squared error over the outputs whose mask is 1.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "qmx_device.h"
#include "qmx_planner.h"
#include "quest_mi355x.h"

namespace qmx {

namespace {

constexpr int kRows = 64;  // samples per workgroup
typedef float floatx4 __attribute__((ext_vector_type(4)));

// round to bfloat16, nearest with ties to even (finite inputs)
__device__ __forceinline__ float bf16Round(float x) {
    const unsigned u = __float_as_uint(x);
    return __uint_as_float((u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u);
}

inline float bf16RoundHost(float x) {
    unsigned u;
    std::memcpy(&u, &x, 4);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

// one 16 x 16 tile of C = sum_k A(i, k) B(k, j), K a multiple of 4;
// the lane receives C[i0 + 4 (lane/16) + v][j0 + lane%16], v = 0..3
template <class FA, class FB>
__device__ __forceinline__ floatx4 tile16(FA A, FB B, int K, int i0, int j0) {
    const int lane = threadIdx.x & 63, c = lane & 15, kq = lane >> 4;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 4)
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A(i0 + c, k0 + kq), B(k0 + kq, j0 + c), acc, 0, 0, 0);
    return acc;
}

struct ChunkArgs {
    const float *w;     // rounded parameters
    const float *norm;  // mean[16], 1/scale[16]
    const float *X, *T, *M;
    const int *idx;     // sample indices of this pass
    int count;
    float invDenom;     // 1 / max(1, sum of the pass's mask)
    float *outY;        // forward: count x 2
    float *grad;        // training: one kNumParams slot per workgroup
    float *loss;        // training: one value per workgroup
};

template <bool TRAIN>
__global__ void __launch_bounds__(kBlock) kMlpChunk(ChunkArgs a) {
    __shared__ float sX[kRows][kFeatures];
    __shared__ float sW1[kFeatures][kHidden];
    __shared__ float sW2[kHidden][kHidden];
    __shared__ float sW3[kHidden][kOutputs];
    __shared__ float sB[2 * kHidden + kOutputs];
    __shared__ float sH1[kRows][kHidden + 1];  // H1, later dZ1
    __shared__ float sH2[kRows][kHidden + 1];  // H2, later dZ2
    __shared__ float sD[kRows][4];             // dY, zero-padded to K = 4
    __shared__ float sRed[kBlock / 64];
    const int tid = threadIdx.x, base = blockIdx.x * kRows;
    for (int e = tid; e < kRows * kFeatures; e += kBlock) {
        const int r = e / kFeatures, f = e % kFeatures, s = base + r;
        sX[r][f] = s < a.count ? (a.X[(size_t)a.idx[s] * kFeatures + f] - a.norm[f]) * a.norm[kFeatures + f] : 0.f;
    }
    for (int e = tid; e < kFeatures * kHidden; e += kBlock) sW1[e / kHidden][e % kHidden] = a.w[kOffW1 + e];
    for (int e = tid; e < kHidden * kHidden; e += kBlock) sW2[e / kHidden][e % kHidden] = a.w[kOffW2 + e];
    for (int e = tid; e < kHidden * kOutputs; e += kBlock) sW3[e / kOutputs][e % kOutputs] = a.w[kOffW3 + e];
    if (tid < kHidden) {
        sB[tid] = a.w[kOffB1 + tid];
        sB[kHidden + tid] = a.w[kOffB2 + tid];
    }
    if (tid < kOutputs) sB[2 * kHidden + tid] = a.w[kOffB3 + tid];
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63, c = lane & 15, kq = lane >> 4;
    const int i0 = wave * 16;  // the wave's sample rows
    for (int jt = 0; jt < kHidden / 16; ++jt) {
        const floatx4 z = tile16([&](int i, int k) { return sX[i][k]; }, [&](int k, int j) { return sW1[k][j]; },
                                 kFeatures, i0, jt * 16);
        for (int v = 0; v < 4; ++v) sH1[i0 + 4 * kq + v][jt * 16 + c] = fmaxf(z[v] + sB[jt * 16 + c], 0.f);
    }
    __syncthreads();
    for (int jt = 0; jt < kHidden / 16; ++jt) {
        const floatx4 z = tile16([&](int i, int k) { return sH1[i][k]; }, [&](int k, int j) { return sW2[k][j]; },
                                 kHidden, i0, jt * 16);
        for (int v = 0; v < 4; ++v)
            sH2[i0 + 4 * kq + v][jt * 16 + c] = fmaxf(z[v] + sB[kHidden + jt * 16 + c], 0.f);
    }
    __syncthreads();
    const floatx4 y = tile16([&](int i, int k) { return sH2[i][k]; },
                             [&](int k, int j) { return j < kOutputs ? sW3[k][j] : 0.f; }, kHidden, i0, 0);
    float lossAcc = 0.f;
    if (c < 4) {
        for (int v = 0; v < 4; ++v) {
            const int r = i0 + 4 * kq + v, s = base + r;
            if (c < kOutputs) {
                const float yv = y[v] + sB[2 * kHidden + c];
                if (!TRAIN) {
                    if (s < a.count) a.outY[(size_t)s * kOutputs + c] = yv;
                } else {
                    float dy = 0.f;
                    if (s < a.count) {
                        const size_t o = (size_t)a.idx[s] * kOutputs + c;
                        const float m = a.M[o], d = yv - a.T[o];
                        dy = 2.f * m * d * a.invDenom;
                        lossAcc += m * d * d * a.invDenom;
                    }
                    sD[r][c] = dy;
                }
            } else if (TRAIN) {
                sD[r][c] = 0.f;
            }
        }
    }
    if (!TRAIN) return;
    for (int off = 32; off >= 1; off >>= 1) lossAcc += __shfl_xor(lossAcc, off, 64);
    if (lane == 0) sRed[wave] = lossAcc;
    __syncthreads();
    if (tid == 0) a.loss[blockIdx.x] = sRed[0] + sRed[1] + sRed[2] + sRed[3];
    float *g = a.grad + (size_t)blockIdx.x * kNumParams;
    // dW3 = H2^T dY (the wave's 16 hidden units), db3
    {
        const floatx4 z = tile16([&](int j, int i) { return sH2[i][j]; },
                                 [&](int i, int m) { return m < kOutputs ? sD[i][m] : 0.f; }, kRows, i0, 0);
        if (c < kOutputs)
            for (int v = 0; v < 4; ++v) g[kOffW3 + (i0 + 4 * kq + v) * kOutputs + c] = z[v];
    }
    if (tid < kOutputs) {
        float s = 0.f;
        for (int i = 0; i < kRows; ++i) s += sD[i][tid];
        g[kOffB3 + tid] = s;
    }
    __syncthreads();
    // dZ2 = (dY W3^T) * relu'(Z2), over H2 (own rows)
    for (int jt = 0; jt < kHidden / 16; ++jt) {
        const floatx4 z = tile16([&](int i, int m) { return sD[i][m]; },
                                 [&](int m, int j) { return m < kOutputs ? sW3[j][m] : 0.f; }, 4, i0, jt * 16);
        for (int v = 0; v < 4; ++v) {
            float &h = sH2[i0 + 4 * kq + v][jt * 16 + c];
            h = h > 0.f ? z[v] : 0.f;
        }
    }
    __syncthreads();
    // dW2 = H1^T dZ2, db2
    for (int jt = 0; jt < kHidden / 16; ++jt) {
        const floatx4 z = tile16([&](int k, int i) { return sH1[i][k]; }, [&](int i, int j) { return sH2[i][j]; },
                                 kRows, i0, jt * 16);
        for (int v = 0; v < 4; ++v) g[kOffW2 + (i0 + 4 * kq + v) * kHidden + jt * 16 + c] = z[v];
    }
    if (tid < kHidden) {
        float s = 0.f;
        for (int i = 0; i < kRows; ++i) s += sH2[i][tid];
        g[kOffB2 + tid] = s;
    }
    __syncthreads();
    // dZ1 = (dZ2 W2^T) * relu'(Z1), over H1 (own rows)
    for (int jt = 0; jt < kHidden / 16; ++jt) {
        const floatx4 z = tile16([&](int i, int j) { return sH2[i][j]; }, [&](int j, int k) { return sW2[k][j]; },
                                 kHidden, i0, jt * 16);
        for (int v = 0; v < 4; ++v) {
            float &h = sH1[i0 + 4 * kq + v][jt * 16 + c];
            h = h > 0.f ? z[v] : 0.f;
        }
    }
    __syncthreads();
    // dW1 = X^T dZ1 (the wave's 16 hidden units as columns), db1
    {
        const floatx4 z = tile16([&](int f, int i) { return sX[i][f]; }, [&](int i, int k) { return sH1[i][k]; },
                                 kRows, 0, i0);
        for (int v = 0; v < 4; ++v) g[kOffW1 + (4 * kq + v) * kHidden + i0 + c] = z[v];
    }
    if (tid < kHidden) {
        float s = 0.f;
        for (int i = 0; i < kRows; ++i) s += sH1[i][tid];
        g[kOffB1 + tid] = s;
    }
}

// sum of the per-workgroup gradient slots (fixed order); Adam when master != 0
__global__ void __launch_bounds__(kBlock) kMlpReduce(const float *partial, int slots, const float *lossPart,
                                                     float *gradOut, float *lossOut, float *master, float *rounded,
                                                     float *m, float *v, float lr, float c1, float c2) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p == 0) {
        float s = 0.f;
        for (int k = 0; k < slots; ++k) s += lossPart[k];
        lossOut[0] += s;
    }
    if (p >= kNumParams) return;
    float g = 0.f;
    for (int k = 0; k < slots; ++k) g += partial[(size_t)k * kNumParams + p];
    if (gradOut) gradOut[p] = g;
    if (!master) return;
    const float mm = 0.9f * m[p] + 0.1f * g;
    const float vv = 0.999f * v[p] + 0.001f * g * g;
    m[p] = mm;
    v[p] = vv;
    const float w = master[p] - lr * (mm / c1) / (sqrtf(vv / c2) + 1e-8f);
    master[p] = w;
    rounded[p] = bf16Round(w);
}

struct Scratch {
    float *X = nullptr, *T = nullptr, *M = nullptr, *Y = nullptr, *partial = nullptr, *lossPart = nullptr,
          *lossOut = nullptr, *grad = nullptr;
    int *idx = nullptr;
    ~Scratch() {
        for (void *p : {(void *)X, (void *)T, (void *)M, (void *)Y, (void *)partial, (void *)lossPart, (void *)lossOut,
                        (void *)grad, (void *)idx})
            if (p) devFree(p);
    }
};

template <typename V>
V *uploadN(const V *h, size_t n) {
    V *d = static_cast<V *>(devAlloc(std::max<size_t>(n, 1) * sizeof(V)));
    if (n) hostToDevice(d, h, n * sizeof(V));
    return d;
}

void setParamsDev(qmxPlanner *p, const float *w) {
    std::vector<float> r(w, w + kNumParams);
    for (float &x : r) x = bf16RoundHost(x);
    hostToDevice(p->dev.master, w, kNumParams * sizeof(float));
    hostToDevice(p->dev.rounded, r.data(), kNumParams * sizeof(float));
}

void uploadNorm(qmxPlanner *p) {
    float n[2 * kFeatures];
    for (int f = 0; f < kFeatures; ++f) {
        n[f] = p->host.mean[f];
        n[kFeatures + f] = p->host.scale[f] != 0.f ? 1.f / p->host.scale[f] : 1.f;
    }
    hostToDevice(p->dev.norm, n, sizeof n);
}

// one forward (+ backward) pass over samples idx[0..count)
void chunkPass(qmxPlanner *p, Scratch &s, int count, float invDenom, bool train) {
    ChunkArgs a{p->dev.rounded, p->dev.norm, s.X, s.T, s.M, s.idx, count, invDenom, s.Y, s.partial, s.lossPart};
    const int chunks = (count + kRows - 1) / kRows;
    if (train)
        hipLaunchKernelGGL(kMlpChunk<true>, dim3(chunks), dim3(kBlock), 0, hstream(), a);
    else
        hipLaunchKernelGGL(kMlpChunk<false>, dim3(chunks), dim3(kBlock), 0, hstream(), a);
    checkLaunch("planner network");
}

bool validSamples(int N, const char *who) {
    if (N < 1) {
        std::fprintf(stderr, "libquest_mi355x: %s: need at least one sample\n", who);
        return false;
    }
    return true;
}

}  // namespace

}  // namespace qmx

using namespace qmx;

extern "C" {

int qmxPlannerNumParams(void) { return kNumParams; }

qmxPlanner *qmxPlannerCreate(unsigned seed) {
    runtimeInit();
    auto *p = new qmxPlanner;
    p->dev.master = static_cast<float *>(devAlloc(kNumParams * sizeof(float)));
    p->dev.rounded = static_cast<float *>(devAlloc(kNumParams * sizeof(float)));
    p->dev.adamM = static_cast<float *>(devAlloc(kNumParams * sizeof(float)));
    p->dev.adamV = static_cast<float *>(devAlloc(kNumParams * sizeof(float)));
    p->dev.norm = static_cast<float *>(devAlloc(2 * kFeatures * sizeof(float)));
    for (int f = 0; f < kFeatures; ++f) {
        p->host.mean[f] = 0.f;
        p->host.scale[f] = 1.f;
    }
    uploadNorm(p);
    // He-uniform weights, zero biases
    std::mt19937 rng(seed);
    std::vector<float> w(kNumParams, 0.f);
    auto fill = [&](int off, int fanIn, int count) {
        std::uniform_real_distribution<float> u(-std::sqrt(6.f / fanIn), std::sqrt(6.f / fanIn));
        for (int i = 0; i < count; ++i) w[off + i] = u(rng);
    };
    fill(kOffW1, kFeatures, kFeatures * kHidden);
    fill(kOffW2, kHidden, kHidden * kHidden);
    fill(kOffW3, kHidden, kHidden * kOutputs);
    setParamsDev(p, w.data());
    std::vector<float> zero(kNumParams, 0.f);
    hostToDevice(p->dev.adamM, zero.data(), kNumParams * sizeof(float));
    hostToDevice(p->dev.adamV, zero.data(), kNumParams * sizeof(float));
    return p;
}

void qmxPlannerDestroy(qmxPlanner *p) {
    if (!p) return;
    for (void *q : {(void *)p->dev.master, (void *)p->dev.rounded, (void *)p->dev.adamM, (void *)p->dev.adamV,
                    (void *)p->dev.norm})
        devFree(q);
    delete p;
}

void qmxPlannerGetParams(const qmxPlanner *p, float *out) {
    deviceToHost(out, p->dev.master, kNumParams * sizeof(float));
}

void qmxPlannerSetParams(qmxPlanner *p, const float *params) {
    setParamsDev(p, params);
    std::vector<float> zero(kNumParams, 0.f);
    hostToDevice(p->dev.adamM, zero.data(), kNumParams * sizeof(float));
    hostToDevice(p->dev.adamV, zero.data(), kNumParams * sizeof(float));
    p->dev.step = 0;
}

void qmxPlannerSetNormalisation(qmxPlanner *p, const float *mean, const float *scale) {
    std::memcpy(p->host.mean, mean, sizeof p->host.mean);
    std::memcpy(p->host.scale, scale, sizeof p->host.scale);
    uploadNorm(p);
}

void qmxPlannerGetNormalisation(const qmxPlanner *p, float *mean, float *scale) {
    std::memcpy(mean, p->host.mean, sizeof p->host.mean);
    std::memcpy(scale, p->host.scale, sizeof p->host.scale);
}

void qmxPlannerSetMargin(qmxPlanner *p, float log2Margin) { p->host.margin = log2Margin; }

float qmxPlannerGetMargin(const qmxPlanner *p) { return p->host.margin; }

void qmxPlannerPredict(qmxPlanner *p, const float *X, int N, float *Y) {
    if (!validSamples(N, "qmxPlannerPredict")) return;
    Scratch s;
    std::vector<int> idx(N);
    for (int i = 0; i < N; ++i) idx[i] = i;
    s.X = uploadN(X, (size_t)N * kFeatures);
    s.idx = uploadN(idx.data(), idx.size());
    s.Y = static_cast<float *>(devAlloc((size_t)N * kOutputs * sizeof(float)));
    chunkPass(p, s, N, 0.f, false);
    deviceToHost(Y, s.Y, (size_t)N * kOutputs * sizeof(float));
}

double qmxPlannerGradient(qmxPlanner *p, const float *X, const float *T, const float *M, int N, float *grad) {
    if (!validSamples(N, "qmxPlannerGradient")) return NAN;
    Scratch s;
    std::vector<int> idx(N);
    double msum = 0;
    for (int i = 0; i < N; ++i) idx[i] = i;
    for (size_t i = 0; i < (size_t)N * kOutputs; ++i) msum += M[i];
    const int chunks = (N + kRows - 1) / kRows;
    s.X = uploadN(X, (size_t)N * kFeatures);
    s.T = uploadN(T, (size_t)N * kOutputs);
    s.M = uploadN(M, (size_t)N * kOutputs);
    s.idx = uploadN(idx.data(), idx.size());
    s.partial = static_cast<float *>(devAlloc((size_t)chunks * kNumParams * sizeof(float)));
    s.lossPart = static_cast<float *>(devAlloc((size_t)chunks * sizeof(float)));
    s.lossOut = static_cast<float *>(devAlloc(sizeof(float)));
    s.grad = static_cast<float *>(devAlloc(kNumParams * sizeof(float)));
    const float zero = 0.f;
    hostToDevice(s.lossOut, &zero, sizeof zero);
    chunkPass(p, s, N, (float)(1.0 / std::max(1.0, msum)), true);
    hipLaunchKernelGGL(kMlpReduce, dim3((kNumParams + kBlock - 1) / kBlock), dim3(kBlock), 0, hstream(), s.partial,
                       chunks, s.lossPart, s.grad, s.lossOut, nullptr, nullptr, nullptr, nullptr, 0.f, 1.f, 1.f);
    checkLaunch("qmxPlannerGradient");
    deviceToHost(grad, s.grad, kNumParams * sizeof(float));
    float loss;
    deviceToHost(&loss, s.lossOut, sizeof loss);
    return loss;
}

double qmxPlannerTrain(qmxPlanner *p, const float *X, const float *T, const float *M, int N, int epochs, int batch,
                       float lr, unsigned seed, float *lossPerEpoch) {
    if (!validSamples(N, "qmxPlannerTrain") || epochs < 1 || batch < 1) return NAN;
    Scratch s;
    s.X = uploadN(X, (size_t)N * kFeatures);
    s.T = uploadN(T, (size_t)N * kOutputs);
    s.M = uploadN(M, (size_t)N * kOutputs);
    s.idx = static_cast<int *>(devAlloc((size_t)N * sizeof(int)));
    const int maxChunks = (std::min(batch, N) + kRows - 1) / kRows;
    s.partial = static_cast<float *>(devAlloc((size_t)maxChunks * kNumParams * sizeof(float)));
    s.lossPart = static_cast<float *>(devAlloc((size_t)maxChunks * sizeof(float)));
    s.lossOut = static_cast<float *>(devAlloc((size_t)epochs * sizeof(float)));
    std::vector<float> zeros(epochs, 0.f);
    hostToDevice(s.lossOut, zeros.data(), epochs * sizeof(float));
    std::vector<double> rowMask(N);
    for (int i = 0; i < N; ++i) {
        double m = 0;
        for (int o = 0; o < kOutputs; ++o) m += M[(size_t)i * kOutputs + o];
        rowMask[i] = m;
    }
    std::vector<int> perm(N);
    std::mt19937 rng(seed);
    const int batches = (N + batch - 1) / batch;
    for (int e = 0; e < epochs; ++e) {
        for (int i = 0; i < N; ++i) perm[i] = i;
        std::shuffle(perm.begin(), perm.end(), rng);
        hostToDevice(s.idx, perm.data(), (size_t)N * sizeof(int));
        for (int b = 0; b < batches; ++b) {
            const int first = b * batch, count = std::min(batch, N - first);
            double msum = 0;
            for (int i = first; i < first + count; ++i) msum += rowMask[perm[i]];
            ChunkArgs a{p->dev.rounded, p->dev.norm, s.X, s.T, s.M, s.idx + first, count,
                        (float)(1.0 / std::max(1.0, msum)), nullptr, s.partial, s.lossPart};
            const int chunks = (count + kRows - 1) / kRows;
            hipLaunchKernelGGL(kMlpChunk<true>, dim3(chunks), dim3(kBlock), 0, hstream(), a);
            ++p->dev.step;
            const float c1 = 1.f - std::pow(0.9f, (float)p->dev.step), c2 = 1.f - std::pow(0.999f, (float)p->dev.step);
            hipLaunchKernelGGL(kMlpReduce, dim3((kNumParams + kBlock - 1) / kBlock), dim3(kBlock), 0, hstream(),
                               s.partial, chunks, s.lossPart, nullptr, s.lossOut + e, p->dev.master, p->dev.rounded,
                               p->dev.adamM, p->dev.adamV, lr, c1, c2);
        }
        checkLaunch("qmxPlannerTrain");
    }
    std::vector<float> losses(epochs);
    deviceToHost(losses.data(), s.lossOut, epochs * sizeof(float));
    for (float &l : losses) l /= batches;
    if (lossPerEpoch) std::memcpy(lossPerEpoch, losses.data(), epochs * sizeof(float));
    return losses.back();
}

}  // extern "C"
