// Internal interface of the learned Trotter planner (MLP on the device +
// host-side decision logic).
#ifndef QMX_PLANNER_H
#define QMX_PLANNER_H

#include <cstddef>

namespace qmx {

constexpr int kFeatures = 16;
constexpr int kHidden = 64;
constexpr int kOutputs = 2;  // size-free log2 repetitions for order 2 and order 4
// flat parameter layout: W1 (16 x 64, row-major), b1, W2 (64 x 64), b2, W3 (64 x 2), b3
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kFeatures * kHidden;
constexpr int kOffW2 = kOffB1 + kHidden;
constexpr int kOffB2 = kOffW2 + kHidden * kHidden;
constexpr int kOffW3 = kOffB2 + kHidden;
constexpr int kOffB3 = kOffW3 + kHidden * kOutputs;
constexpr int kNumParams = kOffB3 + kOutputs;  // 5378

struct PlannerDev {
    float *master = nullptr;  // FP32 parameters (Adam updates these)
    float *rounded = nullptr; // the same rounded to bfloat16 (what the layers use)
    float *adamM = nullptr, *adamV = nullptr;
    float *norm = nullptr;    // mean[16], 1/scale[16]
    int step = 0;
};

struct PlannerHost {
    float mean[kFeatures], scale[kFeatures];
    float margin = 0.f;                    // log2 safety margin added to predictions
    float sizeExp[kOutputs] = {0.25f, 0.125f};  // log2 r* grows as sizeExp * log2(n)
};

}  // namespace qmx

struct qmxPlanner {
    qmx::PlannerDev dev;
    qmx::PlannerHost host;
};

#endif
