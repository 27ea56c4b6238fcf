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
constexpr int kMaxCurve = 16;                   // size-curve knots

struct PlannerDev {
    float *master = nullptr;  // FP32 parameters (Adam updates these)
    float *rounded = nullptr; // the same rounded to bfloat16 (what the layers use)
    float *adamM = nullptr, *adamV = nullptr;
    float *norm = nullptr;    // mean[16], 1/scale[16]
    int step = 0;
};

struct PlannerHost {
    float mean[kFeatures], scale[kFeatures];
    float margin = 0.f;  // log2 safety margin added to predictions
    // size curve: mean log2 r*_o at knots log2(n) (ascending), piecewise linear,
    // extended past the ends with the end segments' slopes; default: log2(n)/(2o)
    int knots = 2;
    float knotLog2n[kMaxCurve] = {0.f, 1.f};
    float knotLog2r[kMaxCurve][kOutputs] = {{0.f, 0.f}, {0.25f, 0.125f}};
};

// mean log2 r*_o of the size curve at log2(n)
inline double sizeCurve(const PlannerHost &h, int o, double x) {
    int k = 0;
    while (k < h.knots - 2 && x > h.knotLog2n[k + 1]) ++k;
    const double x0 = h.knotLog2n[k], x1 = h.knotLog2n[k + 1];
    const double y0 = h.knotLog2r[k][o], y1 = h.knotLog2r[k + 1][o];
    return y0 + (y1 - y0) * (x - x0) / (x1 - x0);
}

}  // namespace qmx

struct qmxPlanner {
    qmx::PlannerDev dev;
    qmx::PlannerHost host;
};

#endif
