/* Precision constants for the QuEST C API served by libquest_mi355x.
 *
 * Same macro names and values as QuEST 3.4 (Code/QuEST340/QuEST/include/
 * QuEST_precision.h:19-68), so caller code that tests against REAL_EPS or
 * formats with REAL_STRING_FORMAT keeps working. As in the reference, QuEST_PREC
 * defaults to 1 (qreal = float), which QuantumLab's Real follows
 * (Code/QuantumLab/Base/QLSetup.h:23-28).
 *
 * Two libraries serve the two settings, and a caller links the one matching the
 * QuEST_PREC it compiles with:
 *   QuEST_PREC == 1   libquest_mi355x_f32.so   qreal = float; registers keep
 *                     FP32 amplitudes on the device unless created with
 *                     createQuregEx(..., QMX_PREC_FP64)
 *   QuEST_PREC == 2   libquest_mi355x.so       qreal = double; FP64 amplitudes
 *                     unless created with createQuregEx(..., QMX_PREC_FP32)
 * Reductions accumulate in FP64 in both. QuEST_PREC == 4 (long double) is not
 * served.
 */
#ifndef QUEST_PRECISION_H
#define QUEST_PRECISION_H

#include <math.h>

#ifndef QuEST_PREC
#define QuEST_PREC 1
#endif

#if QuEST_PREC == 1
#define qreal float
#define REAL_STRING_FORMAT "%.8f"
#define REAL_QASM_FORMAT "%.8g"
#define REAL_EPS 1e-5
#define REAL_SPECIFIER "%f"
#define absReal(X) fabs(X)
#elif QuEST_PREC == 2
#define qreal double
#define REAL_STRING_FORMAT "%.14f"
#define REAL_QASM_FORMAT "%.14g"
#define REAL_EPS 1e-13
#define REAL_SPECIFIER "%lf"
#define absReal(X) fabs(X)
#else
#error "libquest_mi355x serves QuEST_PREC 1 (libquest_mi355x_f32.so) and 2 (libquest_mi355x.so)"
#endif

/* upper bound on sub-registers for the phase-function family (QuEST_precision.h:72) */
#define MAX_NUM_REGS_APPLY_ARBITRARY_PHASE 100

#endif /* QUEST_PRECISION_H */
