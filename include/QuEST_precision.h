/* Precision constants for the QuEST C API served by libquest_mi355x.
 *
 * Same macro names and values as QuEST 3.4 (Code/QuEST340/QuEST/include/
 * QuEST_precision.h:28-68), so caller code that tests against REAL_EPS or
 * formats with REAL_STRING_FORMAT keeps working.
 *
 * The library is built with QuEST_PREC == 2: the API-level qreal is double.
 * Registers can still keep their amplitudes as float on the GPU (see
 * createQuregEx() in quest_mi355x.h); that choice is per register and does not
 * change any type in this header.
 */
#ifndef QUEST_PRECISION_H
#define QUEST_PRECISION_H

#include <math.h>

#ifndef QuEST_PREC
#define QuEST_PREC 2
#endif

#if QuEST_PREC == 2
#define qreal double
#define REAL_STRING_FORMAT "%.14f"
#define REAL_QASM_FORMAT "%.14g"
#define REAL_EPS 1e-13
#define REAL_SPECIFIER "%lf"
#define absReal(X) fabs(X)
#else
#error "libquest_mi355x exposes qreal = double only (QuEST_PREC == 2); per-register FP32 storage is chosen with createQuregEx()"
#endif

/* upper bound on sub-registers for the phase-function family (QuEST_precision.h:72) */
#define MAX_NUM_REGS_APPLY_ARBITRARY_PHASE 100

#endif /* QUEST_PRECISION_H */
