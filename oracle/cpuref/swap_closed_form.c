/* Swap-test probabilities by the closed form the MI355X plan evaluates — test
 * infrastructure only (the like-for-like CPU baseline of the batched swap
 * tests; the product never loads this library).
 *
 * The swap test of real q-qubit registers |a> and |b> (SwapTest.cpp:70-123:
 * H on the ancilla, the two encodings, q controlled-SWAPs (ancilla, i, i+q),
 * H) leaves the ancilla-0 amplitudes (a_x b_y + a_y b_x) / 2 for every pair
 * (x, y) of register indices, so P(ancilla = 0) is the sum of their squares,
 * (1 + <a|b>^2) / 2 for normalised registers. This evaluates all 4^q of them,
 * as the device does (the matrix is symmetric: pairs x < y count twice), in
 * FP64, one OpenMP thread per test.
 */
#include <stddef.h>

/* tests t = 0..numTests-1: registers of q[t] qubits, amplitudes a + aOff[t]
 * and b + bOff[t] (2^q[t] each); P(ancilla = 0) into p0[t] */
void cpurefSwapP0ClosedForm(int numTests, const int *q, const double *a, const long long *aOff, const double *b,
                            const long long *bOff, double *p0) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < numTests; ++t) {
        const double *x = a + aOff[t], *y = b + bOff[t];
        const long long D = 1LL << q[t];
        double diag = 0.0, off = 0.0;
        for (long long i = 0; i < D; ++i) {
            const double m = x[i] * y[i];  /* (a_i b_i + a_i b_i) / 2 */
            diag += m * m;
            const double xi = x[i], yi = y[i];
            double s = 0.0;
#pragma omp simd reduction(+ : s)
            for (long long j = i + 1; j < D; ++j) {
                const double v = 0.5 * (xi * y[j] + x[j] * yi);
                s += v * v;
            }
            off += s;
        }
        p0[t] = diag + 2.0 * off;
    }
}
