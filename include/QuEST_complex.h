/* Complex scalar type of the QuEST 3.4 C API, as exported by libquest_mi355x.
 *
 * Same public names as the reference header
 * Code/QuEST340/QuEST/include/QuEST_complex.h: qcomp (the complex type at
 * qreal's precision), float_complex / double_complex / long_double_complex,
 * toComplex(scalar) -> Complex and fromComplex(Complex) -> qcomp. In C++ they
 * are std::complex types, and creal / cimag / carg / cabs name std::real /
 * std::imag / std::arg / std::abs; in C they are the C99 _Complex types from
 * <tgmath.h>. Callers written against the reference header (QuEST's own tests,
 * QT/utilities.hpp) rely on that header bringing namespace std into scope in
 * C++, and this one does the same.
 */
#ifndef QUEST_COMPLEX_H
#define QUEST_COMPLEX_H

#include "QuEST_precision.h"

#ifdef __cplusplus

#include <cmath>
#include <complex>
using namespace std;

typedef complex<float> float_complex;
typedef complex<double> double_complex;
typedef complex<long double> long_double_complex;

#define creal(x) real(x)
#define cimag(x) imag(x)
#define carg(x) arg(x)
#define cabs(x) abs(x)

#else

#include <tgmath.h>

typedef float complex float_complex;
typedef double complex double_complex;
typedef long double complex long_double_complex;

#define float_complex(r, i) ((float)(r) + ((float)(i)) * I)
#define double_complex(r, i) ((double)(r) + ((double)(i)) * I)
#define long_double_complex(r, i) ((long double)(r) + ((long double)(i)) * I)

#endif

#if QuEST_PREC == 1
#define qcomp float_complex
#else
#define qcomp double_complex
#endif

#define toComplex(scalar) ((Complex){.real = creal(scalar), .imag = cimag(scalar)})
#define fromComplex(comp) qcomp(comp.real, comp.imag)

#endif /* QUEST_COMPLEX_H */
