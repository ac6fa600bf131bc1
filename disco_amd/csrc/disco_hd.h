/* Host/device qualifier shared by the C99 headers that both gcc and hipcc
 * compile.  Under hipcc every physics leaf is a __host__ __device__ inline
 * function; under a plain C compiler it is a static inline function.
 */
#ifndef DISCO_HD_H
#define DISCO_HD_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DHD __host__ __device__ static inline __attribute__((always_inline))
#else
#include <math.h>
#define DHD static inline
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* IEEE-754 bit pattern of a double, for the counter-based noise hash. */
DHD uint64_t disco_bits(double x) {
    union { double d; uint64_t u; } v;
    v.d = x;
    return v.u;
}

/* a / b rounded to nearest, given y = 1 / b rounded to nearest: q = a y lies
 * within one ulp of a / b, the residual a - b q is exact in one fused
 * multiply-add, and q + (a - b q) y then rounds to the correctly rounded
 * quotient (Markstein's theorem, for quotients and residuals clear of the
 * subnormal range).  Several quotients by one divisor share one division:
 * the results are bit-identical to a / b, at three operations each instead
 * of a full FP64 division sequence. */
DHD double div_rn(double a, double b, double y) {
    const double q = a * y;
    const double r = fma(-b, q, a);
    return fma(r, y, q);
}

/* SplitMix64 finaliser (Steele, Lea & Flood 2014). */
DHD uint64_t disco_splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Uniform deviate in [0,1) from a 64-bit hash (53 mantissa bits). */
DHD double disco_u01(uint64_t h) {
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

#endif
