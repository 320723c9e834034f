/* Shared prelude for PDE system sources.
 *
 * A system source defines the reference's physics callbacks F, B, S
 * (pypde/cfuncs.py:6-8; pypde/solvers.py:156-195) and optionally M = dS/dQ.
 * The same text is compiled two ways: by hipRTC as HIP device code (the product
 * path) and by a host C compiler as plain C (callbacks for CPU checkers).
 * PDE_FN hides the difference.  Signatures:
 *
 *   PDE_FN void F(double *out, const double *Q, const double *dQ, int d);  out[V]
 *   PDE_FN void B(double *out, const double *Q, int d);                    out[V*V] row-major
 *   PDE_FN void S(double *out, const double *Q);                           out[V]
 *   PDE_FN void M(double *out, const double *Q);                           out[V*V] = dS/dQ
 *
 * dQ is D x V row-major (dQ[d*V+i] = dQ_i/dx_d); first-order systems ignore it.
 */
#ifndef PDE_PRELUDE_H
#define PDE_PRELUDE_H

#if defined(__HIPCC_RTC__) || defined(__HIP__)
#define PDE_FN __device__ static inline
#define PDE_RINT(x) __builtin_rint(x)
#define PDE_LDEXP(x, k) __builtin_ldexp(x, k)
#define PDE_FMA(a, b, c) __builtin_fma(a, b, c)
#define PDE_FREXP(x, e) __builtin_frexp(x, e)
#else
#include <math.h>
#define PDE_FN static inline
#define PDE_RINT(x) rint(x)
#define PDE_LDEXP(x, k) ldexp(x, k)
#define PDE_FMA(a, b, c) fma(a, b, c)
#define PDE_FREXP(x, e) frexp(x, e)
#endif

/* exp(x) by Cody-Waite argument reduction (x = k ln2 + r, |r| <= ln2/2, ln2 split
 * into a head and a tail) followed by the degree-13 Taylor polynomial of e^r in
 * Horner form with fused multiply-adds, then scaling by 2^k.  The truncation
 * error r^14/14! < 5e-18 is below half an ulp, so the result is within about one
 * ulp of exp(x).  Written out explicitly so that host and device evaluate the
 * same operations in the same order and agree bit for bit. */
PDE_FN double pde_exp(double x) {
  if (x > 709.782712893384) return 1.0 / 0.0;
  if (x < -745.1332191019412) return 0.0;
  const double ln2_hi = 0.6931471805599453;     /* nearest double to ln 2 */
  const double ln2_lo = 2.3190468138462996e-17; /* ln 2 - ln2_hi */
  double k = PDE_RINT(x * 1.4426950408889634);  /* x / ln 2 */
  double r = PDE_FMA(-k, ln2_hi, x);
  r = PDE_FMA(-k, ln2_lo, r);
  /* 1/j! for j = 13 .. 1 */
  double p = 1.6059043836821613e-10;
  p = PDE_FMA(p, r, 2.08767569878681e-09);
  p = PDE_FMA(p, r, 2.505210838544172e-08);
  p = PDE_FMA(p, r, 2.755731922398589e-07);
  p = PDE_FMA(p, r, 2.7557319223985893e-06);
  p = PDE_FMA(p, r, 2.48015873015873e-05);
  p = PDE_FMA(p, r, 0.0001984126984126984);
  p = PDE_FMA(p, r, 0.001388888888888889);
  p = PDE_FMA(p, r, 0.008333333333333333);
  p = PDE_FMA(p, r, 0.041666666666666664);
  p = PDE_FMA(p, r, 0.16666666666666666);
  p = PDE_FMA(p, r, 0.5);
  p = PDE_FMA(p, r, 1.0);
  p = PDE_FMA(p, r, 1.0);
  return PDE_LDEXP(p, (int)k);
}

/* Real cube root of x > 0: x = m 2^(3k) with m in [0.5, 4) (exact exponent
 * split), Newton's iteration y <- y - (y^3 - m) / (3 y^2) from the linear
 * guess 0.62 + 0.26 m (within 9 % of m^(1/3) on [0.5, 4)), six steps (the
 * error squares each step), then 2^k.  Plain IEEE operations in a fixed order,
 * so host and device agree bit for bit.  0 for x = 0, NaN for x < 0 or NaN. */
PDE_FN double pde_cbrt(double x) {
  if (x == 0.0) return 0.0;
  if (!(x > 0.0)) return 0.0 / 0.0;
  if (x > 1.7976931348623157e308) return x;
  int e;
  double m = PDE_FREXP(x, &e); /* x = m 2^e, m in [0.5, 1) */
  int r = e % 3;
  if (r < 0) r += 3;
  const int k = (e - r) / 3;
  m = PDE_LDEXP(m, r);
  double y = 0.62 + 0.26 * m;
  for (int i = 0; i < 6; i++) y = y - (y * y * y - m) / (3.0 * (y * y));
  return PDE_LDEXP(y, k);
}

#endif
