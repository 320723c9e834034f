// Device-side building blocks of the ADER-WENO step, compiled by hipRTC
// together with the constant tables and the user's system source.  The host
// defines before this file:
//   PDE_D, PDE_N, PDE_V                       dimension, order (GL nodes), variables
//   PDE_USE_F/B/S/M, PDE_B_QUIRK, PDE_STIFF   which terms exist, B.dQ form, predictor
//   constexpr tables T_* (see tables.cpp)
//
// Every floating-point operation here mirrors the CPU restatement used as the
// parity checker: dot products are fma chains starting from the first product,
// everything else is evaluated as written (compiled with -ffp-contract=off).
#pragma once

#define DEV __device__ __forceinline__

constexpr int D = PDE_D;
constexpr int N = PDE_N;
constexpr int V = PDE_V;
constexpr int ND = ipow_c(N, D);       // spatial nodes per cell
constexpr int ND1 = ipow_c(N, D - 1);  // transverse nodes per face
constexpr int NT = ND * N;             // space-time nodes per cell
constexpr double FD_EPS = 1.4901161193847656e-08;  // sqrt(2^-52), NumericalDiff.h:9
constexpr double DG_TOL = 6e-6;                     // dg.h:9
constexpr int DG_IT = 50;                           // dg.h:8
constexpr double WENO_EPS = 1e-14;                  // weno.h:10

template <int I> struct IC {
  static constexpr int value = I;
};
// compile-time loop: fn(IC<B>{}), ..., fn(IC<E-1>{})
template <int B, int E, typename Fn>
DEV void static_for(Fn&& fn) {
  if constexpr (B < E) {
    fn(IC<B>{});
    static_for<B + 1, E>(fn);
  }
}

DEV double dfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
DEV double dabs(double a) { return __builtin_fabs(a); }
DEV double dmax(double a, double b) { return __builtin_fmax(a, b); }
DEV double dsqrt(double a) { return __builtin_sqrt(a); }

// node multi-index digit k_d of node idx (row-major, k_0 slowest)
template <int d> DEV constexpr int kdig(int idx) { return (idx / ipow_c(N, D - 1 - d)) % N; }

// ------------------------------------------------------------------- WENO
// One line window (2N-1 values, `st` apart) -> N nodal values (weno.cpp:26-59):
// per stencil c = M^-1 data, p = c^T SIG c + eps, weight o = lambda / p^8, and
// the result sum o c / sum o.  With at most three stencils (N <= 3) the
// weights are formed without a division per stencil: each o is multiplied
// through by prod_r p_r^8, o_s' = lambda_s prod_{r != s} p_r^8 (at least
// (1e-14)^16 = 1e-224, so never below the normal range), and the sums are
// scaled by one reciprocal.  Where that product form overflows (p above
// about 1e19) the division form is used.  Either form gives the same weights
// to a few ulps.
template <typename Get>
DEV void weno_line(Get get, double* out) {
  constexpr int NS = T_NSTEN;
  double c[NS][N], P[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
#pragma unroll
    for (int k = 0; k < N; k++) {
      double acc = T_WINV[s][k][0] * get(T_SOFF[s]);
#pragma unroll
      for (int j = 1; j < N; j++) acc = dfma(T_WINV[s][k][j], get(T_SOFF[s] + j), acc);
      c[s][k] = acc;
    }
    double p = 0.0;
#pragma unroll
    for (int k = 0; k < N; k++) {
      double t = T_SIG[k][0] * c[s][0];
#pragma unroll
      for (int j = 1; j < N; j++) t = dfma(T_SIG[k][j], c[s][j], t);
      p = (k == 0) ? c[s][0] * t : dfma(c[s][k], t, p);
    }
    p = p + WENO_EPS;
    double p2 = p * p;
    double p4 = p2 * p2;
    P[s] = p4 * p4;
  }
  double num[N], den = 0.0;
  if constexpr (NS <= 3) {
#pragma unroll
    for (int k = 0; k < N; k++) num[k] = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
      double o = T_SLAM[s];
#pragma unroll
      for (int r = 0; r < NS; r++)
        if (r != s) o = o * P[r];
#pragma unroll
      for (int k = 0; k < N; k++) num[k] = dfma(o, c[s][k], num[k]);
      den = den + o;
    }
    if (den <= 1.7976931348623157e308) {
      const double rden = 1.0 / den;
#pragma unroll
      for (int k = 0; k < N; k++) out[k] = num[k] * rden;
      return;
    }
    den = 0.0;  // product form overflowed (or NaN): the division form below
  }
#pragma unroll
  for (int k = 0; k < N; k++) num[k] = 0.0;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const double o = T_SLAM[s] / P[s];
#pragma unroll
    for (int k = 0; k < N; k++) num[k] = dfma(o, c[s][k], num[k]);
    den = den + o;
  }
#pragma unroll
  for (int k = 0; k < N; k++) out[k] = num[k] / den;
}

// Ghost index map (boundaries.cpp:7-23); code 2 = reflective wall (an
// addition to the reference's codes).  Returns the interior index, sets
// *flip for mirrors.
DEV int ghost_src(int i, int type, int n, bool* flip) {
  int r = i - N;
  *flip = false;
  if (r < 0) {
    if (type == 0) return 0;
    if (type == 1) return n + r;
    *flip = true;
    return -r - 1;
  }
  if (r >= n) {
    if (type == 0) return n - 1;
    if (type == 1) return r - n;
    *flip = true;
    return 2 * n - r - 1;
  }
  return r;
}

// ------------------------------------------------------------ eigenvalues
// Householder reflector (Golub & Van Loan Alg. 5.1.1, v left unnormalised):
// P x = mu e_1, P = I - beta v v^T, v = (v0, x[1..M-1]), beta = 2 / (v^T v).
// A tail below 2^-900 in square norm counts as reduced (entries are <= 1).
constexpr double HOUSE_TINY = 0x1p-900;
template <int M>
DEV double house(const double* x, double* v) {
  double sigma = 0.0;
#pragma unroll
  for (int i = 1; i < M; i++) sigma = dfma(x[i], x[i], sigma);
  if (sigma <= HOUSE_TINY) {
    v[0] = 1.0;
#pragma unroll
    for (int i = 1; i < M; i++) v[i] = 0.0;
    return 0.0;
  }
  double mu = dsqrt(dfma(x[0], x[0], sigma));
  double v0 = (x[0] <= 0.0) ? x[0] - mu : -sigma / (x[0] + mu);
  v[0] = v0;
#pragma unroll
  for (int i = 1; i < M; i++) v[i] = x[i];
  return 2.0 / dfma(v0, v0, sigma);
}

// The same reflector with one division: v^T v = 2 mu (mu - x0), so beta =
// 1 / (mu (mu - x0)) when x0 <= 0 (v0 = x0 - mu); when x0 > 0 (v0 = -sigma /
// (x0 + mu)) beta = (x0 + mu) / (mu sigma), and both quotients come from r =
// 1 / ((x0 + mu) mu sigma); *mu_out = mu (P x = mu e_1).  For the
// Hessenberg reduction of a matrix scaled to entries <= 1 (every |x_i| <= V):
// a tail below 2^-500 in square norm counts as reduced, which keeps the
// product above 2^-1000.
constexpr double HOUSE1_TINY = 0x1p-500;
template <int M>
DEV double house1(const double* x, double* v, double* mu_out) {
  double sigma = 0.0;
#pragma unroll
  for (int i = 1; i < M; i++) sigma = dfma(x[i], x[i], sigma);
  if (sigma <= HOUSE1_TINY) {
    v[0] = 1.0;
#pragma unroll
    for (int i = 1; i < M; i++) v[i] = 0.0;
    return 0.0;
  }
  const double mu = dsqrt(dfma(x[0], x[0], sigma));
  *mu_out = mu;
  const bool pos = x[0] > 0.0;
  const double s = x[0] + mu;
  const double r = 1.0 / (pos ? (s * mu) * sigma : mu * (mu - x[0]));
  v[0] = pos ? -sigma * ((r * mu) * sigma) : x[0] - mu;
#pragma unroll
  for (int i = 1; i < M; i++) v[i] = x[i];
  return pos ? s * (r * s) : r;
}

// H = P H on rows r0..r0+M-1 for the columns j with lo_c <= j <= hi_c.
template <int M, int R0>
DEV void house_left(double* H, const double* v, double beta, int lo_c, int hi_c) {
#pragma unroll
  for (int j = 0; j < V; j++) {
    if (j >= lo_c && j <= hi_c) {
      double s = v[0] * H[R0 * V + j];
#pragma unroll
      for (int i = 1; i < M; i++) s = dfma(v[i], H[(R0 + i) * V + j], s);
      s = beta * s;
#pragma unroll
      for (int i = 0; i < M; i++) H[(R0 + i) * V + j] = dfma(-s, v[i], H[(R0 + i) * V + j]);
    }
  }
}

// H = H P on columns c0..c0+M-1 for the rows i with lo_r <= i <= hi_r.
template <int M, int C0>
DEV void house_right(double* H, const double* v, double beta, int lo_r, int hi_r) {
#pragma unroll
  for (int i = 0; i < V; i++) {
    if (i >= lo_r && i <= hi_r) {
      double s = v[0] * H[i * V + C0];
#pragma unroll
      for (int j = 1; j < M; j++) s = dfma(v[j], H[i * V + C0 + j], s);
      s = beta * s;
#pragma unroll
      for (int j = 0; j < M; j++) H[i * V + C0 + j] = dfma(-s, v[j], H[i * V + C0 + j]);
    }
  }
}

// H[i][i + off] for a run-time i, every array index static (H stays in
// registers); entries outside the matrix read as 0.
template <int OFF>
DEV double hband(const double* H, int i) {
  double out = 0.0;
#pragma unroll
  for (int r = 0; r < V; r++)
    if (r + OFF >= 0 && r + OFF < V && r == i) out = H[r * V + r + OFF];
  return out;
}

#if PDE_V >= 3
// Characteristic polynomial det(xI - H) of the upper Hessenberg H, monic,
// c[0..V]: leading principal minors expanded along their last column,
//   D_k = (x - h_kk) D_k-1 - sum_{m=1}^{k-1} h_{k-m,k} (prod_{j=k-m+1}^{k} h_{j,j-1}) D_{k-m-1}
// (1-based; Wilkinson, The Algebraic Eigenvalue Problem, ch. 7).
DEV void hess_charpoly(const double* H, double* c) {
  double D[V + 1][V + 1];
  D[0][0] = 1.0;
#pragma unroll
  for (int k = 1; k <= V; k++) {
    const double hkk = H[(k - 1) * V + (k - 1)];
#pragma unroll
    for (int i = 0; i <= k; i++) {
      double v = (i >= 1) ? D[k - 1][i - 1] : 0.0;
      if (i <= k - 1) v = dfma(-hkk, D[k - 1][i], v);
      D[k][i] = v;
    }
    double prod = 1.0;
#pragma unroll
    for (int m = 1; m <= k - 1; m++) {
      prod = prod * H[(k - m) * V + (k - m - 1)];
      const double t = H[(k - m - 1) * V + (k - 1)] * prod;
#pragma unroll
      for (int i = 0; i <= k - m - 1; i++) D[k][i] = dfma(-t, D[k - m - 1][i], D[k][i]);
    }
  }
#pragma unroll
  for (int i = 0; i <= V; i++) c[i] = D[V][i];
}

// K real roots of monic degree-V polynomials, root k of c[k] from x0[k]: by
// Halley's method for V <= 5, by Laguerre's method above.  Both converge
// cubically at simple roots and, from above the largest real root of a
// real-rooted polynomial, decrease monotonically onto it (Halley's step
// p p' / (p'^2 - p p''/2) = 2G / (G^2 + H), G = p'/p = sum 1/(x - r_i), H =
// G^2 - p''/p = sum 1/(x - r_i)^2, stays below the distance to the nearest
// root).  Halley's step needs no square root; from a start far outside the
// spectrum it moves only about 2/(V+1) of the way per step, where Laguerre's
// step lands near the roots at once, which is why the larger systems keep
// Laguerre.  The Laguerre step n / (G +- sqrt((n-1)(n H - G^2))) is evaluated
// multiplied through by p: n p / (p' +- sqrt((n-1)((n-1) p'^2 - n p p''))).
// The K iterations are independent and advance in one loop; each
// stops at its own convergence exactly as it would alone.  The fused Horner
// recurrence for p, p' and p''/2 starts after its first two steps, which are
// exact for the monic c (c[V] = 1, finite x).  Converged: |step| <= 4 eps
// |x|, or |step| <= 1e-8 |x| with |step| |x| <= |previous step|^2
// (superlinear convergence seen, so the remaining error is below
// |step|^2.5).  ok[k] is false when root k does not converge in 40 steps;
// dp[k] is p' where p was last evaluated (at the root, or one step of at most
// 1e-8 |root| before it).
constexpr bool ROOT_HALLEY = V <= 5;
template <int K>
DEV void real_roots(const double (&c)[K][V + 1], const double (&x0)[K], double (&root)[K],
                    double (&dp)[K], bool (&ok)[K]) {
  double x[K], prev[K];
  int state[K];  // 0 running, 1 converged, 2 failed
#pragma unroll
  for (int k = 0; k < K; k++) {
    x[k] = x0[k];
    prev[k] = 0.0;
    state[k] = 0;
    dp[k] = 0.0;
  }
  for (int it = 0; it < 40; it++) {
    bool any = false;
#pragma unroll
    for (int k = 0; k < K; k++) any = any || state[k] == 0;
    if (!any) break;
#pragma unroll
    for (int k = 0; k < K; k++) {
      const double xk = x[k];
      const double b1 = xk + c[k][V - 1];
      double f = 1.0, d = xk + b1, b = dfma(xk, b1, c[k][V - 2]);
#pragma unroll
      for (int j = V - 3; j >= 0; j--) {
        f = dfma(xk, f, d);
        d = dfma(xk, d, b);
        b = dfma(xk, b, c[k][j]);
      }
      double num, den;
      if constexpr (ROOT_HALLEY) {
        num = b * d;
        den = dfma(d, d, -(b * f));
      } else {
        const double r =
            (double)(V - 1) * ((double)(V - 1) * (d * d) - (double)(2 * V) * (f * b));
        const double sq = r > 0.0 ? dsqrt(r) : 0.0;
        num = (double)V * b;
        den = (d >= 0.0) ? d + sq : d - sq;
      }
      const double step = num / den;
      const double xn = xk - step;
      const double as = dabs(step), ax = dabs(xn);
      const bool conv = !(as > 4.0 * 2.220446049250313e-16 * ax) ||
                        as <= 2.2250738585072014e-308 ||
                        (as <= 1e-8 * ax && as * ax <= prev[k] * prev[k]);
      // state update as selects (no branches): a running root records p'; an
      // exact zero of p stops at xk, a zero denominator fails, otherwise the
      // step is taken and convergence (finite root) or failure decided
      const bool run = state[k] == 0;
      const bool exact = b == 0.0;
      const bool step_ok = run && !exact && (den != 0.0);
      const int ns = exact ? 1 : !(den != 0.0) ? 2 : conv ? (ax <= 1e300 ? 1 : 2) : 0;
      dp[k] = run ? d : dp[k];
      x[k] = step_ok ? xn : xk;
      prev[k] = step_ok ? as : prev[k];
      state[k] = run ? ns : state[k];
    }
  }
  // a root converges where its iterate stops: at xk for an exact zero, at
  // the last step's xn otherwise (a failed root is not used)
#pragma unroll
  for (int k = 0; k < K; k++) {
    ok[k] = state[k] == 1;
    root[k] = x[k];
  }
}

// Characteristic polynomial c of the scaled upper Hessenberg H and the
// starting points of the root iterations for its largest and smallest
// real roots: mu +- 1.001 sigma sqrt(V / 2), mu and sigma^2 the mean and
// variance of the roots from c[V-1] and c[V-2] -- for a spectrum of pairs
// u +- a_i around roots at u (the waves of a hyperbolic system) sigma^2 V / 2
// = sum a_i^2, at or above the largest a_i^2 -- or +-||H||_inf when sigma^2
// is not positive.  Any start is safe: the certificate decides.
DEV void charpoly_setup(const double* H, double* c, double* xhi, double* xlo) {
  hess_charpoly(H, c);
  constexpr double inv_v = 1.0 / (double)V;
  const double mu = -(c[V - 1] * inv_v);
  const double s2 = dfma(-mu, mu, dfma(c[V - 1], c[V - 1], -2.0 * c[V - 2]) * inv_v);
  if (s2 > 0.0 && s2 <= 1e300) {
    const double w = 1.001 * dsqrt(s2 * (0.5 * (double)V));
    *xhi = mu + w;
    *xlo = mu - w;
    return;
  }
  double u = 0.0;
#pragma unroll
  for (int i = 0; i < V; i++) {
    double rs = 0.0;
#pragma unroll
    for (int j = 0; j < V; j++) rs = rs + dabs(H[i * V + j]);
    u = dmax(u, rs);
  }
  *xhi = u;
  *xlo = -u;
}

// Given the largest and smallest real roots l1 > l2 of c, with p' = d1, d2
// there, check that they are two roots (l1 - l2 > 1e-9 M), that both are
// accurate -- the rounding error of evaluating p,
// 2 V eps sum_k |c_k| |x|^k, moves a root by that over |p'|, which must stay
// below 1e-11 M, M = max(|l1|, |l2|) (a simple root next to a cluster is
// ill-conditioned as a polynomial root) -- and certify that every
// other root is smaller in modulus: q = p / ((x - l1)(x - l2)) is centred on
// the mean of its roots, alpha, and its roots bounded by |z - alpha| <= R,
// R = 2^(E+1) >= 2 max_k |t_k|^(1/(m-k)) (Fujiwara's bound, in powers of
// two: |t_k| < 2^e_k gives |t_k|^(1/j) < 2^ceil(e_k/j)) for t(y) = q(y +
// alpha).  False (the caller falls back to QR) when the certificate fails:
// complex or repeated extreme eigenvalues.
DEV bool charpoly_certify(const double* c, double l1, double l2, double d1, double d2,
                          double* out) {
  constexpr int m = V - 2;
  const double M = dmax(dabs(l1), dabs(l2));
  // two distinct roots (both runs can land on the one real root)
  if (!(l1 - l2 > 1e-9 * M)) return false;
  {
    const double a1 = dabs(l1), a2 = dabs(l2);
    double E1 = 1.0, E2 = 1.0;
#pragma unroll
    for (int j = V - 1; j >= 0; j--) {
      E1 = dfma(a1, E1, dabs(c[j]));
      E2 = dfma(a2, E2, dabs(c[j]));
    }
    if (!((double)(2 * V) * 2.220446049250313e-16 * E1 <= 1e-11 * M * dabs(d1))) return false;
    if (!((double)(2 * V) * 2.220446049250313e-16 * E2 <= 1e-11 * M * dabs(d2))) return false;
  }
  // q = p / (x - l1) / (x - l2)
  double q1[V], q[V - 1];
  q1[V - 1] = c[V];
#pragma unroll
  for (int j = V - 2; j >= 0; j--) q1[j] = dfma(l1, q1[j + 1], c[j + 1]);
  q[V - 2] = q1[V - 1];
#pragma unroll
  for (int j = V - 3; j >= 0; j--) q[j] = dfma(l2, q[j + 1], q1[j + 1]);
  constexpr double inv_m = 1.0 / (double)m;
  const double alpha = -(q[m - 1] * inv_m);
  double t[m + 1];
#pragma unroll
  for (int i = 0; i <= m; i++) t[i] = q[i];
#pragma unroll
  for (int k = 0; k < m; k++)
#pragma unroll
    for (int j = m - 1; j >= k; j--) t[j] = dfma(alpha, t[j + 1], t[j]);
  double tsum = 0.0;
#pragma unroll
  for (int k = 0; k < m; k++) tsum = tsum + dabs(t[k]);
  if (!(tsum <= 1.7976931348623157e308)) return false;
  int emax = -2000;
#pragma unroll
  for (int k = 0; k < m; k++) {
    if (t[k] != 0.0) {
      int e;
      (void)__builtin_frexp(t[k], &e);  // |t_k| = f 2^e, f in [0.5, 1)
      const int jj = m - k;
      const int ce = e >= 0 ? (e + jj - 1) / jj : -((-e) / jj);
      emax = ce > emax ? ce : emax;
    }
  }
  const double R = emax == -2000 ? 0.0 : __builtin_ldexp(1.0, emax + 1);
  if (!(dabs(alpha) + R < M)) return false;
  *out = M;
  return true;
}
#endif

// Scale A (destroyed) by the power of two 2^-e2 at or above its largest
// |entry| above V = 5 or when that lies outside 2^+-64 (*scaled), and reduce
// it to upper Hessenberg form in place.  False when A is zero, subnormal or
// NaN; *early is then max |lambda|.
DEV bool eig_prepare(double* H, int* e2out, bool* scaled, double* early) {
  double scale = 0.0;
#pragma unroll
  for (int i = 0; i < V * V; i++) scale = dmax(scale, dabs(H[i]));
  if (!(scale >= 2.2250738585072014e-308)) {
    *early = (scale != scale) ? scale : 0.0;
    return false;
  }
  int e2;
  (void)__builtin_frexp(scale, &e2);  // scale = m 2^e2, m in [0.5, 1)
  *e2out = e2;
  // Up to V = 5, with the largest |entry| within 2^+-64, the matrix is left
  // unscaled: every operation of the reduction and of the polynomial path
  // then stays in the normal range (products of degree up to 2 V - 1 < 16)
  // and commutes exactly with scaling by a power of two, so only the QR
  // fallback, whose shifts are absolute, takes the scaled matrix (max_abs_eig
  // scales it).
  *scaled = V > 5 || e2 < -64 || e2 > 64;
  if (*scaled) {
    const double pw = __builtin_ldexp(1.0, -e2);
#pragma unroll
    for (int i = 0; i < V * V; i++) H[i] = H[i] * pw;
  }
  static_for<0, (V >= 3 ? V - 2 : 0)>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    constexpr int m = V - k - 1;
    double x[m], vv[m];
#pragma unroll
    for (int i = 0; i < m; i++) x[i] = H[(k + 1 + i) * V + k];
    double mu = 0.0;
    double beta = house1<m>(x, vv, &mu);
    if (beta != 0.0) {
      // column k is x: P x = mu e_1, set rather than computed
      house_left<m, k + 1>(H, vv, beta, k + 1, V - 1);
      H[(k + 1) * V + k] = mu;
#pragma unroll
      for (int i = k + 2; i < V; i++) H[i * V + k] = 0.0;
      house_right<m, k + 1>(H, vv, beta, 0, V - 1);
    }
  });
  return true;
}

// Francis double-shift QR on the scaled Hessenberg H (destroyed): max
// |lambda| in the scaled units.
DEV double eig_qr(double* H) {
  double anorm = 0.0;
#pragma unroll
  for (int i = 0; i < V * V; i++) anorm = anorm + dabs(H[i]);
  double best = 0.0, ex = 0.0;
  int hi = V - 1, iter = 0, total = 0;
  while (hi >= 0) {
    int lo = 0;
    bool found = false;
#pragma unroll
    for (int l = V - 1; l >= 1; l--) {
      if (!found && l <= hi) {
        double s = dabs(H[(l - 1) * V + l - 1]) + dabs(H[l * V + l]);
        if (s == 0.0) s = anorm;
        if (dabs(H[l * V + l - 1]) <= 2.220446049250313e-16 * s) {
          H[l * V + l - 1] = 0.0;
          lo = l;
          found = true;
        }
      }
    }
    if (lo == hi) {
      best = dmax(best, dabs(hband<0>(H, hi) + ex));
      hi -= 1;
      iter = 0;
      continue;
    }
    if (lo == hi - 1) {
      double a = hband<0>(H, hi - 1) + ex, b = hband<1>(H, hi - 1);
      double c = hband<-1>(H, hi), dd = hband<0>(H, hi) + ex;
      double p = 0.5 * (a - dd);
      double bc = b * c;
      double disc = p * p + bc;
      if (disc >= 0.0) {
        double z = dsqrt(disc);
        z = (p >= 0.0) ? p + z : p - z;
        double l1 = dd + z;
        double l2 = (z != 0.0) ? dd - bc / z : dd;
        best = dmax(best, dmax(dabs(l1), dabs(l2)));
      } else {
        double re = dd + p;
        best = dmax(best, dsqrt(re * re - disc));
      }
      hi -= 2;
      iter = 0;
      continue;
    }
    if constexpr (V >= 3) {
      if (total >= 40 * V) {
#pragma unroll
        for (int i = 0; i < V; i++)
          if (i <= hi) best = dmax(best, dabs(H[i * V + i] + ex));
        break;
      }
      double sx = hband<0>(H, hi), sy = hband<0>(H, hi - 1);
      double sw = hband<-1>(H, hi) * hband<1>(H, hi - 1);
      if (iter == 10) {
        ex = ex + sx;
#pragma unroll
        for (int i = 0; i < V; i++)
          if (i <= hi) H[i * V + i] = H[i * V + i] - sx;
        double s = dabs(hband<-1>(H, hi)) + dabs(hband<-1>(H, hi - 1));
        sx = 0.75 * s;
        sy = sx;
        sw = -0.4375 * s * s;
      } else if (iter == 30) {
        double s = (sy - sx) / 2.0;
        s = s * s + sw;
        if (s > 0.0) {
          s = dsqrt(s);
          if (sy < sx) s = -s;
          s = s + (sy - sx) / 2.0;
          s = sx - sw / s;
          ex = ex + s;
#pragma unroll
          for (int i = 0; i < V; i++)
            if (i <= hi) H[i * V + i] = H[i * V + i] - s;
          sx = 0.964;
          sy = 0.964;
          sw = 0.964;
        }
      }
      iter++;
      total++;
      double h00 = hband<0>(H, lo);
      double r = sx - h00, q = sy - h00;
      double x[3], v[3];
      x[0] = (r * q - sw) / hband<-1>(H, lo + 1) + hband<1>(H, lo);
      x[1] = hband<0>(H, lo + 1) - h00 - r - q;
      x[2] = hband<-1>(H, lo + 2);
      static_for<0, (V >= 3 ? V - 2 : 0)>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (k >= lo && k <= hi - 2) {
          if constexpr (k >= 1) {
            if (k > lo) {
              x[0] = H[k * V + k - 1];
              x[1] = H[(k + 1) * V + k - 1];
              x[2] = H[(k + 2) * V + k - 1];
            }
          }
          double beta = house<3>(x, v);
          if (beta != 0.0) {
            if constexpr (k >= 1) {
              if (k > lo) {
                H[k * V + k - 1] = dsqrt(x[0] * x[0] + (x[1] * x[1] + x[2] * x[2]));
                H[(k + 1) * V + k - 1] = 0.0;
                H[(k + 2) * V + k - 1] = 0.0;
              }
            }
            house_left<3, k>(H, v, beta, k, hi);
            int rr = (k + 3 < hi) ? k + 3 : hi;
            house_right<3, k>(H, v, beta, lo, rr);
          }
        }
      });
      static_for<2, (V >= 3 ? V : 2)>([&](auto hc) {
        constexpr int hh = decltype(hc)::value;  // hi in 2..V-1
        if (hh == hi) {
          x[0] = H[(hh - 1) * V + hh - 2];
          x[1] = H[hh * V + hh - 2];
          double beta = house<2>(x, v);
          if (beta != 0.0) {
            H[(hh - 1) * V + hh - 2] = dsqrt(x[0] * x[0] + x[1] * x[1]);
            H[hh * V + hh - 2] = 0.0;
            house_left<2, hh - 1>(H, v, beta, hh - 1, hh);
            house_right<2, hh - 1>(H, v, beta, lo, hh);
          }
        }
      });
    }
  }
  return best;
}

// max |lambda| of A (V x V, row-major; destroyed), as the reference takes it from
// Eigen's EigenSolver (system.cpp:28-43): scale by the power of two at or above
// the largest |entry| (exact; Eigen's RealSchur divides by the largest |entry|,
// RealSchur.h:258-276), Householder Hessenberg reduction, then
// Francis double-shift QR steps (Golub & Van Loan Alg. 7.5.1-7.5.2, EISPACK
// hqr structure) with deflation on eps-small subdiagonals and origin-moving
// exceptional shifts after 10 and 30 steps without deflation; after 40 V steps
// the remaining diagonal stands in for unresolved eigenvalues.  All array
// indices are compile-time constants; run-time bounds only predicate.
DEV double max_abs_eig(double* H) {
  int e2 = 0;
  bool scaled = true;
  double early;
  if (!eig_prepare(H, &e2, &scaled, &early)) return early;
#if PDE_V >= 3
  {
    double c[2][V + 1], x0[2], root[2], dp[2], r;
    bool ok[2];
    charpoly_setup(H, c[0], &x0[0], &x0[1]);
#pragma unroll
    for (int j = 0; j <= V; j++) c[1][j] = c[0][j];
    real_roots<2>(c, x0, root, dp, ok);
    if (ok[0] && ok[1] && charpoly_certify(c[0], root[0], root[1], dp[0], dp[1], &r))
      return scaled ? __builtin_ldexp(r, e2) : r;
  }
#endif
  if (!scaled) {
    const double pw = __builtin_ldexp(1.0, -e2);
#pragma unroll
    for (int i = 0; i < V * V; i++) H[i] = H[i] * pw;
  }
  return __builtin_ldexp(eig_qr(H), e2);
}

// ------------------------------------------------------ system functions
DEV void callF(double* out, const double* q, const double* dq, int d) {
#if PDE_USE_F
  F(out, q, dq, d);
#else
#pragma unroll
  for (int v = 0; v < V; v++) out[v] = 0.0;
#endif
}

// 1 / h_i for the forward-difference steps h_i = max(FD_EPS |q_i|, FD_EPS) =
// 2^-26 m_i, m_i = max(|q_i|, 1) (exact: FD_EPS = 2^-26) with one division
// for all of them: 1 / h_i = (2^26 / prod m) (prod_{j<i} m_j prod_{j>i} m_j).
// When prod m passes 2^1000 each is divided on its own.
DEV void fd_recips(const double* q, double* rh) {
  double m[V], pre[V + 1], suf[V + 1];
  pre[0] = 1.0;
#pragma unroll
  for (int i = 0; i < V; i++) {
    m[i] = dmax(dabs(q[i]), 1.0);
    pre[i + 1] = pre[i] * m[i];
  }
  if (!(pre[V] <= 0x1p1000)) {
#pragma unroll
    for (int i = 0; i < V; i++) rh[i] = 1.0 / dmax(FD_EPS * dabs(q[i]), FD_EPS);
    return;
  }
  suf[V] = 1.0;
#pragma unroll
  for (int i = V - 1; i >= 0; i--) suf[i] = suf[i + 1] * m[i];
  const double r = 0x1p26 / pre[V];
#pragma unroll
  for (int i = 0; i < V; i++) rh[i] = r * (pre[i] * suf[i + 1]);
}

// A = dF/dQ by forward differences + B(q, d) (system.cpp:6-26,
// NumericalDiff.h:6-44), column i scaled by 1 / h_i (fd_recips; the reference
// divides every entry by h_i); f0 receives F(q).
DEV void system_matrix(const double* q_in, const double* dq, int d, double* A, double* f0) {
  double q[V];
#pragma unroll
  for (int i = 0; i < V; i++) q[i] = q_in[i];
#if PDE_USE_F
  double v1[V], v2[V], rh[V];
  F(v1, q, dq, d);
  fd_recips(q, rh);
#pragma unroll
  for (int i = 0; i < V; i++) {
    double qi0 = q[i];
    double h = dmax(FD_EPS * dabs(qi0), FD_EPS);
    q[i] = qi0 + h;
    F(v2, q, dq, d);
    q[i] = qi0;
#pragma unroll
    for (int r = 0; r < V; r++) A[r * V + i] = (v2[r] - v1[r]) * rh[i];
  }
#pragma unroll
  for (int i = 0; i < V; i++) f0[i] = v1[i];
#else
#pragma unroll
  for (int i = 0; i < V * V; i++) A[i] = 0.0;
#pragma unroll
  for (int i = 0; i < V; i++) f0[i] = 0.0;
#endif
#if PDE_USE_B
  double b[V * V];
  B(b, q, d);
#pragma unroll
  for (int i = 0; i < V * V; i++) A[i] = A[i] + b[i];
#endif
}

DEV double max_abs_eigs(const double* q, const double* dq, int d, double* f0) {
  double A[V * V];
  system_matrix(q, dq, d, A, f0);
  return max_abs_eig(A);
}

#ifndef PDE_SECOND
#define PDE_SECOND 0
#endif
#if PDE_SECOND
// Viscous wave speed (system.cpp:53-61): 2 (N + 1) / dX_d max|lambda(dF_d /
// d(dQ_d))|, the Jacobian by forward differences in dQ[d V + i].
template <int d>
DEV double max_abs_eigs_second_order(const double* q_in, const double* dq_in, double dxd) {
  double q[V], dq[D * V], v1[V], v2[V], A[V * V];
#pragma unroll
  for (int i = 0; i < V; i++) q[i] = q_in[i];
#pragma unroll
  for (int i = 0; i < D * V; i++) dq[i] = dq_in[i];
  F(v1, q, dq, d);
#pragma unroll
  for (int i = 0; i < V; i++) {
    double qi0 = dq[d * V + i];
    double h = dmax(FD_EPS * dabs(qi0), FD_EPS);
    dq[d * V + i] = qi0 + h;
    F(v2, q, dq, d);
    dq[d * V + i] = qi0;
    double rh = 1.0 / h;
#pragma unroll
    for (int r = 0; r < V; r++) A[r * V + i] = (v2[r] - v1[r]) * rh;
  }
  return (double)(2 * (N + 1)) / dxd * max_abs_eig(A);
}
#endif

// B(q) applied to a gradient: matrix-vector product, or (PDE_B_QUIRK) the
// reference's depth-1 outer-product evaluation (quirk 1: dg.cpp:110, b(0,0) dq[v];
// quirk 2: fv.cpp:234, b(v,0) dq[0]).
template <int QUIRK>
DEV void b_times(const double* b, const double* dqd, double* out) {
#pragma unroll
  for (int v = 0; v < V; v++) {
    if constexpr (QUIRK == 1) {
      out[v] = b[0] * dqd[v];
    } else if constexpr (QUIRK == 2) {
      out[v] = b[v * V] * dqd[0];
    } else {
      double acc = b[v * V] * dqd[0];
#pragma unroll
      for (int j = 1; j < V; j++) acc = dfma(b[v * V + j], dqd[j], acc);
      out[v] = acc;
    }
  }
}

// ------------------------------------------------ |A| x for Roe and Osher
#ifndef PDE_FLUX
#define PDE_FLUX 0
#endif
#if PDE_FLUX != 0
DEV double guard_div(double w, double tn) { return (w == 0.0) ? 2.220446049250313e-16 * tn : w; }

// Real Schur form T = Z^T H Z (H overwritten by T): the QR of max_abs_eig
// applied to the whole matrix with the transformations accumulated in Z,
// exceptional shifts restored on the diagonal as blocks deflate, and 2 x 2
// blocks with real eigenvalues split by a rotation (EISPACK hqr2 arrangement).
DEV void real_schur(double* H, double* Z) {
#pragma unroll
  for (int i = 0; i < V * V; i++) Z[i] = (i % (V + 1) == 0) ? 1.0 : 0.0;
  static_for<0, (V >= 3 ? V - 2 : 0)>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    constexpr int m = V - k - 1;
    double x[m], vv[m];
#pragma unroll
    for (int i = 0; i < m; i++) x[i] = H[(k + 1 + i) * V + k];
    double beta = house<m>(x, vv);
    if (beta != 0.0) {
      house_left<m, k + 1>(H, vv, beta, k, V - 1);
      house_right<m, k + 1>(H, vv, beta, 0, V - 1);
      house_right<m, k + 1>(Z, vv, beta, 0, V - 1);
    }
#pragma unroll
    for (int i = k + 2; i < V; i++) H[i * V + k] = 0.0;
  });
  double anorm = 0.0;
#pragma unroll
  for (int i = 0; i < V * V; i++) anorm = anorm + dabs(H[i]);
  double ex = 0.0;
  int hi = V - 1, iter = 0, total = 0;
  while (hi >= 0) {
    int lo = 0;
    bool found = false;
#pragma unroll
    for (int l = V - 1; l >= 1; l--) {
      if (!found && l <= hi) {
        double s = dabs(H[(l - 1) * V + l - 1]) + dabs(H[l * V + l]);
        if (s == 0.0) s = anorm;
        if (dabs(H[l * V + l - 1]) <= 2.220446049250313e-16 * s) {
          H[l * V + l - 1] = 0.0;
          lo = l;
          found = true;
        }
      }
    }
    if (lo == hi) {
#pragma unroll
      for (int i = 0; i < V; i++)
        if (i == hi) H[i * V + i] = H[i * V + i] + ex;
      hi -= 1;
      iter = 0;
      continue;
    }
    if (lo == hi - 1) {
      static_for<1, V>([&](auto hc) {
        constexpr int hh = decltype(hc)::value;
        constexpr int na = hh - 1;
        if (hh == hi) {
          H[na * V + na] = H[na * V + na] + ex;
          H[hh * V + hh] = H[hh * V + hh] + ex;
          double a = H[na * V + na], b = H[na * V + hh], c = H[hh * V + na], d = H[hh * V + hh];
          double p = 0.5 * (a - d);
          double disc = p * p + b * c;
          if (disc >= 0.0) {
            double z = dsqrt(disc);
            z = (p >= 0.0) ? p + z : p - z;
            double xr = c, sr = dabs(xr) + dabs(z);
            double pr = xr / sr, qr = z / sr;
            double rr = dsqrt(pr * pr + qr * qr);
            pr = pr / rr;
            qr = qr / rr;
#pragma unroll
            for (int j = na; j < V; j++) {
              double t = H[na * V + j];
              H[na * V + j] = qr * t + pr * H[hh * V + j];
              H[hh * V + j] = qr * H[hh * V + j] - pr * t;
            }
#pragma unroll
            for (int i = 0; i <= hh; i++) {
              double t = H[i * V + na];
              H[i * V + na] = qr * t + pr * H[i * V + hh];
              H[i * V + hh] = qr * H[i * V + hh] - pr * t;
            }
#pragma unroll
            for (int i = 0; i < V; i++) {
              double t = Z[i * V + na];
              Z[i * V + na] = qr * t + pr * Z[i * V + hh];
              Z[i * V + hh] = qr * Z[i * V + hh] - pr * t;
            }
            H[hh * V + na] = 0.0;
          }
        }
      });
      hi -= 2;
      iter = 0;
      continue;
    }
    if constexpr (V >= 3) {
      if (total >= 40 * V) {
#pragma unroll
        for (int i = 0; i < V; i++)
          if (i <= hi) H[i * V + i] = H[i * V + i] + ex;
        break;
      }
      double sx = hband<0>(H, hi), sy = hband<0>(H, hi - 1);
      double sw = hband<-1>(H, hi) * hband<1>(H, hi - 1);
      if (iter == 10) {
        ex = ex + sx;
#pragma unroll
        for (int i = 0; i < V; i++)
          if (i <= hi) H[i * V + i] = H[i * V + i] - sx;
        double s = dabs(hband<-1>(H, hi)) + dabs(hband<-1>(H, hi - 1));
        sx = 0.75 * s;
        sy = sx;
        sw = -0.4375 * s * s;
      } else if (iter == 30) {
        double s = (sy - sx) / 2.0;
        s = s * s + sw;
        if (s > 0.0) {
          s = dsqrt(s);
          if (sy < sx) s = -s;
          s = s + (sy - sx) / 2.0;
          s = sx - sw / s;
          ex = ex + s;
#pragma unroll
          for (int i = 0; i < V; i++)
            if (i <= hi) H[i * V + i] = H[i * V + i] - s;
          sx = 0.964;
          sy = 0.964;
          sw = 0.964;
        }
      }
      iter++;
      total++;
      double h00 = hband<0>(H, lo);
      double r = sx - h00, q = sy - h00;
      double x[3], v[3];
      x[0] = (r * q - sw) / hband<-1>(H, lo + 1) + hband<1>(H, lo);
      x[1] = hband<0>(H, lo + 1) - h00 - r - q;
      x[2] = hband<-1>(H, lo + 2);
      static_for<0, (V >= 3 ? V - 2 : 0)>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (k >= lo && k <= hi - 2) {
          if constexpr (k >= 1) {
            if (k > lo) {
              x[0] = H[k * V + k - 1];
              x[1] = H[(k + 1) * V + k - 1];
              x[2] = H[(k + 2) * V + k - 1];
            }
          }
          double beta = house<3>(x, v);
          if (beta != 0.0) {
            if constexpr (k >= 1) {
              if (k > lo) {
                H[k * V + k - 1] = dsqrt(x[0] * x[0] + (x[1] * x[1] + x[2] * x[2]));
                H[(k + 1) * V + k - 1] = 0.0;
                H[(k + 2) * V + k - 1] = 0.0;
              }
            }
            house_left<3, k>(H, v, beta, k, V - 1);
            int rr = (k + 3 < hi) ? k + 3 : hi;
            house_right<3, k>(H, v, beta, 0, rr);
            house_right<3, k>(Z, v, beta, 0, V - 1);
          }
        }
      });
      static_for<2, (V >= 3 ? V : 2)>([&](auto hc) {
        constexpr int hh = decltype(hc)::value;
        if (hh == hi) {
          x[0] = H[(hh - 1) * V + hh - 2];
          x[1] = H[hh * V + hh - 2];
          double beta = house<2>(x, v);
          if (beta != 0.0) {
            H[(hh - 1) * V + hh - 2] = dsqrt(x[0] * x[0] + x[1] * x[1]);
            H[hh * V + hh - 2] = 0.0;
            house_left<2, hh - 1>(H, v, beta, hh - 1, V - 1);
            house_right<2, hh - 1>(H, v, beta, 0, hh);
            house_right<2, hh - 1>(Z, v, beta, 0, V - 1);
          }
        }
      });
    }
  }
}

// y = |A| x = R |Lambda| R^-1 x (fluxes.cpp:20-70), in real arithmetic: from
// the real Schur form, a basis Y of each invariant subspace of T (eigenvector
// of a real eigenvalue; real and imaginary parts of the eigenvector of a
// complex pair, on which |A| is |lambda| I), W = Z Y, y = W D W^-1 x.
DEV void abs_apply(const double* A, const double* xin, double* y) {
  double T[V * V], Zm[V * V], Y[V * V], D[V];
  double scale = 0.0;
#pragma unroll
  for (int i = 0; i < V * V; i++) scale = dmax(scale, dabs(A[i]));
  if (!(scale >= 2.2250738585072014e-308)) {
#pragma unroll
    for (int i = 0; i < V; i++) y[i] = (scale != scale) ? scale : 0.0;
    return;
  }
  int e2;
  (void)__builtin_frexp(scale, &e2);
  const double pw = __builtin_ldexp(1.0, -e2);
#pragma unroll
  for (int i = 0; i < V * V; i++) T[i] = A[i] * pw;
  real_schur(T, Zm);
  double tn = 0.0;
#pragma unroll
  for (int i = 0; i < V * V; i++) tn = tn + dabs(T[i]);
#pragma unroll
  for (int i = 0; i < V * V; i++) Y[i] = 0.0;
  bool skip = false;
  static_for<0, V>([&](auto ic) {
    constexpr int en = V - 1 - decltype(ic)::value;
    if (skip) {
      skip = false;
      return;
    }
    bool cpx = false;
    if constexpr (en >= 1) cpx = T[en * V + en - 1] != 0.0;
    if (cpx) {
      if constexpr (en >= 1) {
        constexpr int na = en - 1;
        double a = 0.5 * (T[na * V + na] + T[en * V + en]);
        double pp = 0.5 * (T[na * V + na] - T[en * V + en]);
        double bb = dsqrt(-(pp * pp + T[na * V + en] * T[en * V + na]));
        D[na] = __builtin_ldexp(dsqrt(a * a + bb * bb), e2);
        D[en] = D[na];
        double yr[V], yi[V];
#pragma unroll
        for (int i = 0; i < V; i++) {
          yr[i] = 0.0;
          yi[i] = 0.0;
        }
        yr[en] = 1.0;
        yr[na] = (a - T[en * V + en]) / T[en * V + na];
        yi[na] = bb / T[en * V + na];
        bool skip2 = false;
        static_for<0, na>([&](auto jc) {
          constexpr int i = na - 1 - decltype(jc)::value;
          if (skip2) {
            skip2 = false;
            return;
          }
          bool blk = false;
          if constexpr (i >= 1) blk = T[i * V + i - 1] != 0.0;
          if (blk) {
            if constexpr (i >= 1) {
              constexpr int im = i - 1;
              double s1r = 0.0, s1i = 0.0, s2r = 0.0, s2i = 0.0;
#pragma unroll
              for (int j = i + 1; j <= en; j++) {
                s1r = dfma(T[im * V + j], yr[j], s1r);
                s1i = dfma(T[im * V + j], yi[j], s1i);
                s2r = dfma(T[i * V + j], yr[j], s2r);
                s2i = dfma(T[i * V + j], yi[j], s2i);
              }
              double m11r = T[im * V + im] - a, m11i = -bb, m12 = T[im * V + i];
              double m21 = T[i * V + im], m22r = T[i * V + i] - a, m22i = -bb;
              double detr = (m11r * m22r - m11i * m22i) - m12 * m21;
              double deti = m11r * m22i + m11i * m22r;
              if (detr == 0.0 && deti == 0.0) detr = 2.220446049250313e-16 * tn;
              double n1r = -(s1r * m22r - s1i * m22i) + s2r * m12;
              double n1i = -(s1r * m22i + s1i * m22r) + s2i * m12;
              double n2r = -(s2r * m11r - s2i * m11i) + s1r * m21;
              double n2i = -(s2r * m11i + s2i * m11r) + s1i * m21;
              double dd = detr * detr + deti * deti;
              yr[im] = (n1r * detr + n1i * deti) / dd;
              yi[im] = (n1i * detr - n1r * deti) / dd;
              yr[i] = (n2r * detr + n2i * deti) / dd;
              yi[i] = (n2i * detr - n2r * deti) / dd;
              skip2 = true;
            }
          } else {
            double sr = 0.0, si = 0.0;
#pragma unroll
            for (int j = i + 1; j <= en; j++) {
              sr = dfma(T[i * V + j], yr[j], sr);
              si = dfma(T[i * V + j], yi[j], si);
            }
            double wr = guard_div(T[i * V + i] - a, tn), wi = -bb;
            double dd = wr * wr + wi * wi;
            yr[i] = -(sr * wr + si * wi) / dd;
            yi[i] = -(si * wr - sr * wi) / dd;
          }
        });
#pragma unroll
        for (int r = 0; r < V; r++) {
          Y[r * V + na] = yr[r];
          Y[r * V + en] = yi[r];
        }
        skip = true;
      }
    } else {
      const double lam = T[en * V + en];
      D[en] = __builtin_ldexp(dabs(lam), e2);
      double yv[V];
#pragma unroll
      for (int i = 0; i < V; i++) yv[i] = 0.0;
      yv[en] = 1.0;
      bool skip2 = false;
      static_for<0, en>([&](auto jc) {
        constexpr int i = en - 1 - decltype(jc)::value;
        if (skip2) {
          skip2 = false;
          return;
        }
        bool blk = false;
        if constexpr (i >= 1) blk = T[i * V + i - 1] != 0.0;
        if (blk) {
          if constexpr (i >= 1) {
            constexpr int im = i - 1;
            double s1 = 0.0, s2 = 0.0;
#pragma unroll
            for (int j = i + 1; j <= en; j++) {
              s1 = dfma(T[im * V + j], yv[j], s1);
              s2 = dfma(T[i * V + j], yv[j], s2);
            }
            double m11 = T[im * V + im] - lam, m12 = T[im * V + i];
            double m21 = T[i * V + im], m22 = T[i * V + i] - lam;
            double det = guard_div(m11 * m22 - m12 * m21, tn);
            yv[im] = (-s1 * m22 + s2 * m12) / det;
            yv[i] = (-s2 * m11 + s1 * m21) / det;
            skip2 = true;
          }
        } else {
          double sacc = 0.0;
#pragma unroll
          for (int j = i + 1; j <= en; j++) sacc = dfma(T[i * V + j], yv[j], sacc);
          yv[i] = -sacc / guard_div(T[i * V + i] - lam, tn);
        }
      });
#pragma unroll
      for (int r = 0; r < V; r++) Y[r * V + en] = yv[r];
    }
  });
  // W = Z Y; b = W^-1 x (partial pivoting); b *= D; y = W b
  double W[V * V], LU[V * V], b[V];
#pragma unroll
  for (int i = 0; i < V; i++)
#pragma unroll
    for (int j = 0; j < V; j++) {
      double acc = Zm[i * V] * Y[j];
#pragma unroll
      for (int k = 1; k < V; k++) acc = dfma(Zm[i * V + k], Y[k * V + j], acc);
      W[i * V + j] = acc;
    }
#pragma unroll
  for (int i = 0; i < V * V; i++) LU[i] = W[i];
#pragma unroll
  for (int i = 0; i < V; i++) b[i] = xin[i];
#pragma unroll
  for (int k = 0; k < V; k++) {
    int p = k;
    double best = dabs(LU[k * V + k]);
#pragma unroll
    for (int i = k + 1; i < V; i++) {
      double a = dabs(LU[i * V + k]);
      if (a > best) {
        best = a;
        p = i;
      }
    }
#pragma unroll
    for (int i = k + 1; i < V; i++) {
      if (p == i) {
#pragma unroll
        for (int j = 0; j < V; j++) {
          double t = LU[k * V + j];
          LU[k * V + j] = LU[i * V + j];
          LU[i * V + j] = t;
        }
        double t = b[k];
        b[k] = b[i];
        b[i] = t;
      }
    }
    const double piv_k = guard_div(LU[k * V + k], 1.0);
#pragma unroll
    for (int i = k + 1; i < V; i++) {
      double l = LU[i * V + k] / piv_k;
      LU[i * V + k] = l;
#pragma unroll
      for (int j = k + 1; j < V; j++) LU[i * V + j] = dfma(-l, LU[k * V + j], LU[i * V + j]);
      b[i] = dfma(-l, b[k], b[i]);
    }
  }
#pragma unroll
  for (int i = V - 1; i >= 0; i--) {
    double acc = b[i];
#pragma unroll
    for (int j = i + 1; j < V; j++) acc = dfma(-LU[i * V + j], b[j], acc);
    b[i] = acc / guard_div(LU[i * V + i], 1.0);
  }
#pragma unroll
  for (int i = 0; i < V; i++) b[i] = b[i] * D[i];
#pragma unroll
  for (int i = 0; i < V; i++) {
    double acc = W[i * V] * b[0];
#pragma unroll
    for (int k = 1; k < V; k++) acc = dfma(W[i * V + k], b[k], acc);
    y[i] = acc;
  }
}
#endif
