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
// One line window (2N-1 values, `st` apart) -> N nodal values (weno.cpp:26-59).
template <typename Get>
DEV void weno_line(Get get, double* out) {
  double num[N];
#pragma unroll
  for (int k = 0; k < N; k++) num[k] = 0.0;
  double den = 0.0;
#pragma unroll
  for (int s = 0; s < T_NSTEN; s++) {
    double c[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
      double acc = T_WINV[s][k][0] * get(T_SOFF[s]);
#pragma unroll
      for (int j = 1; j < N; j++) acc = dfma(T_WINV[s][k][j], get(T_SOFF[s] + j), acc);
      c[k] = acc;
    }
    double p = 0.0;
#pragma unroll
    for (int k = 0; k < N; k++) {
      double t = T_SIG[k][0] * c[0];
#pragma unroll
      for (int j = 1; j < N; j++) t = dfma(T_SIG[k][j], c[j], t);
      p = (k == 0) ? c[0] * t : dfma(c[k], t, p);
    }
    p = p + WENO_EPS;
    double p2 = p * p;
    double p4 = p2 * p2;
    double p8 = p4 * p4;
    double o = T_SLAM[s] / p8;
#pragma unroll
    for (int k = 0; k < N; k++) num[k] = dfma(o, c[k], num[k]);
    den = den + o;
  }
#pragma unroll
  for (int k = 0; k < N; k++) out[k] = num[k] / den;
}

// Ghost index map (boundaries.cpp:7-23); code 2 = reflective wall (this
// build's addition).  Returns the interior index, sets *flip for mirrors.
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
  double mu = dsqrt(x[0] * x[0] + sigma);
  double v0 = (x[0] <= 0.0) ? x[0] - mu : -sigma / (x[0] + mu);
  v[0] = v0;
#pragma unroll
  for (int i = 1; i < M; i++) v[i] = x[i];
  return 2.0 / (v0 * v0 + sigma);
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

// element H[r][c] for run-time r, c with every index static (stays in registers)
DEV double hget(const double* H, int r, int c) {
  double out = 0.0;
#pragma unroll
  for (int i = 0; i < V; i++)
#pragma unroll
    for (int j = 0; j < V; j++)
      if (i == r && j == c) out = H[i * V + j];
  return out;
}

// No eigenvalue of the undeflated leading block H[0..hi][0..hi] (shifted back
// by ex) can exceed `best` in modulus when its infinity norm + |ex| <= best:
// the largest modulus is then final.
DEV bool rest_bounded(const double* H, int hi, double ex, double best) {
  if (hi < 0) return false;
  double bound = 0.0;
#pragma unroll
  for (int i = 0; i < V; i++) {
    if (i <= hi) {
      double rs = 0.0;
#pragma unroll
      for (int j = 0; j < V; j++)
        if (j <= hi) rs = rs + dabs(H[i * V + j]);
      bound = dmax(bound, rs);
    }
  }
  return bound + dabs(ex) <= best;
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
  double scale = 0.0;
#pragma unroll
  for (int i = 0; i < V * V; i++) scale = dmax(scale, dabs(H[i]));
  if (!(scale >= 2.2250738585072014e-308)) return (scale != scale) ? scale : 0.0;
  int e2;
  (void)__builtin_frexp(scale, &e2);  // scale = m 2^e2, m in [0.5, 1)
  const double pw = __builtin_ldexp(1.0, -e2);
#pragma unroll
  for (int i = 0; i < V * V; i++) H[i] = H[i] * pw;
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
    }
  });
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
      best = dmax(best, dabs(hget(H, hi, hi) + ex));
      hi -= 1;
      iter = 0;
      if (rest_bounded(H, hi, ex, best)) break;
      continue;
    }
    if (lo == hi - 1) {
      double a = hget(H, hi - 1, hi - 1) + ex, b = hget(H, hi - 1, hi);
      double c = hget(H, hi, hi - 1), dd = hget(H, hi, hi) + ex;
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
      if (rest_bounded(H, hi, ex, best)) break;
      continue;
    }
    if constexpr (V >= 3) {
      if (total >= 40 * V) {
#pragma unroll
        for (int i = 0; i < V; i++)
          if (i <= hi) best = dmax(best, dabs(H[i * V + i] + ex));
        break;
      }
      double sx = hget(H, hi, hi), sy = hget(H, hi - 1, hi - 1);
      double sw = hget(H, hi, hi - 1) * hget(H, hi - 1, hi);
      if (iter == 10) {
        ex = ex + sx;
#pragma unroll
        for (int i = 0; i < V; i++)
          if (i <= hi) H[i * V + i] = H[i * V + i] - sx;
        double s = dabs(hget(H, hi, hi - 1)) + dabs(hget(H, hi - 1, hi - 2));
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
      double h00 = hget(H, lo, lo);
      double r = sx - h00, q = sy - h00;
      double x[3], v[3];
      x[0] = (r * q - sw) / hget(H, lo + 1, lo) + hget(H, lo, lo + 1);
      x[1] = hget(H, lo + 1, lo + 1) - h00 - r - q;
      x[2] = hget(H, lo + 2, lo + 1);
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
  return __builtin_ldexp(best, e2);
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

// A = dF/dQ by forward differences + B(q, d) (system.cpp:6-26,
// NumericalDiff.h:6-44), column i scaled by 1 / h_i (one division per column);
// f0 receives F(q).
DEV void system_matrix(const double* q_in, const double* dq, int d, double* A, double* f0) {
  double q[V];
#pragma unroll
  for (int i = 0; i < V; i++) q[i] = q_in[i];
#if PDE_USE_F
  double v1[V], v2[V];
  F(v1, q, dq, d);
#pragma unroll
  for (int i = 0; i < V; i++) {
    double qi0 = q[i];
    double h = dmax(FD_EPS * dabs(qi0), FD_EPS);
    q[i] = qi0 + h;
    F(v2, q, dq, d);
    q[i] = qi0;
    double rh = 1.0 / h;
#pragma unroll
    for (int r = 0; r < V; r++) A[r * V + i] = (v2[r] - v1[r]) * rh;
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
