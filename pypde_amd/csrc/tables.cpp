// Basis, WENO and DG tables for order N, computed once per solve on the host
// in long double and rounded once to double.  They replace the reference's
// setup code (scipy/math/legendre.cpp:136-197, poly/basis.cpp:7-76,
// solvers/weno/weno_matrices.cpp:8-51, solvers/dg/dg_matrices.cpp:4-55,
// solvers/dg/dg.cpp:22-48).  Then written into the kernel source as exact
// hexadecimal constants.
#include "tables.h"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <sstream>
#include <string>

namespace pypde {

typedef long double ld;
static const int MAXN = TABLE_MAXN;

// Gauss-Legendre nodes and weights by Newton iteration on P_N, ascending,
// mapped to [0, 1].
static void gl_nodes(int n, ld* x, ld* w) {
  for (int i = 0; i < n; i++) {
    ld z = cosl(3.14159265358979323846264338327950288L * ((ld)i + 0.75L) / ((ld)n + 0.5L));
    ld dp = 1.0L;
    for (int it = 0; it < 100; it++) {
      ld p0 = 1.0L, p1 = z;
      for (int k = 2; k <= n; k++) {
        ld p2 = ((ld)(2 * k - 1) * z * p1 - (ld)(k - 1) * p0) / (ld)k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p0 = 1.0L;
      dp = (ld)n * (z * p1 - p0) / (z * z - 1.0L);
      ld z1 = z;
      z = z1 - p1 / dp;
      if (fabsl(z - z1) <= 4.0L * LDBL_EPSILON) break;
    }
    {
      ld p0 = 1.0L, p1 = z;
      for (int k = 2; k <= n; k++) {
        ld p2 = ((ld)(2 * k - 1) * z * p1 - (ld)(k - 1) * p0) / (ld)k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p0 = 1.0L;
      dp = (ld)n * (z * p1 - p0) / (z * z - 1.0L);
    }
    x[n - 1 - i] = (1.0L + z) / 2.0L;
    w[n - 1 - i] = 1.0L / ((1.0L - z * z) * dp * dp);
  }
}

// power-basis coefficients of the Lagrange polynomials at the nodes
static void lagrange_coeffs(int n, const ld* x, ld c[MAXN][MAXN]) {
  for (int j = 0; j < n; j++) {
    ld p[MAXN];
    for (int k = 0; k < n; k++) p[k] = 0.0L;
    p[0] = 1.0L;
    int deg = 0;
    for (int m = 0; m < n; m++) {
      if (m == j) continue;
      ld den = x[j] - x[m];
      for (int k = deg + 1; k >= 1; k--) p[k] = (p[k - 1] - x[m] * p[k]) / den;
      p[0] = (-x[m] * p[0]) / den;
      deg++;
    }
    for (int k = 0; k < n; k++) c[j][k] = p[k];
  }
}

static ld poly_eval(const ld* c, int len, ld s) {
  ld r = 0.0L;
  for (int k = len - 1; k >= 0; k--) r = r * s + c[k];
  return r;
}

static ld poly_int(const ld* c, int n, ld a) {
  ld P[MAXN + 1];
  P[0] = 0.0L;
  for (int k = 0; k < n; k++) P[k + 1] = c[k] / (ld)(k + 1);
  return poly_eval(P, n + 1, a + 1.0L) - poly_eval(P, n + 1, a);
}

static int ld_inverse(int n, ld A[MAXN][MAXN], ld R[MAXN][MAXN]) {
  ld M[MAXN][2 * MAXN];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < 2 * n; j++) M[i][j] = j < n ? A[i][j] : (ld)(j - n == i);
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int i = k + 1; i < n; i++)
      if (fabsl(M[i][k]) > fabsl(M[p][k])) p = i;
    if (M[p][k] == 0.0L) return -1;
    if (p != k)
      for (int j = 0; j < 2 * n; j++) {
        ld t = M[k][j];
        M[k][j] = M[p][j];
        M[p][j] = t;
      }
    ld piv = M[k][k];
    for (int j = 0; j < 2 * n; j++) M[k][j] /= piv;
    for (int i = 0; i < n; i++) {
      if (i == k) continue;
      ld f = M[i][k];
      for (int j = 0; j < 2 * n; j++) M[i][j] -= f * M[k][j];
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) R[i][j] = M[i][n + j];
  return 0;
}

int make_tables(int n, Tables* T) {
  if (n < 1 || n > MAXN) return 1;
  T->N = n;
  ld x[MAXN], w[MAXN], c[MAXN][MAXN];
  gl_nodes(n, x, w);
  lagrange_coeffs(n, x, c);
  for (int i = 0; i < n; i++) {
    T->x[i] = (double)x[i];
    T->w[i] = (double)w[i];
  }
  ld e1[MAXN];
  for (int j = 0; j < n; j++) {
    T->end0[j] = (double)poly_eval(c[j], n, 0.0L);
    e1[j] = poly_eval(c[j], n, 1.0L);
    T->end1[j] = (double)e1[j];
  }
  ld dc[MAXN][MAXN];
  for (int j = 0; j < n; j++) {
    for (int k = 0; k < n; k++) dc[j][k] = 0.0L;
    for (int k = 1; k < n; k++) dc[j][k - 1] = c[j][k] * (ld)k;
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) T->der[i][j] = (double)poly_eval(dc[j], n, x[i]);
  int FN2 = (n - 1) / 2, CN2 = n / 2;
  int off[4] = {-(n - 1), 0, -CN2, -FN2};
  for (int s = 0; s < 4; s++) {
    ld A[MAXN][MAXN], R[MAXN][MAXN];
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) A[i][j] = poly_int(c[j], n, (ld)(i + off[s]));
    if (ld_inverse(n, A, R)) return 2;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        T->wmat[s][i][j] = (double)A[i][j];
        T->winv[s][i][j] = (double)R[i][j];
      }
  }
  {
    ld S[MAXN][MAXN];
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) S[i][j] = 0.0L;
    ld di[MAXN][MAXN];
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++) di[j][k] = c[j][k];
    for (int a = 1; a < n; a++) {
      for (int j = 0; j < n; j++) {
        for (int k = 1; k < n; k++) di[j][k - 1] = di[j][k] * (ld)k;
        di[j][n - 1] = 0.0L;
      }
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
          ld acc = 0.0L;
          for (int p = 0; p < n; p++)
            for (int q = 0; q < n; q++) acc += di[i][p] * di[j][q] / (ld)(p + q + 1);
          S[i][j] += acc;
        }
    }
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) T->sig[i][j] = (double)S[i][j];
  }
  ld G[MAXN][MAXN], GI[MAXN][MAXN];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      ld acc = 0.0L;
      for (int p = 0; p < n; p++)
        for (int q = 0; q < n; q++) acc += dc[i][p] * c[j][q] / (ld)(p + q + 1);
      G[i][j] = e1[i] * e1[j] - acc;
    }
  if (ld_inverse(n, G, GI)) return 3;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      T->dgmat[i][j] = (double)G[i][j];
      T->dginv[i][j] = (double)GI[i][j];
    }
  // Faddeev-LeVerrier on Tw = diag(1/W) DG_MAT (stiff predictor Newton solve)
  {
    ld Tw[MAXN][MAXN], Mk[MAXN][MAXN], TM[MAXN][MAXN], cc[MAXN + 1];
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) Tw[i][j] = G[i][j] / w[i];
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) Mk[i][j] = (ld)(i == j);
    cc[n] = 1.0L;
    for (int k = 1; k <= n; k++) {
      if (k > 1) {
        for (int i = 0; i < n; i++)
          for (int j = 0; j < n; j++) Mk[i][j] = TM[i][j] + (i == j ? cc[n - k + 1] : 0.0L);
      }
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
          ld acc = 0.0L;
          for (int p = 0; p < n; p++) acc += Tw[i][p] * Mk[p][j];
          TM[i][j] = acc;
        }
      ld tr = 0.0L;
      for (int i = 0; i < n; i++) tr += TM[i][i];
      cc[n - k] = -tr / (ld)k;
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) T->flm[k - 1][i][j] = (double)Mk[i][j];
    }
    for (int k = 0; k < n; k++) T->flc[k] = (double)cc[k];
  }
  // stencil list in the reference's order (weno.cpp:47-55)
  T->nsten = 2;
  T->soff[0] = 0;
  T->slam[0] = 1.0;
  T->soff[1] = n - 1;
  T->slam[1] = 1.0;
  T->soff[2] = T->soff[3] = 0;
  T->slam[2] = T->slam[3] = 0.0;
  if (n > 2) {
    T->soff[2] = FN2;
    T->slam[2] = 1e5;
    T->nsten = 3;
    if (n % 2 == 0) {
      T->soff[3] = CN2;
      T->slam[3] = 1e5;
      T->nsten = 4;
    }
  }
  return 0;
}

static std::string hx(double v) {
  char b[64];
  snprintf(b, sizeof b, "%a", v);
  return b;
}

std::string tables_source(const Tables& T, int D) {
  int n = T.N;
  std::ostringstream o;
  auto arr1 = [&](const char* name, const double* a, int len) {
    o << "__device__ constexpr double " << name << "[" << len << "] = {";
    for (int i = 0; i < len; i++) o << (i ? ", " : "") << hx(a[i]);
    o << "};\n";
  };
  auto arr2 = [&](const char* name, const double a[TABLE_MAXN][TABLE_MAXN]) {
    o << "__device__ constexpr double " << name << "[" << n << "][" << n << "] = {";
    for (int i = 0; i < n; i++) {
      o << (i ? ", {" : "{");
      for (int j = 0; j < n; j++) o << (j ? ", " : "") << hx(a[i][j]);
      o << "}";
    }
    o << "};\n";
  };
  arr1("T_X", T.x, n);
  arr1("T_W", T.w, n);
  arr1("T_END0", T.end0, n);
  arr1("T_END1", T.end1, n);
  arr2("T_DER", T.der);
  arr2("T_SIG", T.sig);
  arr2("T_DGMAT", T.dgmat);
  arr2("T_DGINV", T.dginv);
  arr1("T_FLC", T.flc, n);
  o << "__device__ constexpr double T_FLM[" << n << "][" << n << "][" << n << "] = {";
  for (int k = 0; k < n; k++) {
    o << (k ? ", {" : "{");
    for (int i = 0; i < n; i++) {
      o << (i ? ", {" : "{");
      for (int j = 0; j < n; j++) o << (j ? ", " : "") << hx(T.flm[k][i][j]);
      o << "}";
    }
    o << "}";
  }
  o << "};\n";
  o << "__device__ constexpr double T_WINV[4][" << n << "][" << n << "] = {";
  for (int s = 0; s < 4; s++) {
    o << (s ? ", {" : "{");
    for (int i = 0; i < n; i++) {
      o << (i ? ", {" : "{");
      for (int j = 0; j < n; j++) o << (j ? ", " : "") << hx(T.winv[s][i][j]);
      o << "}";
    }
    o << "}";
  }
  o << "};\n";
  o << "constexpr int T_NSTEN = " << T.nsten << ";\n";
  o << "__device__ constexpr int T_SOFF[4] = {" << T.soff[0] << ", " << T.soff[1] << ", "
    << T.soff[2] << ", " << T.soff[3] << "};\n";
  arr1("T_SLAM", T.slam, 4);
  // per-node quadrature weights prod_d W[k_d] (stepper.cpp:31-40), left to right
  int nd = 1;
  for (int d = 0; d < D; d++) nd *= n;
  o << "__device__ constexpr double T_WP[" << nd << "] = {";
  for (int idx = 0; idx < nd; idx++) {
    int k[3], r = idx;
    for (int d = D - 1; d >= 0; d--) {
      k[d] = r % n;
      r /= n;
    }
    double p = T.w[k[0]];
    for (int d = 1; d < D; d++) p = p * T.w[k[d]];
    o << (idx ? ", " : "") << hx(p);
  }
  o << "};\n";
  return o.str();
}

}  // namespace pypde
