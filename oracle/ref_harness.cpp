// Test-infrastructure harness around the reference's own C++ core.
//
// Compiled together with the reference sources (where they lie, under the
// reference tree) into oracle/_ref/libpypde_ref.so by oracle/Makefile.  It adds
// stage-level C entry points so that tests can pin each stage of the oracle
// restatement (oracle/oracle.c) against the reference itself.  Nothing here is
// product code; only tests and bench.py's CPU-baseline leg load it.
//
// The reference already exports `pde_solver` and `weno_solver` (src/api.cpp:5-48);
// those are used unchanged.  DGSolver::obj is private in src/solvers/dg/dg.h:259,
// so this one translation unit widens access to reach it.
#include <functional>
#include <iostream>
#include <sstream>
#include <vector>

#include "eigs/system.h"
#include "grid/boundaries.h"
#include "poly/basis.h"
#include "scipy/math/legendre.h"
#include "solvers/dg/dg_matrices.h"
#include "solvers/fv/fluxes.h"
#include "solvers/fv/fv.h"
#include "solvers/stepper.h"
#include "solvers/weno/weno_matrices.h"
#include "types.h"
// Every standard and Eigen header is already included above, so the access
// widening below reaches only the DGSolver class declaration.
#define private public
#include "solvers/dg/dg.h"
#undef private

typedef void (*Ffn)(double *, double *, double *, int);
typedef void (*Bfn)(double *, double *, int);
typedef void (*Sfn)(double *, double *);

double _max_abs_eig(Matr M); // src/eigs/system.cpp:28

extern "C" {

// Basis tables: src/poly/basis.cpp:7-76, src/solvers/weno/weno_matrices.cpp:8-51,
// src/solvers/dg/dg_matrices.cpp:28-55, src/solvers/dg/dg.cpp:38-40.
// Outputs (row-major): nodes[N], wghts[N], endvals[2N], dervals[N*N],
// weno[4*N*N] (mL, mR, mCL, mCR), sig[N*N], dgmat[N*N].
void ref_tables(int N, double *nodes, double *wghts, double *endvals,
                double *dervals, double *weno, double *sig, double *dgmat) {
  std::vector<poly> basis = basis_polys(N);
  Vec x = scaled_nodes(N), w = scaled_weights(N);
  Mat ev = end_values(basis), dv = derivative_values(basis, x);
  int FN2 = (int)floor((N - 1) / 2.), CN2 = (int)ceil((N - 1) / 2.);
  std::vector<Mat> cm = coefficient_matrices(basis, FN2, CN2);
  Mat sg = oscillation_indicator(basis);
  Mat dg = end_value_products(basis) -
           derivative_products(basis, x, w).transpose();
  for (int i = 0; i < N; i++) {
    nodes[i] = x(i);
    wghts[i] = w(i);
    endvals[i] = ev(0, i);
    endvals[N + i] = ev(1, i);
    for (int j = 0; j < N; j++) {
      dervals[i * N + j] = dv(i, j);
      sig[i * N + j] = sg(i, j);
      dgmat[i * N + j] = dg(i, j);
      for (int s = 0; s < 4; s++)
        weno[(s * N + i) * N + j] = cm[s](i, j);
    }
  }
}

// Ghost extension: src/grid/boundaries.cpp:25-53.  ub has prod(nX+2N) rows.
void ref_boundaries(double *ub, double *u, int *nX, int ndim, int *bc, int N,
                    int V) {
  iVecMap nx(nX, ndim), b(bc, ndim);
  MatMap um(u, nx.prod(), V, OuterStride(V));
  Mat r = boundaries(um, nx, b, N);
  for (int i = 0; i < r.rows(); i++)
    for (int v = 0; v < V; v++)
      ub[i * V + v] = r(i, v);
}

// Largest |eigenvalue| of a VxV row-major matrix: src/eigs/system.cpp:28-43.
double ref_max_abs_eig(double *m, int V) {
  Mat M = MatMap(m, V, V, OuterStride(V));
  return _max_abs_eig(M);
}

// FD system matrix dF/dQ + B: src/eigs/system.cpp:6-26, NumericalDiff.h:6-44.
void ref_system_matrix(Ffn F, Bfn B, double *q, double *dq, int d, int ndim,
                       int V, double *out) {
  Vec qq = VecMap(q, V);
  Mat dqq = MatMap(dq, ndim, V, OuterStride(V));
  Mat M = system_matrix(F, B, qq, dqq, d);
  for (int i = 0; i < V; i++)
    for (int j = 0; j < V; j++)
      out[i * V + j] = M(i, j);
}

// CFL timestep over a slab of reconstructed cells: src/solvers/stepper.cpp:35-76.
// wh has nwcells * N^ndim rows of V.
double ref_timestep(Ffn F, Bfn B, double *wh, int nwcells, double *dX,
                    int ndim, int N, int V, double CFL, double tf, double t,
                    int count, int secondOrder) {
  VecMap dx(dX, ndim);
  TimeStepper ts(F, B, dx, N, V, CFL, tf, secondOrder != 0);
  int rows = nwcells * (int)std::pow(N, ndim);
  MatMap w(wh, rows, V, OuterStride(V));
  return ts.step(w, t, count);
}

// Space-time predictor: src/solvers/dg/dg.cpp:204-224.  qh gets nwcells*N^(ndim+1)
// rows of V.
void ref_predictor(Ffn F, Bfn B, Sfn S, double *wh, int nwcells, double *dX,
                   int ndim, int N, int V, double dt, int stiff, double *qh) {
  VecMap dx(dX, ndim);
  DGSolver dg(F, B, S, dx, stiff != 0, N, V);
  int Nd = (int)std::pow(N, ndim);
  MatMap w(wh, nwcells * Nd, V, OuterStride(V));
  Mat q = dg.predictor(w, dt);
  for (int i = 0; i < q.rows(); i++)
    for (int v = 0; v < V; v++)
      qh[i * V + v] = q(i, v);
}

// Residual obj(q) of one cell's space-time system: src/solvers/dg/dg.cpp:128-171.
// w: N^ndim x V reconstruction, q: N^(ndim+1) x V iterate, res: same shape.
void ref_dg_obj(Ffn F, Bfn B, Sfn S, double *w, double *q, double *dX,
                int ndim, int N, int V, double dt, double *res) {
  VecMap dx(dX, ndim);
  DGSolver dg(F, B, S, dx, true, N, V);
  int Nd = (int)std::pow(N, ndim);
  MatMap wm(w, Nd, V, OuterStride(V));
  Mat Ww(N * Nd, V);
  dg.initial_condition(Ww, wm);
  Vec qv = VecMap(q, N * Nd * V);
  Vec r = dg.obj(qv, Ww, dt);
  for (int i = 0; i < r.size(); i++)
    res[i] = r(i);
}

// Corrector: src/solvers/fv/fv.cpp:200-207 (centers, then interfaces) on an
// interior grid u (prod(nX) x V, updated in place) from qh over prod(nX+2) cells.
void ref_fv_apply(Ffn F, Bfn B, Sfn S, double *u, double *qh, int *nX, int ndim,
                  double *dX, int FLUX, int N, int V, double dt,
                  int secondOrder) {
  iVecMap nx(nX, ndim);
  VecMap dx(dX, ndim);
  iVec nx2 = nx;
  nx2.array() += 2;
  int Nd1 = (int)std::pow(N, ndim + 1);
  FVSolver fv(F, B, S, nx, dx, FLUX, N, V, secondOrder != 0);
  MatMap um(u, nx.prod(), V, OuterStride(V));
  MatMap qm(qh, nx2.prod() * Nd1, V, OuterStride(V));
  fv.apply(um, qm, dt);
}

// One face-point flux: src/solvers/fv/fluxes.cpp:80-101 (+ Bint :103-116).
void ref_flux(Ffn F, Bfn B, double *qL, double *qR, int d, int ndim, int N,
              int V, int FLUX, double *f, double *bint) {
  Vec x = scaled_nodes(N), w = scaled_weights(N);
  Vec dx = Vec::Ones(ndim);
  FluxGenerator fg(F, B, x, w, dx, V, FLUX, ndim);
  Vec l = VecMap(qL, V), r = VecMap(qR, V), out(V), b(V);
  Mat dq0 = Mat::Zero(ndim, V), dq1 = Mat::Zero(ndim, V);
  if (F) {
    fg.flux(out, l, r, dq0, dq1, d, false);
    for (int v = 0; v < V; v++)
      f[v] = out(v);
  }
  if (B) {
    fg.Bint(b, l, r, d);
    for (int v = 0; v < V; v++)
      bint[v] = b(v);
  }
}
}
