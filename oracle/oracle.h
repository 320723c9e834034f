/* CPU restatement of the reference ADER-WENO solver (haranjackson/PyPDE, src/).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Every function follows the reference algorithm cited in oracle.c and fixes one
 * floating-point operation order (explicit fma() in every dot-product chain,
 * everything else evaluated left to right with -ffp-contract=off).  The HIP
 * kernels issue the same operations in the same order, which is what makes
 * bit-level agreement possible despite the reference's finite-difference
 * Jacobians (a last-bit change in a state moves a wave speed by ~1e-8).
 */
#ifndef PYPDE_ORACLE_H
#define PYPDE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef void (*oracle_F)(double *out, double *Q, double *dQ, int d);
typedef void (*oracle_B)(double *out, double *Q, int d);
typedef void (*oracle_S)(double *out, double *Q);
typedef void (*oracle_M)(double *out, double *Q);

typedef struct {
  double newton_tol;     /* stiff predictor: stop when |dq| <= tol (1 + |q|) elementwise */
  int newton_maxit;      /* stiff predictor iteration cap */
  long max_steps;        /* stop after this many steps even if t < tf (<= 0: no cap) */
  const double *reflect; /* ndim*V multipliers for reflective walls (NULL: all +1) */
  const double *guess;   /* predictor initial guesses, nw x N^(D+1) x V (NULL: w broadcast) */
  int guess_check;       /* with guess: fall back to the plain guess when the residual grows */
  int b_quirk;           /* 1: evaluate B.dQ as the reference does (dg.cpp:110, fv.cpp:234) */
} oracle_opts;

typedef struct {
  long steps;
  double t;
  long pred_iters;   /* predictor iterations summed over cells and steps */
  long pred_cells;   /* cell solves summed over steps */
  long pred_maxit;   /* largest per-cell iteration count */
  long pred_capped;  /* cell solves that reached the iteration cap */
  long guess_fallbacks;
} oracle_stats;

/* Tables for order N (N Gauss-Legendre nodes).  Any pointer may be NULL.
 * weno_inv: 4 x N x N inverses of the stencil matrices (L, R, CL, CR);
 * fl_c: c_0..c_{N-1} and fl_m: M_1..M_N (N x N each) of the Faddeev-LeVerrier
 * expansion of Tw = diag(1/W) DG_MAT. */
int oracle_tables(int N, double *nodes, double *wghts, double *endvals,
                  double *dervals, double *weno_mat, double *weno_inv,
                  double *sig, double *dgmat, double *dginv, double *fl_c,
                  double *fl_m);

double oracle_exp(double x);
double oracle_max_abs_eig(const double *A, int V);
/* y = |A| x = R |Lambda| R^-1 x (Roe and Osher dissipation, fluxes.cpp:20-70) */
void oracle_abs_apply(const double *A, const double *x, int V, double *y);
void oracle_system_matrix(oracle_F F, oracle_B B, const double *q,
                          const double *dq, int d, int ndim, int V, double *out);

/* ub = ghost extension of u by N cells per side (bc: 0 transmissive,
 * 1 periodic, 2 reflective). */
int oracle_boundaries(double *ub, const double *u, const int *nX, int ndim,
                      const int *bc, const double *reflect, int N, int V);
/* Reference weno_solver semantics: u has shape nX (no ghosts); the result has
 * shape (nX - 2(N-1)) x N^ndim x V. */
int oracle_weno(double *wh, const double *u, const int *nX, int ndim, int N,
                int V);
double oracle_timestep(oracle_F F, oracle_B B, const double *wh, long nw,
                       const double *dX, int ndim, int N, int V, double CFL,
                       double tf, double t, long count);
/* qh: nw x N^(ndim+1) x V.  iters (optional): per-cell iteration counts. */
int oracle_predictor(oracle_F F, oracle_B B, oracle_S S, oracle_M M,
                     const double *wh, long nw, const double *dX, int ndim,
                     int N, int V, double dt, int stiff, const oracle_opts *opts,
                     double *qh, int *iters);
/* res = obj(q) (reference dg.cpp:128-154) for one cell. */
int oracle_dg_obj(oracle_F F, oracle_B B, oracle_S S, const double *w,
                  const double *q, const double *dX, int ndim, int N, int V,
                  double dt, int b_quirk, double *res);
int oracle_fv(oracle_F F, oracle_B B, oracle_S S, double *u, const double *qh,
              const int *nX, int ndim, const double *dX, int flux, int N, int V,
              double dt, int b_quirk);
/* Full solve with the reference's pde_solver argument list (api.h:4-10) plus
 * the stiff-model Jacobian M, options and statistics.  Returns 0, or -(step+1)
 * when NaNs appear, or a positive error code for bad arguments. */
int oracle_pde_solver(oracle_F F, oracle_B B, oracle_S S, oracle_M M, int useF,
                      int useB, int useS, double *u, double tf, const int *nX,
                      int ndim, const double *dX, double CFL, const int *bc,
                      int stiff, int flux, int N, int V, int ndt,
                      int secondOrder, double *ret, const oracle_opts *opts,
                      oracle_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
