/* pypde_hip: MI355X-native ADER-WENO solver, C ABI.
 *
 * Drop-in for the compute path of haranjackson/PyPDE (libpypde.so).  The
 * physics arrive as HIP device-function source (F, B, S and optionally
 * M = dS/dQ; signatures in pypde_amd/systems/pde_prelude.h) instead of C
 * function pointers, and are compiled with hipRTC together with kernels
 * specialised for (ndim, N, V).
 *
 * Return codes: 0 success, >0 argument / compile / runtime errors (text via
 * pypde_hip_last_error), <0 NaN detected at step -rc.  A run stops at its
 * first NaN step; the reference instead prints "NaNs found", writes the
 * previous and the NaN state into rows pushCount and pushCount+1 of ret, and
 * keeps stepping (iterator.cpp:141-147).  Rows pushed before the NaN step
 * and the final row (the NaN state) agree with the reference.
 */
#ifndef PYPDE_HIP_H
#define PYPDE_HIP_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYPDE_OK 0
#define PYPDE_ERR_ARG 1
#define PYPDE_ERR_COMPILE 2
#define PYPDE_ERR_HIP 3
#define PYPDE_ERR_UNSUPPORTED 4

/* Options beyond the reference's argument list (defaults in brackets). */
typedef struct {
  int struct_size;          /* sizeof(pypde_options) */
  double newton_tol;        /* stiff predictor: |dq| <= tol (1 + |q|) elementwise [1e-12] */
  int newton_maxit;         /* stiff predictor iteration cap [100] */
  long max_steps;           /* stop after this many steps even if t < tf; <= 0: none [0] */
  const double *reflect;    /* ndim*V multipliers for reflective walls (bc 2) [all +1] */
  int b_quirk;              /* 1: B.dQ evaluated as the reference does (dg.cpp:110,
                               fv.cpp:234) instead of the matrix-vector product [0] */
  int use_M;                /* the source defines M(out, Q) = dS/dQ [0: finite differences] */
  int device;               /* HIP device ordinal [0] */
  int profile;              /* time every kernel with HIP events [0] */
  const double *guess;      /* host array: predictor initial guesses for the first step,
                               nwc x N^(ndim+1) x V (testing the learned-guess path) */
  int use_graph;            /* capture one timestep in a HIP graph and replay it [0]
                               (ignored while profiling) */
  int mixed;                /* mixed precision: an FP32 phase of the predictor (contractions,
                               F and S in FP32; stiff Newton solves with the FP64 source
                               Jacobian) runs until its corrections fall below mixed_tol,
                               then the FP64 predictor iterates on from that iterate to the
                               unchanged tolerance [0].  Systems with F, S (and M) only */
  double mixed_tol;         /* FP32 phase: |dq| <= mixed_tol (1 + |q|) [6e-6] */
  int mixed_maxit;          /* FP32 phase iteration cap [20] */
} pypde_options;

typedef struct {
  int struct_size;          /* sizeof(pypde_stats) */
  long steps;
  double t;
  long pred_iters;          /* predictor iterations summed over cells and steps */
  long pred_cells;
  long pred_maxit;
  long pred_capped;         /* cell solves that hit the iteration cap */
  long guess_fallbacks;     /* learned guesses abandoned for the plain guess */
  int nan_step;             /* first step with NaNs (0: none) */
  double ms_loop;           /* wall time of the step loop */
  double ms_kernel[8];      /* profile: summed time per kernel (weno, dt, learned guess,
                               predictor, flux0, flux1, update; the last slot is kept
                               for layout and holds only the cost of one event pair,
                               the step's bookkeeping running inside dt and update) */
  long pred_node_full;      /* stiff predictor: per-node Newton corrections that solved
                               with the source Jacobian, summed over nodes and steps */
  long pred_node_short;     /* ... and those that took the A = 0 shortcut (max |dt dS/dQ|
                               < 1e-6); both counts run from t = 0 like pred_iters */
  long pred_iters_f32;      /* mixed precision: FP32-phase iterations summed over cells */
} pypde_stats;

void pypde_default_options(pypde_options *opts);

/* One full solve.  Argument list of the reference's pde_solver (api.h:4-10)
 * with the three callbacks replaced by `src`, followed by options/statistics.
 * u (prod(nX) x V, row-major, variables fastest) is overwritten with the final
 * state, as in the reference (api.cpp:18, iterator.cpp:129); ret receives ndt
 * snapshot rows (iterator.cpp:136-150).  FLUX: 0 Rusanov, 1 Roe, 2 Osher
 * (fluxes.h:6-8).  secondOrder: F depends on the gradient dQ (viscous terms;
 * face gradients and viscous wave speeds as fv.cpp:87-122, system.cpp:53-61). */
int pde_solver_hip(const char *src, bool useF, bool useB, bool useS, double *u, double tf,
                   int *nX, int ndim, double *dX, double CFL, int *boundaryTypes, bool STIFF,
                   int FLUX, int N, int V, int ndt, bool secondOrder, double *ret,
                   const pypde_options *opts, pypde_stats *stats);

/* GPU replacement of the reference's weno_solver (api.h:12-13, api.cpp:32-48):
 * same arguments and semantics (no ghost cells; ret has shape
 * (nX - 2(N-1)) x N^ndim x V).  The reference returns void; this returns an
 * int status, 0 on success. */
int weno_solver(double *ret, double *u, int *nX, int ndim, int N, int V);

typedef struct pypde_ctx pypde_ctx;

/* Allocation-free variant of weno_solver, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  u and ret are
 * device pointers in weno_solver's layouts (u: prod(nX) x V, ret:
 * prod(nX - 2(N-1)) x N^ndim x V); ndim, N and V are the context's, whose
 * loaded kernels it launches.  Its staging (u as planes, the reconstruction
 * planes) lives in one 64 MiB device array per GPU, held until the process
 * ends: 8 (V prod(nX) + N^ndim V prod(nX - 2(N-1))) bytes (the first part
 * rounded up to 256 bytes) must fit, else PYPDE_ERR_ARG.  It calls no
 * hipMalloc or hipFree.  With stream NULL it runs on the context's stream and
 * returns when done; otherwise it only enqueues on `stream`.  Every call on
 * one GPU shares that array: calls that overlap on one GPU race. */
int weno_solver_noalloc(pypde_ctx *c, double *ret, const double *u, const int *nX,
                        void *stream);

/* Persistent solver context: compile once, keep the state resident in HBM. */
pypde_ctx *pypde_create(const char *src, bool useF, bool useB, bool useS, int *nX, int ndim,
                        double *dX, double CFL, int *boundaryTypes, bool STIFF, int FLUX, int N,
                        int V, bool secondOrder, const pypde_options *opts, int *rc);
int pypde_set_state(pypde_ctx *c, const double *u);          /* host, row-major */
int pypde_get_state(pypde_ctx *c, double *u);                /* host, row-major */
/* Advance from t = 0 until tf or max_steps; ndt snapshot rows into ret (may be
 * NULL when ndt == 0).  Device state stays resident between calls. */
int pypde_run(pypde_ctx *c, double tf, long max_steps, int ndt, double *ret, pypde_stats *stats);
/* Advance `steps` more steps from the current (t, step count). */
int pypde_step(pypde_ctx *c, long steps, double tf, pypde_stats *stats);
/* Device pointers of the resident state (structure of arrays, [V][cell]) and
 * of the last reconstruction/predictor outputs; stream used for all launches. */
void *pypde_device_state(pypde_ctx *c);
void *pypde_stream(pypde_ctx *c);
void pypde_destroy(pypde_ctx *c);

/* Ensembles: n independent grids of one system (same ndim, order, nvar,
 * physics and boundary types), grid i with nX[i*ndim ..] cells and spacing
 * dX[i*ndim ..], each with its own dt.  Every timestep of all members is one
 * HIP graph launch (each member on its own stream inside the graph), so small
 * grids run side by side.  No reference counterpart: the reference solves
 * one grid per pde_solver call (api.cpp:9-30). */
typedef struct pypde_ensemble pypde_ensemble;
pypde_ensemble *pypde_ensemble_create(int n, const char *src, bool useF, bool useB, bool useS,
                                      const int *nX, int ndim, const double *dX, double CFL,
                                      const int *boundaryTypes, bool STIFF, int FLUX, int N, int V,
                                      bool secondOrder, const pypde_options *opts, int *rc);
int pypde_ensemble_set_state(pypde_ensemble *e, int i, const double *u);
int pypde_ensemble_get_state(pypde_ensemble *e, int i, double *u);
/* Advance every member from t = 0 to tf (or max_steps); stats: n entries or NULL. */
int pypde_ensemble_run(pypde_ensemble *e, double tf, long max_steps, pypde_stats *stats,
                       double *ms_loop);
void pypde_ensemble_destroy(pypde_ensemble *e);

/* Learned Newton initial guess (no counterpart in the reference; its hook is
 * DGSolver::initial_guess, dg.cpp:12-20,216).  A 3-layer SiLU MLP, weights as
 * device fp32 arrays (Wt: out x in, row-major), inputs (wh, dt) normalised by
 * in_mean / in_rstd, outputs (qh - w) / dt de-normalised by out_std / out_mean.
 * Once attached, every step infers guesses on the matrix cores (BF16, FP32
 * accumulate) before the stiff predictor; Newton still runs to newton_tol and
 * abandons a guess whose residual grows.  The weights are read (packed into
 * BF16 fragments) when the model is attached: attach again after changing
 * them.  NULL detaches. */
typedef struct {
  const float *W1t, *b1, *W2t, *b2, *W3t, *b3;
  const float *in_mean, *in_rstd, *out_mean, *out_std;
} pypde_mlp;
int pypde_set_guess_model(pypde_ctx *c, const pypde_mlp *m);
/* One step that also writes training pairs as fp32 device columns col0..:
 * X (V N^ndim + 1 rows: wh, dt) and Y (V N^(ndim+1) rows: (qh - w) / dt), leading
 * dimension ldx. */
int pypde_harvest_step(pypde_ctx *c, float *X, float *Y, long ldx, long col0, pypde_stats *stats);
/* Autotuner (no counterpart in the reference, whose only partition knob is its
 * thread-slab split, iterator.cpp:68-84): before its first step a context
 * times every launch variant of its tunable kernels in real trial steps
 * (state restored afterwards) and keeps the fastest, or reuses the choice
 * recorded for the same problem in <code-object cache>/tuning.txt.
 * choices[5]: WENO tile (0 8x32, 1 4x64, 2 16x16, 3 16x32), flux workgroup (0
 * 256, 1 128, 2 64), corrector (0 one thread per cell, 1..4 one variable and
 * 1, 2, 4, 8 x-rows per thread), predictor (0 static, 1 work queue, 2 the
 * stiff predictor's cold/hot split, 3 the same with the cold pass built for
 * three wavefronts per SIMD), source (0 defaults, 1 measured, 2 tuning table,
 * 3 PYPDE_TUNE_FORCE); ms[4]: the chosen WENO, predictor, flux (both
 * directions) and corrector trial times. */
int pypde_get_tuning(pypde_ctx *c, int *choices, double *ms);

/* Host copy of the last inferred guesses, nwc x N^(ndim+1) x V. */
int pypde_last_guess(pypde_ctx *c, double *out);
int pypde_geometry(pypde_ctx *c, long *nwc, int *n_in, int *n_out);

/* Diagnostics: the GPU wave-speed routines on n host matrices (V x V,
 * row-major, V of the context's system), for checking them against a CPU
 * restatement.  which 0: max |lambda| of each matrix (the eigenvalue step of
 * the reference's max_abs_eigs, system.cpp:28-43), out[n]; 2: |A_i| x_i (the
 * |A| product of the Roe and Osher fluxes, fluxes.cpp:20-70; a context with
 * FLUX 1 or 2), x and out n x V.  No counterpart in the reference's C API. */
int pypde_eig_batch(pypde_ctx *c, int which, const double *A, const double *x, long n,
                    double *out);

/* Training kernels (device fp32 pointers; stream = HIP stream or NULL):
 * C = alpha op(A) op(B) + beta C on the matrix cores (BF16 operands, FP32
 * accumulate); Z += b, H = SiLU(Z) (act 1) or Z (act 0); dZ = dH SiLU'(Z);
 * db = row sums; MSE loss and gradient; Adam step. */
int pypde_mlp_gemm(const float *A, int lda, int transA, const float *B, int ldb, int transB,
                   float *C, int ldc, int M, int N, int K, float alpha, float beta, void *stream);
int pypde_mlp_bias_act(float *Z, const float *b, float *H, int rows, int cols, int act,
                       void *stream);
int pypde_mlp_act_grad(float *dZ, const float *dH, const float *Z, long n, void *stream);
int pypde_mlp_rowsum(const float *dZ, float *db, int rows, int cols, void *stream);
int pypde_mlp_mse(const float *Y, const float *T, float *dY, float *loss, int rows, int cols,
                  void *stream);
int pypde_mlp_adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1,
                   float b2, float eps, int step, void *stream);
int pypde_mlp_compile(void);

/* Device memory for callers without their own HIP runtime (kind: 1 host to
 * device, 2 device to host, 3 device to device). */
void *pypde_dev_alloc(size_t bytes);
int pypde_dev_free(void *p);
int pypde_memcpy(void *dst, const void *src, size_t bytes, int kind);
int pypde_dev_memset(void *p, int value, size_t bytes);
int pypde_dev_sync(void);
int pypde_device_count(void);

/* Compile (or fetch from the code-object cache) without touching a GPU; FLUX and
 * secondOrder as in pde_solver_hip. */
int pypde_hip_compile(const char *src, bool useF, bool useB, bool useS, int ndim, int N, int V,
                      bool STIFF, int FLUX, bool secondOrder, const pypde_options *opts);
/* The same for the code object of ensemble members, which also holds the batched
 * ensemble kernels (one launch per stage over every member). */
int pypde_hip_compile_ensemble(const char *src, bool useF, bool useB, bool useS, int ndim, int N,
                               int V, bool STIFF, int FLUX, bool secondOrder,
                               const pypde_options *opts);
/* Tables exactly as the kernels see them (for checks). */
int pypde_hip_tables(int N, double *nodes, double *wghts, double *endvals, double *dervals,
                     double *weno_inv, double *sig, double *dgmat, double *dginv, double *fl_c,
                     double *fl_m);
const char *pypde_hip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
