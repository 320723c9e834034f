// ADER-WENO step kernels for MI355X (gfx950).  Compiled at run time by hipRTC
// after ader_common.h, with the problem's D, N, V, tables and system source.
//
// Device data layout (structure of arrays, cell index fastest):
//   u   [V][ncell]                          interior state
//   wh  [N^D * V][nwc]                      reconstruction, plane = node*V + v
//   tr  [2*D*N*N^(D-1)*V][nwc]              face traces of the predictor,
//                                           plane = (((d*2+e)*N+t)*N^(D-1)+j)*V+v
//   vol [V][nwc]                            volume (source / B) integral
//   ff, fb [N*N^(D-1)*V][nf_d] per d        face flux / B path integral
// nwc = prod(nx_d + 2) reconstructed cells (one ghost layer, iterator.cpp:76).
#pragma once

struct Geom {
  int nx[2];     // interior cells per axis (nx[1] = 1 when D = 1)
  int nw[2];     // reconstructed cells per axis (nx + 2; nw[1] = 1 when D = 1)
  int nb[2];     // ghost-extended cells per axis (nx + 2N)
  int bc[2];     // 0 transmissive, 1 periodic, 2 reflective
  long ncell;    // prod nx
  long nwc;      // prod nw
  long nf[2];    // faces per direction: (nx_d + 1) * prod_{e != d} nx_e
  double dX[2];
  double refl[2][17];  // reflective-wall multipliers per axis (V <= 17)
  int ndt;
  // Flux-item index arithmetic without run-time division: n / nf[d] and n /
  // (faces per row along d: nx[1] for d = 0, nx[1] + 1 for d = 1) as
  // (n * mul) >> sh, exact for n < 2^31 (udiv); idx32 is 1 when every flux
  // item index and every w-cell index is below 2^31.
  unsigned fmul[2], fsh[2], rmul[2], rsh[2];
  int idx32;
};

// n / d for n < 2^31 from d's multiplier m = ceil(2^sh / d), sh = 31 +
// ceil(log2 d): with n = q d + r, n m / 2^sh = q + (r + n e / 2^sh) / d, e =
// m d - 2^sh in [0, d), and n e < 2^31 d <= 2^sh keeps the bracket below d.
DEV unsigned udiv(unsigned n, unsigned m, unsigned sh) {
  return (unsigned)(((unsigned long long)n * m) >> sh);
}

struct StepState {
  double t, dt, tf, cfl;
  long long count, max_steps;
  int push, done, nan_step, pad;
  unsigned long long maxbits;   // unused (the wave-speed maximum lives in pstat[.][8])
  unsigned int nan_flag, pad2;
  unsigned long long pred_iters, pred_cells, pred_capped, pred_fallbacks;
  unsigned int pred_maxit, pad3;
  unsigned long long qhead;     // predictor work queue: next unclaimed w-cell
  // stiff predictor: Newton corrections per spatial node with the full source
  // Jacobian solve and with the A = 0 shortcut, summed over nodes and steps
  unsigned long long pred_nfull, pred_nshort;
  unsigned long long pred_iters32;  // mixed precision: FP32-phase iterations over cells
  unsigned long long hot_count;     // split stiff predictor: cells handed to the hot pass
  // workgroups of k_dt and of the corrector that have finished this step: the
  // last one of each folds the step's partial results (dt; the step's
  // bookkeeping) and resets its ticket
  unsigned int dt_ticket, step_ticket;
  // predictor statistics spread over 64 slots, one 128-byte line each, so the
  // wavefronts' atomics do not queue on one address; the corrector's last
  // workgroup folds them into the totals above and clears them ([0]
  // iterations, [1] cells, [2] capped, [3] fallbacks, [4] max iterations, [5]
  // full and [6] shortcut node corrections, [7] FP32-phase iterations; [8]
  // k_dt's partial maximum, folded by k_dt's last workgroup)
  unsigned long long pstat[64][16];
};

// A workgroup's coordinates within one grid's launch, and the number of
// workgroups of that launch: the hardware's own for a single grid, or one
// member's share of a batched ensemble launch.
struct Blk {
  unsigned int x, y, z, n;
};
DEV Blk hw_blk() {
  return {blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x * gridDim.y * gridDim.z};
}

// Position of workgroup i of n in an order that gives each of the 8 XCDs
// runs of XCD_RUN consecutive positions: workgroups are dispatched to the
// XCDs round-robin (i to XCD i mod 8), so neighbouring positions, whose
// cells share cache lines, meet in one XCD's L2, while costly regions of a
// grid (a reaction zone) still spread over every XCD.  Runs of
// 8 x XCD_RUN positions go round the XCDs; the n mod (8 x XCD_RUN) last
// positions keep their order.  A bijection of [0, n).
constexpr unsigned int XCD_RUN = 16;
DEV unsigned int xcd_order(unsigned int i, unsigned int n) {
  constexpr unsigned int SPAN = 8u * XCD_RUN;
  if (i >= n - n % SPAN) return i;
  const unsigned int x = i & 7u, k = i >> 3;
  return (k / XCD_RUN) * SPAN + x * XCD_RUN + k % XCD_RUN;
}

DEV double wave_max(double x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x = dmax(x, __shfl_xor(x, o, 64));
  return x;
}
// True in every thread of the workgroup that is the last of the grid to
// reach this point; that workgroup also resets the ticket.  What the other
// workgroups hand over must be written with device-scope atomics (performed
// at the device's coherence point) and read back with atomics: each thread
// waits for its own atomics to be acknowledged before the workgroup draws its
// ticket, so the last ticket is drawn after every other workgroup's atomics
// have completed.  No device-scope fence is used: on gfx950 every such fence
// writes back or invalidates the XCD's L2, which costs more than a separate
// one-workgroup launch.  n is the number of workgroups drawing from the
// ticket; every thread of each of them must call it.
DEV bool last_workgroup(unsigned int* ticket, unsigned int n) {
  __shared__ unsigned int is_last;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int prev = atomicAdd(ticket, 1u);
    is_last = (prev == n - 1) ? 1u : 0u;
    if (is_last) atomicExch(ticket, 0u);
  }
  __syncthreads();
  return is_last != 0u;
}
DEV void dt_finalize(StepState* st);

DEV unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
DEV unsigned int wave_max_u32(unsigned int x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    unsigned int y = __shfl_xor(x, o, 64);
    x = x > y ? x : y;
  }
  return x;
}

// --------------------------------------------------------------- WENO
// Boundary ghosts fused into the load, both sweeps staged in LDS
// (boundaries.cpp:25-53, weno.cpp:61-121).  ghost = false gives the
// reference's weno_solver semantics (no ghost extension).
// 2D: one variable per workgroup (coordinate z), a TX x TY tile of output cells
// per workgroup (256 threads); sweep 1 along axis 0 over the TX + 2N - 2 rows
// and TY + 2N - 2 columns the tile needs, then sweep 2 along axis 1 for every
// coefficient index k0.  The autotuner picks the tile per problem.
template <int TX, int TY>
DEV void weno_2d(const double* __restrict__ u, double* __restrict__ wh, const Geom& g, int ghost,
                 int nin0, int nin1, long out_stride, const Blk& b) {
  constexpr int H = 2 * N - 2;
  const int tid = threadIdx.x;
  constexpr int RX = TX + H, RY = TY + H, RYP = RY + 1;
  __shared__ double s0[RX][RYP];
  __shared__ double s1[N][TX][RYP];
  const int v = b.z;
  const int i0 = b.y * TX, j0 = b.x * TY;
  const int nb0 = ghost ? g.nb[0] : nin0, nb1 = ghost ? g.nb[1] : nin1;
  const int no0 = nb0 - H, no1 = nb1 - H;
  const long in_stride = ghost ? g.ncell : (long)nin0 * nin1;
  const int in_row = ghost ? g.nx[1] : nin1;
  for (int e = tid; e < RX * RY; e += 256) {
    int r = e / RY, c = e % RY;
    int I = i0 + r, J = j0 + c;
    double val = 0.0;
    if (I < nb0 && J < nb1) {
      bool f0 = false, f1 = false;
      int s0i = ghost ? ghost_src(I, g.bc[0], g.nx[0], &f0) : I;
      int s1i = ghost ? ghost_src(J, g.bc[1], g.nx[1], &f1) : J;
      double sg = 1.0;
      if (f0) sg = sg * g.refl[0][v];
      if (f1) sg = sg * g.refl[1][v];
      val = u[(long)v * in_stride + (long)s0i * in_row + s1i] * sg;
    }
    s0[r][c] = val;
  }
  __syncthreads();
  for (int e = tid; e < TX * RY; e += 256) {
    int r = e / RY, c = e % RY;
    if (i0 + r < no0) {
      double out[N];
      weno_line([&](int m) { return s0[r + m][c]; }, out);
#pragma unroll
      for (int k = 0; k < N; k++) s1[k][r][c] = out[k];
    }
  }
  __syncthreads();
  for (int e = tid; e < N * TX * TY; e += 256) {
    int c = e % TY, r = (e / TY) % TX, k0 = e / (TY * TX);
    int i = i0 + r, j = j0 + c;
    if (i < no0 && j < no1) {
      double out[N];
      weno_line([&](int m) { return s1[k0][r][c + m]; }, out);
      long cell = (long)i * no1 + j;
#pragma unroll
      for (int k1 = 0; k1 < N; k1++) wh[(long)((k0 * N + k1) * V + v) * out_stride + cell] = out[k1];
    }
  }
}

// 1D: 256 output cells per workgroup, all variables.
DEV void weno_1d(const double* __restrict__ u, double* __restrict__ wh, const Geom& g, int ghost,
                 int nin0, long out_stride, const Blk& b) {
  constexpr int H = 2 * N - 2;
  const int tid = threadIdx.x;
  constexpr int TX = 256;
  constexpr int RX = TX + H;
  __shared__ double s0[V][RX];
  const int i0 = b.x * TX;
  const int nb0 = ghost ? g.nb[0] : nin0;
  const int no0 = nb0 - H;
  for (int e = tid; e < V * RX; e += 256) {
    int v = e / RX, r = e % RX;
    int I = i0 + r;
    double val = 0.0;
    if (I < nb0) {
      bool f0 = false;
      int s = ghost ? ghost_src(I, g.bc[0], g.nx[0], &f0) : I;
      double sg = 1.0;
      if (f0) sg = sg * g.refl[0][v];
      val = u[(long)v * (ghost ? g.ncell : nin0) + s] * sg;
    }
    s0[v][r] = val;
  }
  __syncthreads();
  for (int e = tid; e < V * TX; e += 256) {
    int v = e / TX, r = e % TX;
    int i = i0 + r;
    if (i < no0) {
      double out[N];
      weno_line([&](int m) { return s0[v][r + m]; }, out);
#pragma unroll
      for (int k = 0; k < N; k++) wh[(long)(k * V + v) * out_stride + i] = out[k];
    }
  }
}

#define PYPDE_WENO_KERNEL(NAME, TX, TY)                                                        \
  extern "C" __global__ void __launch_bounds__(256)                                            \
  NAME(const double* __restrict__ u, double* __restrict__ wh, const StepState* __restrict__ st, \
       Geom g, int ghost, int nin0, int nin1, long out_stride) {                               \
    if (st && st->done) return;                                                                \
    if constexpr (D == 1)                                                                      \
      weno_1d(u, wh, g, ghost, nin0, out_stride, hw_blk());                                    \
    else                                                                                       \
      weno_2d<TX, TY>(u, wh, g, ghost, nin0, nin1, out_stride, hw_blk());                      \
  }
PYPDE_WENO_KERNEL(k_weno, 8, 32)
PYPDE_WENO_KERNEL(k_weno_4x64, 4, 64)
PYPDE_WENO_KERNEL(k_weno_16x16, 16, 16)
PYPDE_WENO_KERNEL(k_weno_16x32, 16, 32)

// ------------------------------------------------------------ timestep
// Largest sum_d lambda_d / dX_d over every reconstructed cell
// (stepper.cpp:42-65), reduced as IEEE bit patterns (all values >= 0).
template <int d>
DEV void cell_deriv(const double (&w)[ND][V], int idx, double rdx, double* out) {
  constexpr int stride = ipow_c(N, D - 1 - d);
  // idx is a compile-time constant at every call site after unrolling
  const int kd = (idx / stride) % N;
  const int base = idx - kd * stride;
#pragma unroll
  for (int v = 0; v < V; v++) {
    double acc = T_DER[kd][0] * w[base][v];
#pragma unroll
    for (int j = 1; j < N; j++) acc = dfma(T_DER[kd][j], w[base + j * stride][v], acc);
    out[v] = acc * rdx;
  }
}

// In 2D (V <= 5, no viscous term) two neighbouring lanes share a w-cell and
// take one direction each: twice the wavefronts (a 1024 x 256 grid's 4140
// fill 1.35 rounds of the 3072 slots at three per SIMD) with half the
// eigenvalue work per lane.  Both lanes form the cell's (l_0 / dX_0) + (l_1 /
// dX_1) from the pair's two terms.  The host launches DT_LANES * nwc threads.
constexpr bool DT_PAIR = (D == 2) && (V <= 5) && !PDE_SECOND;
constexpr int DT_LANES = DT_PAIR ? 2 : 1;

DEV void dt_body(const double* __restrict__ wh, StepState* st, const Geom& g, const Blk& b) {
  const long item = (long)b.x * 256 + threadIdx.x;
  const long c = DT_PAIR ? (item >> 1) : item;
  double sum = 0.0;
  if constexpr (DT_PAIR) {
    if (c < g.nwc) {
      const int dl = (int)(item & 1);
      double w[ND][V];
#pragma unroll
      for (int idx = 0; idx < ND; idx++)
#pragma unroll
        for (int v = 0; v < V; v++) w[idx][v] = wh[(long)(idx * V + v) * g.nwc + c];
      double q[V], dq[D * V];
#pragma unroll
      for (int v = 0; v < V; v++) {
        double acc = T_WP[0] * w[0][v];
#pragma unroll
        for (int idx = 1; idx < ND; idx++) acc = dfma(T_WP[idx], w[idx][v], acc);
        q[v] = acc;
      }
      // a first-order F takes no gradient (the reference gives it none)
#pragma unroll
      for (int i = 0; i < D * V; i++) dq[i] = 0.0;
      double f0[V];
      const double part = max_abs_eigs(q, dq, dl, f0) / (dl ? g.dX[1] : g.dX[0]);
      const double other = __shfl_xor(part, 1, 64);
      sum = dl ? other + part : part + other;
      if (!(sum > 0.0)) sum = 0.0;
    }
  } else if (c < g.nwc) {
    double w[ND][V];
#pragma unroll
    for (int idx = 0; idx < ND; idx++)
#pragma unroll
      for (int v = 0; v < V; v++) w[idx][v] = wh[(long)(idx * V + v) * g.nwc + c];
    double q[V], dq[D * V];
#pragma unroll
    for (int v = 0; v < V; v++) {
      double acc = T_WP[0] * w[0][v];
#pragma unroll
      for (int idx = 1; idx < ND; idx++) acc = dfma(T_WP[idx], w[idx][v], acc);
      q[v] = acc;
    }
    static_for<0, D>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      double acc[V], tmp[V];
#pragma unroll
      for (int idx = 0; idx < ND; idx++) {
        cell_deriv<d>(w, idx, 1.0 / g.dX[d], tmp);
#pragma unroll
        for (int v = 0; v < V; v++) acc[v] = (idx == 0) ? T_WP[0] * tmp[v] : dfma(T_WP[idx], tmp[v], acc[v]);
      }
#pragma unroll
      for (int v = 0; v < V; v++) dq[d * V + v] = acc[v];
    });
    static_for<0, D>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      double f0[V];
      double lam = max_abs_eigs(q, dq, d, f0);
#if PDE_SECOND
      lam = lam + max_abs_eigs_second_order<d>(q, dq, g.dX[d]);
#endif
      sum = sum + lam / g.dX[d];
    });
    if (!(sum > 0.0)) sum = 0.0;
  }
  double m = wave_max(sum);
  if ((threadIdx.x & 63) == 0 && m > 0.0)
    atomicMax(&st->pstat[b.x & 63][8], (unsigned long long)__double_as_longlong(m));
  // the last workgroup to finish takes dt from the 64 slots
  if (last_workgroup(&st->dt_ticket, b.n)) dt_finalize(st);
}

extern "C" __global__ void __launch_bounds__(256)
k_dt(const double* __restrict__ wh, StepState* st, Geom g) {
  if (st->done) return;
  dt_body(wh, st, g, hw_blk());
}

// dt = CFL / MAX, x0.2 while count <= 5, clamped to tf - t (stepper.cpp:67-75),
// MAX the maximum over the 64 slots k_dt spreads its atomics over (IEEE bits
// of values >= 0 order as integers); each slot is read and cleared with one
// atomic exchange.  Run by the first wavefront of the last workgroup.
DEV void dt_finalize(StepState* st) {
  if (threadIdx.x >= 64) return;
  unsigned long long b = atomicExch(&st->pstat[threadIdx.x][8], 0ull);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    unsigned long long y = __shfl_xor(b, o, 64);
    b = b > y ? b : y;
  }
  if (threadIdx.x != 0) return;
  double MAX = __longlong_as_double((long long)b);
  double dt = st->cfl / MAX;
  if (st->count <= 5) dt = dt * 0.2;
  if (st->t + dt > st->tf) dt = st->tf - st->t;
  st->dt = dt;
  st->qhead = 0ull;
  st->hot_count = 0ull;
}

// ------------------------------------------------------- eigen diagnostics
// The wave-speed and |A| routines on a batch of V x V matrices (row-major,
// one per thread), for checking them against the CPU restatement:
// which 0: max_abs_eig, out of n entries; 2: abs_apply(A_i, x_i) (Roe and
// Osher modules), out of n V.
extern "C" __global__ void __launch_bounds__(256)
k_eig_batch(const double* __restrict__ A, const double* __restrict__ x, double* __restrict__ out,
            long n, int which) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double H[V * V];
#pragma unroll
  for (int k = 0; k < V * V; k++) H[k] = A[i * V * V + k];
  if (which == 0) {
    out[i] = max_abs_eig(H);
    return;
  }
#if PDE_FLUX != 0
  double xv[V], y[V];
#pragma unroll
  for (int v = 0; v < V; v++) xv[v] = x[i * V + v];
  abs_apply(H, xv, y);
#pragma unroll
  for (int v = 0; v < V; v++) out[i * V + v] = y[v];
#else
  (void)x;
#endif
}
