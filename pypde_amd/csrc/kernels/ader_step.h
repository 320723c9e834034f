// Predictor, face flux, corrector and step bookkeeping kernels (see
// ader_kernels.h for layouts).
#pragma once

// ----------------------------------------------------------- predictor
// One 64-lane wavefront holds CPW cells; lane = slot * ND + spatial node.
// Each lane keeps its node's values at all N time nodes; spatial derivatives
// exchange through LDS.  Cells converge independently (their lanes go idle).
constexpr int CPW = 64 / ND;
static_assert(CPW >= 1, "N^D must not exceed 64");

// Every predictor workgroup is one wavefront (64 lanes: __launch_bounds__(64)
// on each predictor kernel, 64-thread blocks at every launch).  The LDS
// operations of one wavefront are performed in issue order, so exchanging
// values between its lanes through LDS needs only that the compiler keep the
// accesses in program order across this point; a workgroup barrier would also
// wait for every outstanding global load and store of the wavefront.
DEV void pred_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#ifndef PDE_GUESS
#define PDE_GUESS 0
#endif

template <bool B>
struct BConst {
  static constexpr bool value = B;
};

DEV double tw(int i) {
  double r = 0.0;
#pragma unroll
  for (int k = 0; k < N; k++)
    if (k == i) r = T_W[k];
  return r;
}
DEV double tder(int i, int j) {
  double r = 0.0;
#pragma unroll
  for (int k = 0; k < N; k++)
    if (k == i) r = T_DER[k][j];
  return r;
}

// Per-node constants of the space-time system: the DERVALS matrix, and per
// spatial node W[k_d], prod_d W[k_d] (WP), 1 / WP and 1 / (W_t WP).  The
// stiff predictor keeps them in LDS (one table per wavefront) rather than in
// registers, where they would hold some 40 per lane and cost it a wavefront
// per SIMD; the Picard predictor, far from its register limit, keeps them in
// registers (reading them from LDS in its short iterations costs 8 % there),
// and so does the learned-guess variant, which stays at one wavefront per
// SIMD either way (the LDS reads cost it 6 %).
#ifndef PDE_NODE_TAB_LDS
#define PDE_NODE_TAB_LDS (PDE_STIFF && !PDE_GUESS)
#endif
constexpr int NT_W = 0, NT_WP = D, NT_IWP = D + 1, NT_ISC = D + 2, NT_COUNT = D + 2 + N;
struct NodeTab {
  double der[N][N];
  double node[ND][NT_COUNT];
};

struct NodeCtx {
  int slot, idx;
  int kd[D];
  int base[D];
#if PDE_NODE_TAB_LDS
  const NodeTab* nt;
  DEV double der(int d, int j) const { return nt->der[kd[d]][j]; }  // DERVALS row k_d
  DEV double wk(int d) const { return nt->node[idx][NT_W + d]; }    // W[k_d]
  DEV double WP() const { return nt->node[idx][NT_WP]; }
  DEV double IWP() const { return nt->node[idx][NT_IWP]; }
  DEV double isc(int t) const { return nt->node[idx][NT_ISC + t]; }
#else
  double der_[D][N], wk_[D], WP_, IWP_, isc_[N];
  DEV double der(int d, int j) const { return der_[d][j]; }
  DEV double wk(int d) const { return wk_[d]; }
  DEV double WP() const { return WP_; }
  DEV double IWP() const { return IWP_; }
  DEV double isc(int t) const { return isc_[t]; }
#endif
  // W_t prod W (dg.cpp:105-116) and psi_t(0) prod W (dg.cpp:156-171), the
  // products taken in the order d = 0, 1
  DEV double ct(int t) const {
    double a = T_W[t];
#pragma unroll
    for (int d = 0; d < D; d++) a = a * wk(d);
    return a;
  }
  DEV double c0(int t) const {
    double b = T_END0[t];
#pragma unroll
    for (int d = 0; d < D; d++) b = b * wk(d);
    return b;
  }
};

// Lane indices and the per-node constants; with the LDS table, the first
// ND lanes write it and a workgroup barrier follows (every thread of the
// workgroup must call it).
DEV void node_ctx(NodeCtx& c, int lane, NodeTab* nt) {
  c.slot = lane / ND;
  c.idx = lane % ND;
  double wk[D];
#pragma unroll
  for (int d = 0; d < D; d++) {
    int stride = ipow_c(N, D - 1 - d);
    c.kd[d] = (c.idx / stride) % N;
    c.base[d] = c.idx - c.kd[d] * stride;
    wk[d] = tw(c.kd[d]);
  }
  double p = wk[0];
#pragma unroll
  for (int d = 1; d < D; d++) p = p * wk[d];
#if PDE_NODE_TAB_LDS
  c.nt = nt;
  if (lane < N * N) nt->der[lane / N][lane % N] = tder(lane / N, lane % N);
  if (lane < ND) {
    double* row = nt->node[lane];
#pragma unroll
    for (int d = 0; d < D; d++) row[NT_W + d] = wk[d];
    row[NT_WP] = p;
    row[NT_IWP] = 1.0 / p;
#pragma unroll
    for (int t = 0; t < N; t++) row[NT_ISC + t] = 1.0 / (T_W[t] * p);
  }
  pred_sync();
#else
  (void)nt;
#pragma unroll
  for (int d = 0; d < D; d++) {
    c.wk_[d] = wk[d];
#pragma unroll
    for (int j = 0; j < N; j++) c.der_[d][j] = tder(c.kd[d], j);
  }
  c.WP_ = p;
  c.IWP_ = 1.0 / p;
#pragma unroll
  for (int t = 0; t < N; t++) c.isc_[t] = 1.0 / (T_W[t] * p);
#endif
}

// Whether a predictor term reads the iterate's spatial gradient: a viscous F
// (second order) or B.  A first-order F gets dQ = 0, as the face fluxes pass
// zero face gradients: in the reference a first-order F has no gradient
// argument (cfuncs.py:11-27 choose its arity), so none can read it; the
// iterate then need not be staged in LDS every iteration.
constexpr bool PRED_DQ = PDE_SECOND || PDE_USE_B;

// derivative along d at this lane's node of buffer buf[slot][t][node][v]
template <int d, int NTB>
DEV void lds_deriv(const NodeCtx& c, const double (*buf)[NTB][ND][V], int t, double rdx, double* out) {
  constexpr int stride = ipow_c(N, D - 1 - d);
#pragma unroll
  for (int v = 0; v < V; v++) {
    double acc = c.der(d, 0) * buf[c.slot][t][c.base[d]][v];
#pragma unroll
    for (int j = 1; j < N; j++) acc = dfma(c.der(d, j), buf[c.slot][t][c.base[d] + j * stride][v], acc);
    out[v] = acc * rdx;
  }
}

DEV void callS(double* out, const double* q) {
#if PDE_USE_S
  S(out, q);
#else
#pragma unroll
  for (int v = 0; v < V; v++) out[v] = 0.0;
#endif
}

// F is staged in LDS one time node at a time (NTF = 1), except for a viscous
// F, whose gradient traces need every time node's derivatives (NTF = N).
constexpr int NTF = PDE_SECOND ? N : 1;
struct PredLDS {
  double q[CPW][N][ND][V];
#if PDE_USE_F
  double f[D][CPW][NTF][ND][V];
#endif
#if PDE_GUESS
  double red[64];  // per-lane residuals of the learned-guess check
#endif
#if PDE_NODE_TAB_LDS
  NodeTab nt;
#endif
};
#if PDE_NODE_TAB_LDS
#define PRED_NODE_TAB(sm) (&(sm).nt)
#else
#define PRED_NODE_TAB(sm) ((NodeTab*)nullptr)
#endif

#if PDE_SECOND
// Gradient traces for the viscous flux (fv.cpp:87-122): the derivatives of
// the converged q(t) along every d_ at this lane's node (staged in sm.f), then
// extrapolated along each face direction d to both cell ends; planes
// [d][side][t][transverse node][d_][v] of trd.  Call with every lane of the
// wavefront (it synchronises); `on` marks lanes whose cell is done, and
// sm.q must hold the converged q.
DEV void grad_traces(const NodeCtx& c, PredLDS& sm, const Geom& g, long cell, bool on,
                     double* __restrict__ trd) {
  if (on) {
    static_for<0, D>([&](auto dc_) {
      constexpr int d_ = decltype(dc_)::value;
#pragma unroll
      for (int t = 0; t < N; t++) {
        double tmp[V];
        lds_deriv<d_>(c, sm.q, t, 1.0 / g.dX[d_], tmp);
#pragma unroll
        for (int v = 0; v < V; v++) sm.f[d_][c.slot][t][c.idx][v] = tmp[v];
      }
    });
  }
  pred_sync();
  if (on) {
    static_for<0, D>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      constexpr int stride = ipow_c(N, D - 1 - d);
      if (c.kd[d] == 0) {
        const int j = (c.idx / (stride * N)) * stride + c.idx % stride;
#pragma unroll
        for (int e = 0; e < 2; e++)
#pragma unroll
          for (int t = 0; t < N; t++)
#pragma unroll
            for (int d_ = 0; d_ < D; d_++)
#pragma unroll
              for (int v = 0; v < V; v++) {
                double acc = (e ? T_END1[0] : T_END0[0]) * sm.f[d_][c.slot][t][c.idx][v];
#pragma unroll
                for (int m = 1; m < N; m++)
                  acc = dfma(e ? T_END1[m] : T_END0[m], sm.f[d_][c.slot][t][c.idx + m * stride][v], acc);
                const int plane = ((((d * 2 + e) * N + t) * ND1 + j) * D + d_) * V + v;
                trd[(long)plane * g.nwc + cell] = acc;
              }
      }
    });
  }
  pred_sync();
}
#endif

// rhs of the space-time system at this lane's node for all t (dg.cpp:50-126):
// r = W_t prod W (S(q) - sum_d (B_d dq_d + d_d F_d)),  rhs = r dt + Ww.
// Assumes sm.q holds the current iterate when PRED_DQ; leaves the lane's dq
// in dqs (zero when no term reads it).  With
// NTF = 1, F of one time node at a time passes through LDS (two barriers per
// time node); every (t, v) result is the same either way.
DEV void node_rhs(const NodeCtx& c, PredLDS& sm, const double (&q)[N][V], const double* w,
                  double dt, const Geom& g, bool active, double (&rhs)[N][V],
                  double (&dqs)[N][D][V]) {
  auto eval_f = [&](int t, int slot_t) {
    static_for<0, D>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      if constexpr (PRED_DQ) {
        lds_deriv<d>(c, sm.q, t, 1.0 / g.dX[d], dqs[t][d]);
      } else {
#pragma unroll
        for (int v = 0; v < V; v++) dqs[t][d][v] = 0.0;
      }
    });
#if PDE_USE_F
#pragma unroll
    for (int d = 0; d < D; d++) {
      double f[V];
      F(f, q[t], &dqs[t][0][0], d);
#pragma unroll
      for (int v = 0; v < V; v++) sm.f[d][c.slot][slot_t][c.idx][v] = f[v];
    }
#else
    (void)slot_t;
#endif
  };
  auto finish = [&](int t, int slot_t) {
    double r[V];
    callS(r, q[t]);
    static_for<0, D>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
#if PDE_USE_B
      double b[V * V], bd[V];
      B(b, q[t], d);
      b_times<PDE_B_QUIRK ? 1 : 0>(b, dqs[t][d], bd);
#pragma unroll
      for (int v = 0; v < V; v++) r[v] = r[v] - bd[v];
#endif
#if PDE_USE_F
      double df[V];
      lds_deriv<d>(c, sm.f[d], slot_t, 1.0 / g.dX[d], df);
#pragma unroll
      for (int v = 0; v < V; v++) r[v] = r[v] - df[v];
#else
      (void)slot_t;
#endif
    });
#pragma unroll
    for (int v = 0; v < V; v++) {
      double rr = r[v] * c.ct(t);
      rhs[t][v] = rr * dt + c.c0(t) * w[v];
    }
  };
  if constexpr (NTF == N) {
    if (active) {
#pragma unroll
      for (int t = 0; t < N; t++) eval_f(t, t);
    }
#if PDE_USE_F
    pred_sync();
#endif
    if (active) {
#pragma unroll
      for (int t = 0; t < N; t++) finish(t, t);
    }
  } else {
#pragma unroll
    for (int t = 0; t < N; t++) {
      if (active) eval_f(t, 0);
#if PDE_USE_F
      pred_sync();
#endif
      if (active) finish(t, 0);
#if PDE_USE_F
      if (t + 1 < N) pred_sync();
#endif
    }
  }
}

// node_rhs for an iterate that is constant in time (q(t) = w at every time
// node: the first iteration from the plain guess).  F, B, S and the spatial
// derivatives are then the same at every t, so they are evaluated once, at
// t = 0, and reused; every value equals the one node_rhs computes.
DEV void node_rhs_const(const NodeCtx& c, PredLDS& sm, const double (&q)[N][V], const double* w,
                        double dt, const Geom& g, bool active, double (&rhs)[N][V],
                        double (&dqs)[N][D][V]) {
  if (active) {
    static_for<0, D>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      if constexpr (PRED_DQ) {
        lds_deriv<d>(c, sm.q, 0, 1.0 / g.dX[d], dqs[0][d]);
      } else {
#pragma unroll
        for (int v = 0; v < V; v++) dqs[0][d][v] = 0.0;
      }
    });
#if PDE_USE_F
#pragma unroll
    for (int d = 0; d < D; d++) {
      double f[V];
      F(f, q[0], &dqs[0][0][0], d);
#pragma unroll
      for (int v = 0; v < V; v++) sm.f[d][c.slot][0][c.idx][v] = f[v];
    }
#endif
  }
#if PDE_USE_F
  pred_sync();
#endif
  if (active) {
    double r[V];
    callS(r, q[0]);
    static_for<0, D>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
#if PDE_USE_B
      double b[V * V], bd[V];
      B(b, q[0], d);
      b_times<PDE_B_QUIRK ? 1 : 0>(b, dqs[0][d], bd);
#pragma unroll
      for (int v = 0; v < V; v++) r[v] = r[v] - bd[v];
#endif
#if PDE_USE_F
      double df[V];
      lds_deriv<d>(c, sm.f[d], 0, 1.0 / g.dX[d], df);
#pragma unroll
      for (int v = 0; v < V; v++) r[v] = r[v] - df[v];
#endif
    });
#pragma unroll
    for (int t = 0; t < N; t++)
#pragma unroll
      for (int v = 0; v < V; v++) {
        double rr = r[v] * c.ct(t);
        rhs[t][v] = rr * dt + c.c0(t) * w[v];
      }
#pragma unroll
    for (int t = 1; t < N; t++)
#pragma unroll
      for (int d = 0; d < D; d++)
#pragma unroll
        for (int v = 0; v < V; v++) dqs[t][d][v] = dqs[0][d][v];
  }
}

#if PDE_STIFF
constexpr double NEWTON_A_TINY = 1e-6;  // max |dt dS/dQ| below which the step drops A
// Newton correction at one spatial node (see the CPU restatement for the
// derivation): g_t = G_t / (W_t WP), A = dt dS/dQ at the time-weighted mean,
// h_t = sum_j A^j sum_s M_{N-j}[t][s] g_s, P = p(A), dq = -P^-1 h.
// Returns true where the source Jacobian was solved, false for the shortcut.
// FULL = false builds only the shortcut: it returns true, leaving G as it
// was, where the full solve would be needed (the caller hands the cell on).
template <bool FULL = true>
DEV bool newton_node(const NodeCtx& c, const double (&q)[N][V], double (&G)[N][V], double dt) {
  double g[N][V];
#pragma unroll
  for (int t = 0; t < N; t++)
#pragma unroll
    for (int v = 0; v < V; v++) g[t][v] = G[t][v] * c.isc(t);
  double qbar[V];
#pragma unroll
  for (int v = 0; v < V; v++) {
    double acc = T_W[0] * q[0][v];
#pragma unroll
    for (int t = 1; t < N; t++) acc = dfma(T_W[t], q[t][v], acc);
    qbar[v] = acc;
  }
  double A[V * V];
#if PDE_USE_M
  M(A, qbar);
#elif PDE_USE_S
  {
    double s0[V], s1[V], qq[V];
#pragma unroll
    for (int v = 0; v < V; v++) qq[v] = qbar[v];
    S(s0, qq);
#pragma unroll
    for (int j = 0; j < V; j++) {
      double q0 = qq[j];
      double hh = dmax(FD_EPS * dabs(q0), FD_EPS);
      qq[j] = q0 + hh;
      S(s1, qq);
      qq[j] = q0;
      double rhh = 1.0 / hh;
#pragma unroll
      for (int i = 0; i < V; i++) A[i * V + j] = (s1[i] - s0[i]) * rhh;
    }
  }
#else
#pragma unroll
  for (int i = 0; i < V * V; i++) A[i] = 0.0;
#endif
  // max |dt dS/dQ| as |dt| max |dS/dQ|: rounding is monotone, so this is the
  // maximum of the products, bit for bit (and 0 for a NaN dt, as the maximum
  // that skips NaN entries gives); dt A itself is formed in the full path only
  double amax = 0.0;
#pragma unroll
  for (int i = 0; i < V * V; i++) amax = dmax(amax, dabs(A[i]));
  amax = dmax(0.0, dabs(dt) * amax);
  if (amax < NEWTON_A_TINY) {
    // negligible source Jacobian: dq_t = -(sum_s M_N[t][s] g_s) (1 / c_0)
#pragma unroll
    for (int t = 0; t < N; t++)
#pragma unroll
      for (int v = 0; v < V; v++) {
        double acc = T_FLM[N - 1][t][0] * g[0][v];
#pragma unroll
        for (int s2 = 1; s2 < N; s2++) acc = dfma(T_FLM[N - 1][t][s2], g[s2][v], acc);
        G[t][v] = -(acc * (1.0 / T_FLC[0]));
      }
    return false;
  }
  if constexpr (!FULL) return true;
#pragma unroll
  for (int i = 0; i < V * V; i++) A[i] = dt * A[i];
  // h_t (Horner in A); result left in G
#pragma unroll
  for (int t = 0; t < N; t++) {
    double h[V];
#pragma unroll
    for (int j = N - 1; j >= 0; j--) {
      double u[V];
#pragma unroll
      for (int v = 0; v < V; v++) {
        double acc = T_FLM[N - j - 1][t][0] * g[0][v];
#pragma unroll
        for (int s = 1; s < N; s++) acc = dfma(T_FLM[N - j - 1][t][s], g[s][v], acc);
        u[v] = acc;
      }
      if (j == N - 1) {
#pragma unroll
        for (int v = 0; v < V; v++) h[v] = u[v];
      } else {
        double Ah[V];
#pragma unroll
        for (int i = 0; i < V; i++) {
          double acc = A[i * V] * h[0];
#pragma unroll
          for (int k = 1; k < V; k++) acc = dfma(A[i * V + k], h[k], acc);
          Ah[i] = acc;
        }
#pragma unroll
        for (int v = 0; v < V; v++) h[v] = Ah[v] + u[v];
      }
    }
#pragma unroll
    for (int v = 0; v < V; v++) G[t][v] = h[v];
  }
  // P = p(A) = A^N + c_{N-1} A^{N-1} + ... + c_0 I
  double P[V * V];
#pragma unroll
  for (int i = 0; i < V * V; i++) P[i] = A[i];
#pragma unroll
  for (int i = 0; i < V; i++) P[i * V + i] = P[i * V + i] + T_FLC[N - 1];
#pragma unroll
  for (int j = N - 2; j >= 0; j--) {
    // P <- P A + c_j I, one row at a time (row i of P A needs only row i of P)
#pragma unroll
    for (int i = 0; i < V; i++) {
      double row[V];
#pragma unroll
      for (int k = 0; k < V; k++) {
        double acc = P[i * V] * A[k];
#pragma unroll
        for (int m = 1; m < V; m++) acc = dfma(P[i * V + m], A[m * V + k], acc);
        row[k] = acc;
      }
#pragma unroll
      for (int k = 0; k < V; k++) P[i * V + k] = row[k];
      P[i * V + i] = P[i * V + i] + T_FLC[j];
    }
  }
  // LU with partial pivoting (first largest pivot), static indices only; each
  // pivot inverted once (rd[k] = 1 / u_kk); the row swaps, done with selects,
  // are skipped for the whole wavefront where no lane pivots
  int piv[V];
  double rd[V];
  bool swapped = false;
#pragma unroll
  for (int k = 0; k < V; k++) {
    int p = k;
    double best = dabs(P[k * V + k]);
#pragma unroll
    for (int i = k + 1; i < V; i++) {
      double a = dabs(P[i * V + k]);
      if (a > best) {
        best = a;
        p = i;
      }
    }
    piv[k] = p;
    swapped = swapped || p != k;
    if (__ballot(p != k) != 0ull) {
#pragma unroll
      for (int i = k + 1; i < V; i++) {
        if (p == i) {
#pragma unroll
          for (int j = 0; j < V; j++) {
            double tmp = P[k * V + j];
            P[k * V + j] = P[i * V + j];
            P[i * V + j] = tmp;
          }
        }
      }
    }
    rd[k] = 1.0 / P[k * V + k];
#pragma unroll
    for (int i = k + 1; i < V; i++) {
      double l = P[i * V + k] * rd[k];
      P[i * V + k] = l;
#pragma unroll
      for (int j = k + 1; j < V; j++) P[i * V + j] = dfma(-l, P[k * V + j], P[i * V + j]);
    }
  }
  const bool any_swap = __ballot(swapped) != 0ull;
#pragma unroll
  for (int t = 0; t < N; t++) {
    double* b = G[t];
    if (any_swap) {
#pragma unroll
      for (int k = 0; k < V; k++) {
#pragma unroll
        for (int i = k + 1; i < V; i++) {
          if (piv[k] == i) {
            double tmp = b[k];
            b[k] = b[i];
            b[i] = tmp;
          }
        }
      }
    }
#pragma unroll
    for (int i = 1; i < V; i++) {
      double acc = b[i];
#pragma unroll
      for (int j = 0; j < i; j++) acc = dfma(-P[i * V + j], b[j], acc);
      b[i] = acc;
    }
#pragma unroll
    for (int i = V - 1; i >= 0; i--) {
      double acc = b[i];
#pragma unroll
      for (int j = i + 1; j < V; j++) acc = dfma(-P[i * V + j], b[j], acc);
      b[i] = -(acc * rd[i]);
    }
  }
  return true;
}
#endif

// Persistent, work-queued: each wavefront claims cells from st->qhead (reset
// by k_dt's last workgroup every step), iterates its CPW slots, and as soon as a slot's
// cell has converged writes that cell's outputs and refills the slot with the
// next unclaimed cell, so cells with many Newton iterations do not hold the
// lanes of cells that need few.  Every cell's arithmetic is independent of
// which slot or wavefront runs it.
extern "C" __global__ void __launch_bounds__(64)
k_predictor(const double* __restrict__ wh, const double* __restrict__ guess,
            double* __restrict__ tr, double* __restrict__ vol, double* __restrict__ qh,
            StepState* st, Geom g, double newton_tol, int newton_maxit,
            double* __restrict__ trd) {
  if (st->done) return;
  __shared__ PredLDS sm;
  const double dt = st->dt;
  const int lane = threadIdx.x;
  NodeCtx c;
  node_ctx(c, lane, PRED_NODE_TAB(sm));
  const bool lane_ok = c.slot < CPW;
  const unsigned long long cellmask =
      (ND == 64 ? ~0ull : ((1ull << ND) - 1ull)) << (lane_ok ? c.slot * ND : 0);
#if PDE_STIFF
  const int maxit = newton_maxit;
#else
  const int maxit = DG_IT;
#endif
  double w[V], q[N][V], rhs[N][V], dqs[N][D][V];
  long cell = -1;
  bool active = false;
  int it = 0;
  int learned = 0;
  int fellback = 0;
  double r0 = 0.0;
  // per-lane statistics over the cells this lane's slot has run
  unsigned long long s_iters = 0ull, s_cells = 0ull, s_cap = 0ull, s_fb = 0ull;
  unsigned int s_max = 0u;
  unsigned int n_full = 0u, n_short = 0u;  // node corrections of this lane

  auto start_cell = [&](long cl) {
    cell = cl;
    active = cl >= 0;
    it = 0;
    fellback = 0;
    r0 = 0.0;
    learned = (PDE_GUESS && guess) ? 1 : 0;
    if (active) {
#pragma unroll
      for (int v = 0; v < V; v++) w[v] = wh[(long)(c.idx * V + v) * g.nwc + cl];
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int v = 0; v < V; v++)
          q[t][v] = (PDE_GUESS && guess) ? guess[(long)((t * ND + c.idx) * V + v) * g.nwc + cl]
                                         : w[v];
    }
  };
  {
    long long base = 0;
    if (lane == 0) base = (long long)atomicAdd(&st->qhead, (unsigned long long)CPW);
    base = __shfl(base, 0, 64);
    long cl = (long)base + c.slot;
    start_cell((lane_ok && cl < g.nwc) ? cl : -1);
  }

  while (__ballot(active) != 0ull) {
    if (PRED_DQ && active) {
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int v = 0; v < V; v++) sm.q[c.slot][t][c.idx][v] = q[t][v];
    }
    pred_sync();
    node_rhs(c, sm, q, w, dt, g, active, rhs, dqs);
    bool more = false;
    bool skip = false;
#if PDE_STIFF
    // residual G = rhs - (DG_MAT (x) W) q   (dg.cpp:128-154)
    double rmax = 0.0;
    if (active) {
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int v = 0; v < V; v++) {
          double acc = (T_DGMAT[t][0] * c.WP()) * q[0][v];
#pragma unroll
          for (int s2 = 1; s2 < N; s2++) acc = dfma(T_DGMAT[t][s2] * c.WP(), q[s2][v], acc);
          rhs[t][v] = rhs[t][v] - acc;
          double a = dabs(rhs[t][v]);
          rmax = (a > rmax || a != a) ? a : rmax;  // keeps a NaN
        }
    }
#if PDE_GUESS
    // while a cell is still on its learned guess, a residual that is not finite
    // or larger than at the previous iterate sends it back to the plain guess
    if (__ballot(active && learned) != 0ull) {
      pred_sync();
      sm.red[lane] = rmax;
      pred_sync();
      double cm = 0.0;
#pragma unroll
      for (int i = 0; i < ND; i++) {
        double e = sm.red[(lane_ok ? c.slot : 0) * ND + i];
        cm = (e > cm || e != e) ? e : cm;  // max that keeps a NaN
      }
      if (active && learned) {
        if (!(cm <= (it > 0 ? r0 : 1.7976931348623157e308))) {
#pragma unroll
          for (int t = 0; t < N; t++)
#pragma unroll
            for (int v = 0; v < V; v++) q[t][v] = w[v];
          fellback = 1;
          learned = 0;
          skip = true;
        } else {
          r0 = cm;
        }
      }
    }
#endif
    if (active && !skip) {
      if (newton_node(c, q, rhs, dt)) n_full++;
      else n_short++;
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int v = 0; v < V; v++) {
          if (!(dabs(rhs[t][v]) <= newton_tol * (1.0 + dabs(q[t][v])))) more = true;
          q[t][v] = q[t][v] + rhs[t][v];
        }
    }
    if (skip) more = true;
#else
    (void)learned;
    (void)r0;
    (void)newton_tol;
    if (active) {
      // Picard: q1 = DG_MAT^-1 rhs / prod W  (dg.cpp:187-202)
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int v = 0; v < V; v++) {
          double acc = T_DGINV[t][0] * rhs[0][v];
#pragma unroll
          for (int s2 = 1; s2 < N; s2++) acc = dfma(T_DGINV[t][s2], rhs[s2][v], acc);
          double q1 = acc * c.IWP();
          if (dabs(q1 - q[t][v]) > DG_TOL * (1.0 + dabs(q[t][v]))) more = true;
          q[t][v] = q1;
        }
    }
#endif
    const unsigned long long mb = __ballot(more);
    bool fin = false;
    if (active) {
      it++;
      if ((mb & cellmask) == 0ull || it >= maxit) {
        active = false;
        fin = true;
      }
    }
    pred_sync();
    if (__ballot(fin) == 0ull) continue;

    // ---- finished slots: traces, volume term, optional full qh
    if (fin) {
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int v = 0; v < V; v++) sm.q[c.slot][t][c.idx][v] = q[t][v];
    }
    pred_sync();
    if (fin) {
      static_for<0, D>([&](auto dc) {
        constexpr int d = decltype(dc)::value;
        constexpr int stride = ipow_c(N, D - 1 - d);
        if (c.kd[d] == 0) {
          const int j = (c.idx / (stride * N)) * stride + c.idx % stride;
#pragma unroll
          for (int e = 0; e < 2; e++)
#pragma unroll
            for (int t = 0; t < N; t++)
#pragma unroll
              for (int v = 0; v < V; v++) {
                double acc = (e ? T_END1[0] : T_END0[0]) * sm.q[c.slot][t][c.idx][v];
#pragma unroll
                for (int m = 1; m < N; m++)
                  acc = dfma(e ? T_END1[m] : T_END0[m], sm.q[c.slot][t][c.idx + m * stride][v], acc);
                const int plane = (((d * 2 + e) * N + t) * ND1 + j) * V + v;
                tr[(long)plane * g.nwc + cell] = acc;
              }
        }
      });
      if (qh) {
#pragma unroll
        for (int t = 0; t < N; t++)
#pragma unroll
          for (int v = 0; v < V; v++) qh[(long)((t * ND + c.idx) * V + v) * g.nwc + cell] = q[t][v];
      }
    }
#if PDE_SECOND
    grad_traces(c, sm, g, cell, fin, trd);
#endif
#if PDE_USE_S || PDE_USE_B
    // volume term (fv.cpp:33-85): per node sum over t, then over nodes in order
    double part[V];
    if (fin) {
#pragma unroll
      for (int t = 0; t < N; t++) {
        double sv[V];
        callS(sv, q[t]);
#if PDE_USE_B
        static_for<0, D>([&](auto dc) {
          constexpr int d = decltype(dc)::value;
          double dqd[V], b[V * V], bd[V];
          lds_deriv<d>(c, sm.q, t, 1.0 / g.dX[d], dqd);
          B(b, q[t], d);
          b_times<PDE_B_QUIRK ? 2 : 0>(b, dqd, bd);
#pragma unroll
          for (int v = 0; v < V; v++) sv[v] = sv[v] - bd[v];
        });
#endif
        double tmp = dt * T_W[t];
#pragma unroll
        for (int d = 0; d < D; d++) tmp = tmp * c.wk(d);
#pragma unroll
        for (int v = 0; v < V; v++) part[v] = (t == 0) ? tmp * sv[v] : dfma(tmp, sv[v], part[v]);
      }
    }
    pred_sync();
    if (fin) {
#pragma unroll
      for (int v = 0; v < V; v++) sm.q[c.slot][0][c.idx][v] = part[v];
    }
    pred_sync();
    if (fin && c.idx == 0) {
#pragma unroll
      for (int v = 0; v < V; v++) {
        double acc = sm.q[c.slot][0][0][v];
#pragma unroll
        for (int i = 1; i < ND; i++) acc = acc + sm.q[c.slot][0][i][v];
        vol[(long)v * g.nwc + cell] = acc;
      }
    }
    pred_sync();
#endif
    if (fin && c.idx == 0) {
      s_iters += (unsigned long long)it;
      s_cells += 1ull;
      s_cap += (it >= maxit) ? 1ull : 0ull;
      s_fb += (unsigned long long)fellback;
      s_max = (unsigned int)it > s_max ? (unsigned int)it : s_max;
    }
    // ---- refill finished slots with the next unclaimed cells, in slot order
    const unsigned long long lead = __ballot(fin && c.idx == 0);
    long long base = 0;
    if (lane == 0) base = (long long)atomicAdd(&st->qhead, (unsigned long long)__popcll(lead));
    base = __shfl(base, 0, 64);
    if (fin) {
      const unsigned long long below = lead & ((1ull << (c.slot * ND)) - 1ull);
      long cl = (long)base + __popcll(below);
      start_cell(cl < g.nwc ? cl : -1);
    }
  }
  // ---- statistics
  s_iters = wave_sum_u64(s_iters);
  s_cells = wave_sum_u64(s_cells);
  s_cap = wave_sum_u64(s_cap);
  s_fb = wave_sum_u64(s_fb);
  s_max = wave_max_u32(s_max);
  const unsigned long long w_full = wave_sum_u64(n_full), w_short = wave_sum_u64(n_short);
  if (lane == 0 && s_cells) {
    unsigned long long* ps = st->pstat[blockIdx.x & 63];
    atomicAdd(&ps[0], s_iters);
    atomicAdd(&ps[1], s_cells);
    if (s_cap) atomicAdd(&ps[2], s_cap);
    if (s_fb) atomicAdd(&ps[3], s_fb);
    atomicMax(&ps[4], (unsigned long long)s_max);
    if (w_full) atomicAdd(&ps[5], w_full);
    if (w_short) atomicAdd(&ps[6], w_short);
  }
}

// The space-time DG predictor over all reconstructed cells (dg.cpp:204-224):
// initial guess (dg.cpp:12-20, 156-171), Picard (dg.cpp:187-202) or the
// per-node stiff Newton in place of the reference's Newton-Krylov
// (dg.cpp:173-185), then the face traces (evaluations.cpp:49-70,
// fv.cpp:102-122) and the cell-volume term (fv.cpp:33-85) for the corrector.
// Static schedule: one wavefront per group of CPW consecutive cells, which
// stays until its slowest cell has converged.
// MODE 0: every cell, by block.  MODE 1 (stiff, plain guess): the same cells,
// but built without the full source-Jacobian solve: a cell where a node would
// need it is dropped (no outputs, no statistics) and appended to hot[], so
// this kernel needs far fewer registers.  MODE 2: the cells listed in hot[]
// by MODE 1, with the full solve.  Every cell gets the same arithmetic as in
// MODE 0.
template <int MODE = 0>
DEV void predictor_static_body(PredLDS& sm, const double* __restrict__ wh,
                               const double* __restrict__ guess, double* __restrict__ tr,
                               double* __restrict__ vol, double* __restrict__ qh, StepState* st,
                               const Geom& g, double newton_tol, int newton_maxit,
                               double* __restrict__ trd, const Blk& b,
                               unsigned int* __restrict__ hot = nullptr) {
  if (st->done) return;
  if (MODE == 2 && (unsigned long long)b.x * CPW >= st->hot_count) return;
  const double dt = st->dt;
  const int lane = threadIdx.x;
  NodeCtx c;
  node_ctx(c, lane, PRED_NODE_TAB(sm));
  long cell;
  bool valid;
  if constexpr (MODE == 2) {
    const unsigned long long i = (unsigned long long)b.x * CPW + c.slot;
    valid = (c.slot < CPW) && (i < st->hot_count);
    cell = valid ? (long)hot[i] : 0;
  } else {
    cell = (long)xcd_order(b.x, b.n) * CPW + c.slot;
    valid = (c.slot < CPW) && (cell < g.nwc);
  }
  double w[V], q[N][V], rhs[N][V], dqs[N][D][V];
  if (valid) {
#pragma unroll
    for (int v = 0; v < V; v++) w[v] = wh[(long)(c.idx * V + v) * g.nwc + cell];
#pragma unroll
    for (int t = 0; t < N; t++)
#pragma unroll
      for (int v = 0; v < V; v++)
        q[t][v] = (PDE_GUESS && guess) ? guess[(long)((t * ND + c.idx) * V + v) * g.nwc + cell] : w[v];
  }
  const unsigned long long cellmask =
      (ND == 64 ? ~0ull : ((1ull << ND) - 1ull)) << (c.slot < CPW ? c.slot * ND : 0);
#if PDE_STIFF
  const int maxit = newton_maxit;
#else
  const int maxit = DG_IT;
#endif
  bool active = valid;
  int it = 0;
  int learned = (PDE_GUESS && guess) ? 1 : 0;
  int fellback = 0;
  double r0 = 0.0;
  unsigned int n_full = 0u, n_short = 0u;  // node corrections of this lane
  bool cell_hot = false;                   // MODE 1: the cell needs the full solve
  // One iteration for every cell still active; FIRST: the first iteration
  // from the plain guess, whose iterate is constant in time (the wave-uniform
  // q = w), so the right-hand side needs F, S and the derivatives once.  It is
  // peeled off the loop so that the loop holds one right-hand side only
  // (the stiff cold pass then needs 208 registers instead of 234).
  auto iterate = [&](auto firstc) {
    constexpr bool FIRST = decltype(firstc)::value;
    if (PRED_DQ && active) {
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int v = 0; v < V; v++) sm.q[c.slot][t][c.idx][v] = q[t][v];
    }
    pred_sync();
    if constexpr (FIRST)
      node_rhs_const(c, sm, q, w, dt, g, active, rhs, dqs);
    else
      node_rhs(c, sm, q, w, dt, g, active, rhs, dqs);
    bool more = false;
    bool skip = false;
#if PDE_STIFF
    // residual G = rhs - (DG_MAT (x) W) q   (dg.cpp:128-154)
    double rmax = 0.0;
    if (active) {
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int v = 0; v < V; v++) {
          double acc = (T_DGMAT[t][0] * c.WP()) * q[0][v];
#pragma unroll
          for (int s = 1; s < N; s++) acc = dfma(T_DGMAT[t][s] * c.WP(), q[s][v], acc);
          rhs[t][v] = rhs[t][v] - acc;
          double a = dabs(rhs[t][v]);
          rmax = (a > rmax || a != a) ? a : rmax;  // keeps a NaN
        }
    }
#if PDE_GUESS
    // while a cell is still on its learned guess, a residual that is not finite
    // or larger than at the previous iterate sends it back to the plain guess
    if (__ballot(active && learned) != 0ull) {
      pred_sync();
      sm.red[lane] = rmax;
      pred_sync();
      double cm = 0.0;
#pragma unroll
      for (int i = 0; i < ND; i++) {
        double e = sm.red[(c.slot < CPW ? c.slot : 0) * ND + i];
        cm = (e > cm || e != e) ? e : cm;  // max that keeps a NaN
      }
      if (active && learned) {
        if (!(cm <= (it > 0 ? r0 : 1.7976931348623157e308))) {
#pragma unroll
          for (int t = 0; t < N; t++)
#pragma unroll
            for (int v = 0; v < V; v++) q[t][v] = w[v];
          fellback = 1;
          learned = 0;
          skip = true;
        } else {
          r0 = cm;
        }
      }
    }
#endif
    bool node_hot = false;
    if (active && !skip) {
      if (newton_node<MODE != 1>(c, q, rhs, dt)) {
        n_full++;
        node_hot = MODE == 1;
      } else {
        n_short++;
      }
      if (!node_hot) {
#pragma unroll
        for (int t = 0; t < N; t++)
#pragma unroll
          for (int v = 0; v < V; v++) {
            if (!(dabs(rhs[t][v]) <= newton_tol * (1.0 + dabs(q[t][v])))) more = true;
            q[t][v] = q[t][v] + rhs[t][v];
          }
      }
    }
    if (skip) more = true;
    if constexpr (MODE == 1) {
      // a cell with a node that needs the full solve leaves for the hot pass
      if ((__ballot(node_hot) & cellmask) != 0ull && active) {
        cell_hot = true;
        active = false;
      }
    }
#else
    (void)learned;
    (void)r0;
    (void)newton_tol;
    if (active) {
      // Picard: q1 = DG_MAT^-1 rhs / prod W  (dg.cpp:187-202)
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int v = 0; v < V; v++) {
          double acc = T_DGINV[t][0] * rhs[0][v];
#pragma unroll
          for (int s = 1; s < N; s++) acc = dfma(T_DGINV[t][s], rhs[s][v], acc);
          double q1 = acc * c.IWP();
          if (dabs(q1 - q[t][v]) > DG_TOL * (1.0 + dabs(q[t][v]))) more = true;
          q[t][v] = q1;
        }
    }
#endif
    const unsigned long long mb = __ballot(more);
    if (active) {
      it++;
      if ((mb & cellmask) == 0ull || it >= maxit) active = false;
    }
    pred_sync();
  };
  if (!(PDE_GUESS && guess) && __ballot(active) != 0ull) iterate(BConst<true>{});
  while (__ballot(active) != 0ull) iterate(BConst<false>{});

  // ---- epilogue: traces, volume term, optional full qh
  if constexpr (MODE == 1) {
    // hand hot cells to the hot pass: one atomic per wavefront, slots in order
    const unsigned long long lead = __ballot(cell_hot && c.idx == 0);
    unsigned long long base = 0;
    if (lane == 0 && lead) base = atomicAdd(&st->hot_count, (unsigned long long)__popcll(lead));
    base = __shfl(base, 0, 64);
    if (cell_hot && c.idx == 0)
      hot[base + __popcll(lead & ((1ull << (c.slot * ND)) - 1ull))] = (unsigned int)cell;
  }
  const bool done_here = valid && !cell_hot;
  if (done_here) {
#pragma unroll
    for (int t = 0; t < N; t++)
#pragma unroll
      for (int v = 0; v < V; v++) sm.q[c.slot][t][c.idx][v] = q[t][v];
  }
  pred_sync();
  if (done_here) {
    // traces with every lane busy: the N ND1 = ND lanes of a cell take one
    // (time node t, transverse node j) each and write both ends e of every
    // direction d and variable v
    const int t = c.idx / ND1, j = c.idx % ND1;
    static_for<0, D>([&](auto dc) {
      constexpr int d = decltype(dc)::value;
      constexpr int stride = ipow_c(N, D - 1 - d);
      const int idx0 = (j / stride) * stride * N + j % stride;  // node with k_d = 0
      double qm[N][V];
#pragma unroll
      for (int m = 0; m < N; m++)
#pragma unroll
        for (int v = 0; v < V; v++) qm[m][v] = sm.q[c.slot][t][idx0 + m * stride][v];
#pragma unroll
      for (int e = 0; e < 2; e++)
#pragma unroll
        for (int v = 0; v < V; v++) {
          double acc = (e ? T_END1[0] : T_END0[0]) * qm[0][v];
#pragma unroll
          for (int m = 1; m < N; m++) acc = dfma(e ? T_END1[m] : T_END0[m], qm[m][v], acc);
          const int plane = (((d * 2 + e) * N + t) * ND1 + j) * V + v;
          tr[(long)plane * g.nwc + cell] = acc;
        }
    });
    if (qh) {
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int v = 0; v < V; v++) qh[(long)((t * ND + c.idx) * V + v) * g.nwc + cell] = q[t][v];
    }
  }
#if PDE_SECOND
  grad_traces(c, sm, g, cell, done_here, trd);
#endif
#if PDE_USE_S || PDE_USE_B
  // volume term (fv.cpp:33-85): per node sum over t, then over nodes in order
  double part[V];
  if (done_here) {
#pragma unroll
    for (int t = 0; t < N; t++) {
      double s[V];
      callS(s, q[t]);
#if PDE_USE_B
      static_for<0, D>([&](auto dc) {
        constexpr int d = decltype(dc)::value;
        double dqd[V], b[V * V], bd[V];
        lds_deriv<d>(c, sm.q, t, 1.0 / g.dX[d], dqd);
        B(b, q[t], d);
        b_times<PDE_B_QUIRK ? 2 : 0>(b, dqd, bd);
#pragma unroll
        for (int v = 0; v < V; v++) s[v] = s[v] - bd[v];
      });
#endif
      double tmp = dt * T_W[t];
#pragma unroll
      for (int d = 0; d < D; d++) tmp = tmp * c.wk(d);
#pragma unroll
      for (int v = 0; v < V; v++) part[v] = (t == 0) ? tmp * s[v] : dfma(tmp, s[v], part[v]);
    }
  }
  pred_sync();
  if (done_here) {
#pragma unroll
    for (int v = 0; v < V; v++) sm.q[c.slot][0][c.idx][v] = part[v];
  }
  pred_sync();
  // the cell's lane idx sums variable v = idx (+ ND, ...) over the nodes in order
  if (done_here) {
#pragma unroll
    for (int v0 = 0; v0 < V; v0 += ND) {
      const int v = v0 + c.idx;
      if (v < V) {
        double acc = sm.q[c.slot][0][0][v];
#pragma unroll
        for (int i = 1; i < ND; i++) acc = acc + sm.q[c.slot][0][i][v];
        vol[(long)v * g.nwc + cell] = acc;
      }
    }
  }
#endif
  // ---- statistics (one contribution per cell)
  unsigned long long mine = (done_here && c.idx == 0) ? (unsigned long long)it : 0ull;
  unsigned long long cnt = (done_here && c.idx == 0) ? 1ull : 0ull;
  unsigned long long cap = (done_here && c.idx == 0 && it >= maxit) ? 1ull : 0ull;
  unsigned long long fb = (done_here && c.idx == 0) ? (unsigned long long)fellback : 0ull;
  unsigned int mx = (done_here && c.idx == 0) ? (unsigned int)it : 0u;
  if (!done_here) n_full = n_short = 0u;
  mine = wave_sum_u64(mine);
  cnt = wave_sum_u64(cnt);
  cap = wave_sum_u64(cap);
  fb = wave_sum_u64(fb);
  mx = wave_max_u32(mx);
  const unsigned long long w_full = wave_sum_u64(n_full), w_short = wave_sum_u64(n_short);
  if (lane == 0) {
    unsigned long long* ps = st->pstat[b.x & 63];
    atomicAdd(&ps[0], mine);
    atomicAdd(&ps[1], cnt);
    if (cap) atomicAdd(&ps[2], cap);
    if (fb) atomicAdd(&ps[3], fb);
    atomicMax(&ps[4], (unsigned long long)mx);
    if (w_full) atomicAdd(&ps[5], w_full);
    if (w_short) atomicAdd(&ps[6], w_short);
  }
}

extern "C" __global__ void __launch_bounds__(64)
k_predictor_static(const double* __restrict__ wh, const double* __restrict__ guess,
                   double* __restrict__ tr, double* __restrict__ vol, double* __restrict__ qh,
                   StepState* st, Geom g, double newton_tol, int newton_maxit,
                   double* __restrict__ trd) {
  __shared__ PredLDS sm;
  predictor_static_body(sm, wh, guess, tr, vol, qh, st, g, newton_tol, newton_maxit, trd,
                        hw_blk());
}

#if PDE_STIFF && !PDE_GUESS
// The split stiff predictor: the cold pass (no full source-Jacobian solve in
// the code, so fewer registers and more wavefronts per SIMD) over every
// cell, then the hot pass over the cells it handed on.
extern "C" __global__ void __launch_bounds__(64)
k_predictor_cold(const double* __restrict__ wh, double* __restrict__ tr, double* __restrict__ vol,
                 double* __restrict__ qh, StepState* st, Geom g, double newton_tol,
                 int newton_maxit, unsigned int* __restrict__ hot) {
  __shared__ PredLDS sm;
  predictor_static_body<1>(sm, wh, nullptr, tr, vol, qh, st, g, newton_tol, newton_maxit,
                           nullptr, hw_blk(), hot);
}
// The cold pass with the register budget of three wavefronts per SIMD (168
// registers; whatever does not fit spills), for the autotuner to weigh
// against the two-wavefront build.
extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_predictor_cold3(const double* __restrict__ wh, double* __restrict__ tr, double* __restrict__ vol,
                  double* __restrict__ qh, StepState* st, Geom g, double newton_tol,
                  int newton_maxit, unsigned int* __restrict__ hot) {
  __shared__ PredLDS sm;
  predictor_static_body<1>(sm, wh, nullptr, tr, vol, qh, st, g, newton_tol, newton_maxit,
                           nullptr, hw_blk(), hot);
}
extern "C" __global__ void __launch_bounds__(64)
k_predictor_hot(const double* __restrict__ wh, double* __restrict__ tr, double* __restrict__ vol,
                double* __restrict__ qh, StepState* st, Geom g, double newton_tol,
                int newton_maxit, unsigned int* __restrict__ hot) {
  __shared__ PredLDS sm;
  predictor_static_body<2>(sm, wh, nullptr, tr, vol, qh, st, g, newton_tol, newton_maxit,
                           nullptr, hw_blk(), hot);
}
#endif

// The same with the register budget of two wavefronts per SIMD (spilling the
// rest), for latency hiding where the single-wavefront kernel stalls.
extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_predictor_static2(const double* __restrict__ wh, const double* __restrict__ guess,
                    double* __restrict__ tr, double* __restrict__ vol, double* __restrict__ qh,
                    StepState* st, Geom g, double newton_tol, int newton_maxit,
                   double* __restrict__ trd) {
  __shared__ PredLDS sm;
  predictor_static_body(sm, wh, guess, tr, vol, qh, st, g, newton_tol, newton_maxit, trd,
                        hw_blk());
}

#if PDE_MIXED
// ---------------------------------------------- mixed precision: FP32 phase
// The predictor's space-time iteration with its contractions (spatial
// derivatives of q and of F, the time operator of the residual), F and S in
// FP32; the stiff Newton correction still solves with the FP64 source
// Jacobian (newton_node).  It runs until the FP32 corrections fall below
// tol32 (1 + |q|), at most maxit32 iterations, and leaves its iterate as the
// initial guess of the FP64 predictor, which iterates on from it to the
// unchanged tolerance (the FP64 residual correction).  A cell whose FP32
// iterate is not finite hands over the plain guess instead.
struct PredLDS32 {
  float q[CPW][N][ND][V];
#if PDE_USE_F
  float f[D][CPW][N][ND][V];
#endif
#if PDE_NODE_TAB_LDS
  NodeTab nt;
#endif
};

DEV void F32(float* out, const float* q, const float* dq, int d) {
  f32::R o[V], qq[V], dd[D * V];
#pragma unroll
  for (int v = 0; v < V; v++) qq[v] = f32::R(q[v]);
#pragma unroll
  for (int i = 0; i < D * V; i++) dd[i] = f32::R(dq[i]);
  f32::F(o, qq, dd, d);
#pragma unroll
  for (int v = 0; v < V; v++) out[v] = o[v].v;
}

DEV void S32(float* out, const float* q) {
#if PDE_USE_S
  f32::R o[V], qq[V];
#pragma unroll
  for (int v = 0; v < V; v++) qq[v] = f32::R(q[v]);
  f32::S(o, qq);
#pragma unroll
  for (int v = 0; v < V; v++) out[v] = o[v].v;
#else
#pragma unroll
  for (int v = 0; v < V; v++) out[v] = 0.0f;
#endif
}

template <int d>
DEV void lds_deriv32(const NodeCtx& c, const float (*buf)[N][ND][V], int t, float rdx, float* out) {
  constexpr int stride = ipow_c(N, D - 1 - d);
#pragma unroll
  for (int v = 0; v < V; v++) {
    float acc = (float)c.der(d, 0) * buf[c.slot][t][c.base[d]][v];
#pragma unroll
    for (int j = 1; j < N; j++)
      acc = __builtin_fmaf((float)c.der(d, j), buf[c.slot][t][c.base[d] + j * stride][v], acc);
    out[v] = acc * rdx;
  }
}

// rhs of the space-time system in FP32 (node_rhs); CONST: q is constant in
// time, so F, S and derivatives are evaluated once and reused for every t.
template <bool CONST>
DEV void node_rhs32(const NodeCtx& c, PredLDS32& sm, const float (&q)[N][V], const float* w,
                    float dt, const float* rdx, bool active, float (&rhs)[N][V]) {
  constexpr int NTE = CONST ? 1 : N;  // time nodes evaluated
  float dqs[NTE][D][V];
  if (active) {
#pragma unroll
    for (int t = 0; t < NTE; t++) {
      static_for<0, D>([&](auto dc) {
        constexpr int d = decltype(dc)::value;
        lds_deriv32<d>(c, sm.q, t, rdx[d], dqs[t][d]);
      });
#if PDE_USE_F
#pragma unroll
      for (int d = 0; d < D; d++) {
        float f[V];
        F32(f, q[t], &dqs[t][0][0], d);
#pragma unroll
        for (int v = 0; v < V; v++) sm.f[d][c.slot][t][c.idx][v] = f[v];
      }
#endif
    }
  }
#if PDE_USE_F
  pred_sync();
#endif
  if (active) {
    float r[NTE][V];
#pragma unroll
    for (int t = 0; t < NTE; t++) {
      S32(r[t], q[t]);
#if PDE_USE_F
      static_for<0, D>([&](auto dc) {
        constexpr int d = decltype(dc)::value;
        float df[V];
        lds_deriv32<d>(c, sm.f[d], t, rdx[d], df);
#pragma unroll
        for (int v = 0; v < V; v++) r[t][v] = r[t][v] - df[v];
      });
#endif
    }
#pragma unroll
    for (int t = 0; t < N; t++)
#pragma unroll
      for (int v = 0; v < V; v++)
        rhs[t][v] = (r[CONST ? 0 : t][v] * (float)c.ct(t)) * dt + (float)c.c0(t) * w[v];
  }
}

extern "C" __global__ void __launch_bounds__(64)
k_predictor_f32(const double* __restrict__ wh, const double* __restrict__ gin,
                double* __restrict__ gout, StepState* st, Geom g, float tol32, int maxit32) {
  if (st->done) return;
  __shared__ PredLDS32 sm;
  const float dt = (float)st->dt;
  const double dtd = st->dt;
  const int lane = threadIdx.x;
  NodeCtx c;
  node_ctx(c, lane, PRED_NODE_TAB(sm));
  const long cell = (long)blockIdx.x * CPW + c.slot;
  const bool valid = (c.slot < CPW) && (cell < g.nwc);
  float rdx[D];
#pragma unroll
  for (int d = 0; d < D; d++) rdx[d] = (float)(1.0 / g.dX[d]);
  float w[V], q[N][V], rhs[N][V];
  if (valid) {
#pragma unroll
    for (int v = 0; v < V; v++) w[v] = (float)wh[(long)(c.idx * V + v) * g.nwc + cell];
#pragma unroll
    for (int t = 0; t < N; t++)
#pragma unroll
      for (int v = 0; v < V; v++)
        q[t][v] = gin ? (float)gin[(long)((t * ND + c.idx) * V + v) * g.nwc + cell] : w[v];
  }
  const unsigned long long cellmask =
      (ND == 64 ? ~0ull : ((1ull << ND) - 1ull)) << (c.slot < CPW ? c.slot * ND : 0);
  bool active = valid;
  int it = 0;
  while (__ballot(active) != 0ull) {
    if (active) {
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int v = 0; v < V; v++) sm.q[c.slot][t][c.idx][v] = q[t][v];
    }
    pred_sync();
    if (it == 0 && !gin)
      node_rhs32<true>(c, sm, q, w, dt, rdx, active, rhs);
    else
      node_rhs32<false>(c, sm, q, w, dt, rdx, active, rhs);
    bool more = false;
    if (active) {
#if PDE_STIFF
      // residual G = rhs - (DG_MAT (x) W) q, then the FP64 Newton correction
      double qd[N][V], G[N][V];
      const float wp = (float)c.WP();
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int v = 0; v < V; v++) {
          float acc = ((float)T_DGMAT[t][0] * wp) * q[0][v];
#pragma unroll
          for (int s2 = 1; s2 < N; s2++) acc = __builtin_fmaf((float)T_DGMAT[t][s2] * wp, q[s2][v], acc);
          G[t][v] = (double)(rhs[t][v] - acc);
          qd[t][v] = (double)q[t][v];
        }
      (void)newton_node(c, qd, G, dtd);
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int v = 0; v < V; v++) {
          const float dq = (float)G[t][v];
          if (!(__builtin_fabsf(dq) <= tol32 * (1.0f + __builtin_fabsf(q[t][v])))) more = true;
          q[t][v] = q[t][v] + dq;
        }
#else
      // Picard: q1 = DG_MAT^-1 rhs / prod W
      const float iwp = (float)c.IWP();
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int v = 0; v < V; v++) {
          float acc = (float)T_DGINV[t][0] * rhs[0][v];
#pragma unroll
          for (int s2 = 1; s2 < N; s2++) acc = __builtin_fmaf((float)T_DGINV[t][s2], rhs[s2][v], acc);
          const float q1 = acc * iwp;
          if (!(__builtin_fabsf(q1 - q[t][v]) <= tol32 * (1.0f + __builtin_fabsf(q[t][v]))))
            more = true;
          q[t][v] = q1;
        }
#endif
    }
    const unsigned long long mb = __ballot(more);
    if (active) {
      it++;
      if ((mb & cellmask) == 0ull || it >= maxit32) active = false;
    }
    pred_sync();
  }
  // hand over: the FP32 iterate, or the plain guess where it is not finite
  bool bad = false;
  if (valid) {
#pragma unroll
    for (int t = 0; t < N; t++)
#pragma unroll
      for (int v = 0; v < V; v++)
        if (!(__builtin_fabsf(q[t][v]) <= 3.4e38f)) bad = true;
  }
  const bool cell_bad = (__ballot(bad) & cellmask) != 0ull;
  if (valid) {
#pragma unroll
    for (int t = 0; t < N; t++)
#pragma unroll
      for (int v = 0; v < V; v++)
        gout[(long)((t * ND + c.idx) * V + v) * g.nwc + cell] =
            cell_bad ? wh[(long)(c.idx * V + v) * g.nwc + cell] : (double)q[t][v];
  }
  unsigned long long mine = (valid && c.idx == 0) ? (unsigned long long)it : 0ull;
  mine = wave_sum_u64(mine);
  if (lane == 0 && mine) atomicAdd(&st->pstat[blockIdx.x & 63][7], mine);
}
#endif

// ------------------------------------------------------------ face flux
// Rusanov flux and B path integral at every face point (fluxes.cpp:72-116).
// Faces along d: (nx_d + 1) x prod_{e != d} nx_e, all touching an interior cell.
template <int d>
DEV void face_cells(const Geom& g, long f, long* wl, long* wr) {
  if constexpr (D == 1) {
    *wl = f;
    *wr = f + 1;
  } else if constexpr (d == 0) {
    long f0 = f / g.nx[1], f1 = f % g.nx[1];
    *wl = f0 * g.nw[1] + (f1 + 1);
    *wr = *wl + g.nw[1];
  } else {
    long f0 = f / (g.nx[1] + 1), f1 = f % (g.nx[1] + 1);
    *wl = (f0 + 1) * g.nw[1] + f1;
    *wr = *wl + 1;
  }
}

// Numerical flux at one face point along axis d (fv.cpp:124-198): Rusanov
// (fluxes.cpp:72-101), Roe or Osher (fluxes.cpp:20-70), the viscous term
// (fluxes.cpp:96-100) and the non-conservative path integral Bint
// (fluxes.cpp:103-116), from the two cells' face traces.
template <int d, int BS = 256>
DEV void flux_dir(const double* __restrict__ tr, double* __restrict__ ff, double* __restrict__ fb,
                  const StepState* st, const Geom& g, const double* __restrict__ trd,
                  const Blk& b) {
  const long nf = g.nf[d];
  long f, wl, wr;
  int j, t;
  if (g.idx32) {
    // 32-bit indices, divisions by nf and by the faces per row as
    // multiply-shifts (nf N ND1 and nwc below 2^31)
    const unsigned item = b.x * (unsigned)BS + threadIdx.x;
    if (item >= (unsigned)nf * (unsigned)(N * ND1)) return;
    const unsigned rest = udiv(item, g.fmul[d], g.fsh[d]);
    const unsigned f32 = item - rest * (unsigned)nf;
    j = (int)(rest % ND1);
    t = (int)(rest / ND1);
    f = f32;
    if constexpr (D == 1) {
      wl = f32;
      wr = f32 + 1;
    } else {
      const unsigned row = d == 0 ? (unsigned)g.nx[1] : (unsigned)g.nx[1] + 1;
      const unsigned f0 = udiv(f32, g.rmul[d], g.rsh[d]), f1 = f32 - f0 * row;
      const unsigned nw1 = (unsigned)g.nw[1];
      const unsigned l = d == 0 ? f0 * nw1 + (f1 + 1) : (f0 + 1) * nw1 + f1;
      wl = l;
      wr = d == 0 ? l + nw1 : l + 1;
    }
  } else {
    const long item = (long)b.x * BS + threadIdx.x;
    if (item >= nf * N * ND1) return;
    f = item % nf;
    j = (int)((item / nf) % ND1);
    t = (int)(item / (nf * ND1));
    face_cells<d>(g, f, &wl, &wr);
  }
  double qL[V], qR[V];
#pragma unroll
  for (int v = 0; v < V; v++) {
    qL[v] = tr[(long)((((d * 2 + 1) * N + t) * ND1 + j) * V + v) * g.nwc + wl];
    qR[v] = tr[(long)((((d * 2 + 0) * N + t) * ND1 + j) * V + v) * g.nwc + wr];
  }
  const long o = (long)((t * ND1 + j) * V) * nf + f;
#if PDE_USE_F
  {
    // face gradients (viscous F only; zero otherwise, as the reference passes
    // unused dQ)
    double dqL[D * V], dqR[D * V];
#pragma unroll
    for (int i = 0; i < D * V; i++) {
#if PDE_SECOND
      dqL[i] = trd[(long)((((d * 2 + 1) * N + t) * ND1 + j) * D * V + i) * g.nwc + wl];
      dqR[i] = trd[(long)((((d * 2 + 0) * N + t) * ND1 + j) * D * V + i) * g.nwc + wr];
#else
      dqL[i] = 0.0;
      dqR[i] = 0.0;
#endif
    }
    double fl[V], fr[V], fv[V];
#if PDE_FLUX == 0
    // Rusanov: max(lam(qL), lam(qR)) (qL - qR) + F(qL) + F(qR)  (fluxes.cpp:72-101)
    double l1 = max_abs_eigs(qL, dqL, d, fl);
    double l2 = max_abs_eigs(qR, dqR, d, fr);
    double s = dmax(l1, l2);
#pragma unroll
    for (int v = 0; v < V; v++) fv[v] = s * (qL[v] - qR[v]) + (fl[v] + fr[v]);
#else
    // Roe: |sum_i W_i A(q_i)| (qL - qR); Osher: sum_i W_i |A(q_i)| (qL - qR), on
    // the path q_i = qR + x_i (qL - qR)  (fluxes.cpp:20-70)
    F(fl, qL, dqL, d);
    F(fr, qR, dqR, d);
    double Dq[V], q[V], A[V * V], ftmp[V], Dsum[V], dqp[D * V];
#if PDE_FLUX == 1
    double Mx[V * V];
#endif
#pragma unroll
    for (int v = 0; v < V; v++) Dq[v] = qL[v] - qR[v];
#pragma unroll
    for (int i = 0; i < N; i++) {
#pragma unroll
      for (int v = 0; v < V; v++) q[v] = dfma(T_X[i], Dq[v], qR[v]);
#pragma unroll
      for (int k = 0; k < D * V; k++) dqp[k] = dfma(T_X[i], dqL[k] - dqR[k], dqR[k]);
      system_matrix(q, dqp, d, A, ftmp);
#if PDE_FLUX == 1
#pragma unroll
      for (int k = 0; k < V * V; k++) Mx[k] = (i == 0) ? T_W[0] * A[k] : dfma(T_W[i], A[k], Mx[k]);
#else
      double yv[V];
      abs_apply(A, Dq, yv);
#pragma unroll
      for (int v = 0; v < V; v++) Dsum[v] = (i == 0) ? T_W[0] * yv[v] : dfma(T_W[i], yv[v], Dsum[v]);
#endif
    }
#if PDE_FLUX == 1
    abs_apply(Mx, Dq, Dsum);
#endif
#pragma unroll
    for (int v = 0; v < V; v++) fv[v] = Dsum[v] + (fl[v] + fr[v]);
#endif
#if PDE_SECOND
    // viscous dissipation (fluxes.cpp:96-100)
    const double m1 = max_abs_eigs_second_order<d>(qL, dqL, g.dX[d]);
    const double m2 = max_abs_eigs_second_order<d>(qR, dqR, g.dX[d]);
    const double s2 = dmax(m1, m2);
#pragma unroll
    for (int v = 0; v < V; v++) fv[v] = fv[v] + s2 * (qL[v] - qR[v]);
#endif
#pragma unroll
    for (int v = 0; v < V; v++) ff[o + (long)v * nf] = fv[v];
  }
#endif
#if PDE_USE_B
  {
    double Dq[V], Mx[V * V], b[V * V], q[V];
#pragma unroll
    for (int v = 0; v < V; v++) Dq[v] = qR[v] - qL[v];
#pragma unroll
    for (int i = 0; i < N; i++) {
#pragma unroll
      for (int v = 0; v < V; v++) q[v] = dfma(T_X[i], Dq[v], qL[v]);
      B(b, q, d);
#pragma unroll
      for (int k = 0; k < V * V; k++) Mx[k] = (i == 0) ? T_W[0] * b[k] : dfma(T_W[i], b[k], Mx[k]);
    }
#pragma unroll
    for (int v = 0; v < V; v++) {
      double acc = Mx[v * V] * Dq[0];
#pragma unroll
      for (int k = 1; k < V; k++) acc = dfma(Mx[v * V + k], Dq[k], acc);
      fb[o + (long)v * nf] = acc;
    }
  }
#endif
}

// Workgroup sizes 256 (k_flux0, k_flux1), 128 and 64 (..._128, ..._64): the
// autotuner picks one per problem; every face point's arithmetic is the same.
#define PYPDE_FLUX_KERNEL(NAME, DIR, BS)                                                       \
  extern "C" __global__ void __launch_bounds__(BS)                                             \
  NAME(const double* __restrict__ tr, double* __restrict__ ff, double* __restrict__ fb,        \
       const StepState* st, Geom g, const double* __restrict__ trd) {                          \
    if (st->done) return;                                                                      \
    flux_dir<DIR, BS>(tr, ff, fb, st, g, trd, hw_blk());                                       \
  }
PYPDE_FLUX_KERNEL(k_flux0, 0, 256)
PYPDE_FLUX_KERNEL(k_flux0_128, 0, 128)
PYPDE_FLUX_KERNEL(k_flux0_64, 0, 64)
#if PDE_D >= 2
PYPDE_FLUX_KERNEL(k_flux1, 1, 256)
PYPDE_FLUX_KERNEL(k_flux1_128, 1, 128)
PYPDE_FLUX_KERNEL(k_flux1_64, 1, 64)
#endif

// ------------------------------------------------------------- corrector
// u += volume term, then the face contributions in the reference's
// face-cell order (fv.cpp:148-197): faces c+1-e_d (c on the right, d
// ascending, each over t then transverse node), then face-cell c+1 (c on the
// left, over t, then d, then transverse node).  Also writes the snapshot row
// and raises the NaN flag.
struct FaceBufs {
  const double* ff[2];
  const double* fb[2];
};

DEV void advance_step(StepState* st, int ndt);

DEV void update_cell(double* __restrict__ u, const double* __restrict__ vol, const FaceBufs& fbufs,
                     double* __restrict__ ret, StepState* st, const Geom& g, const Blk& b) {
  const long c = (long)b.x * 256 + threadIdx.x;
  if (c >= g.ncell) return;
  const double dt = st->dt;
  int cc[2];
  cc[0] = (D == 1) ? (int)c : (int)(c / g.nx[1]);
  cc[1] = (D == 1) ? 0 : (int)(c % g.nx[1]);
  double uc[V];
#pragma unroll
  for (int v = 0; v < V; v++) uc[v] = u[(long)v * g.ncell + c];
#if PDE_USE_S || PDE_USE_B
  {
    long wc = (D == 1) ? (long)(cc[0] + 1) : (long)(cc[0] + 1) * g.nw[1] + (cc[1] + 1);
#pragma unroll
    for (int v = 0; v < V; v++) uc[v] = uc[v] + vol[(long)v * g.nwc + wc];
  }
#endif
#if PDE_USE_F || PDE_USE_B
  long fidx[2][2];  // [side][d]
#pragma unroll
  for (int d = 0; d < D; d++)
#pragma unroll
    for (int side = 0; side < 2; side++) {
      if (D == 1) {
        fidx[side][d] = cc[0] + side;
      } else if (d == 0) {
        fidx[side][d] = (long)(cc[0] + side) * g.nx[1] + cc[1];
      } else {
        fidx[side][d] = (long)cc[0] * (g.nx[1] + 1) + cc[1] + side;
      }
    }
  auto contrib = [&](int d, int side, int t) {
    const double c0 = (dt * T_W[t]) / (2.0 * g.dX[d]);
    const long nf = g.nf[d];
#pragma unroll
    for (int j = 0; j < ND1; j++) {
      double tmp = c0;
      if (D == 2) tmp = tmp * T_W[j];
      const long o = (long)((t * ND1 + j) * V) * nf + fidx[side][d];
#pragma unroll
      for (int v = 0; v < V; v++) {
#if PDE_USE_F
        double f = fbufs.ff[d][o + (long)v * nf];
#else
        double f = 0.0;
#endif
#if PDE_USE_B
        double b = fbufs.fb[d][o + (long)v * nf];
#else
        double b = 0.0;
#endif
        uc[v] = side == 0 ? uc[v] - tmp * (b - f) : uc[v] - tmp * (b + f);
      }
    }
  };
#pragma unroll
  for (int d = 0; d < D; d++)
#pragma unroll
    for (int t = 0; t < N; t++) contrib(d, 0, t);
#pragma unroll
  for (int t = 0; t < N; t++)
#pragma unroll
    for (int d = 0; d < D; d++) contrib(d, 1, t);
#endif
  bool bad = false;
#pragma unroll
  for (int v = 0; v < V; v++) {
    u[(long)v * g.ncell + c] = uc[v];
    if (uc[v] != uc[v]) bad = true;
  }
  if (bad) atomicOr(&st->nan_flag, 1u);
  const double tn = st->t + dt;
  const int push = st->push;
  if (ret && push < g.ndt && tn >= (double)(push + 1) / (double)g.ndt * st->tf) {
#pragma unroll
    for (int v = 0; v < V; v++) ret[((long)push * g.ncell + c) * V + v] = uc[v];
  }
}

// Every corrector kernel ends the step: its last workgroup runs advance_step
// once all the others have written u and raised the NaN flag.
extern "C" __global__ void __launch_bounds__(256)
k_update(double* __restrict__ u, const double* __restrict__ vol, FaceBufs fbufs,
         double* __restrict__ ret, StepState* st, Geom g) {
  if (st->done) return;
  const Blk b = hw_blk();
  update_cell(u, vol, fbufs, ret, st, g, b);
  if (last_workgroup(&st->step_ticket, b.n)) advance_step(st, g.ndt);
}

// The corrector with one variable per thread and R consecutive x-rows per
// thread (thread = (v, row group, y)): each x-face flux row between two of
// its cells is read once and kept in registers for the next row, so the face
// fluxes are read close to once (x faces (R+1)/R times).  Every cell and
// variable accumulates the same terms in the same order as k_update, so the
// results are identical; the autotuner picks R per problem.
template <int R>
DEV void update_rows(double* __restrict__ u, const double* __restrict__ vol, const FaceBufs& fbufs,
                     double* __restrict__ ret, StepState* st, const Geom& g, const Blk& b) {
  const int nx = g.nx[0], ny = g.nx[1];
  const int nxg = (nx + R - 1) / R;
  const long idx = (long)b.x * 256 + threadIdx.x;
  if (idx >= (long)V * nxg * ny) return;
  const int cy = (int)(idx % ny);
  const long rest = idx / ny;
  const int xg = (int)(rest % nxg);
  const int v = (int)(rest / nxg);
  const double dt = st->dt;
  const double tn = st->t + dt;
  const int push = st->push;
  const bool snap = ret && push < g.ndt && tn >= (double)(push + 1) / (double)g.ndt * st->tf;
  double c0[D][N];  // (dt W_t) / (2 dX_d)
#pragma unroll
  for (int d = 0; d < D; d++)
#pragma unroll
    for (int t = 0; t < N; t++) c0[d][t] = (dt * T_W[t]) / (2.0 * g.dX[d]);
  auto coef = [&](int d, int t, int j) {
    double tmp = c0[d][t];
    if (D == 2) tmp = tmp * T_W[j];
    return tmp;
  };
  auto plane = [&](int t, int j) { return (long)((t * ND1 + j) * V + v); };
  const long nf0 = g.nf[0];
  const int cx0 = xg * R;
  // left x face of the first row: f = cx0 * ny + cy
  double xlf[N][ND1], xlb[N][ND1];
#pragma unroll
  for (int t = 0; t < N; t++)
#pragma unroll
    for (int j = 0; j < ND1; j++) {
      const long f = (long)cx0 * ny + cy;
      xlf[t][j] = PDE_USE_F ? fbufs.ff[0][plane(t, j) * nf0 + f] : 0.0;
      xlb[t][j] = PDE_USE_B ? fbufs.fb[0][plane(t, j) * nf0 + f] : 0.0;
    }
  bool bad = false;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int cx = cx0 + r;
    if (cx >= nx) break;
    const long c = (long)cx * ny + cy;
    double uc = u[(long)v * g.ncell + c];
#if PDE_USE_S || PDE_USE_B
    {
      const long wc = (D == 1) ? (long)(cx + 1) : (long)(cx + 1) * g.nw[1] + (cy + 1);
      uc = uc + vol[(long)v * g.nwc + wc];
    }
#endif
    double xrf[N][ND1], xrb[N][ND1];
#pragma unroll
    for (int t = 0; t < N; t++)
#pragma unroll
      for (int j = 0; j < ND1; j++) {
        const long f = (long)(cx + 1) * ny + cy;
        xrf[t][j] = PDE_USE_F ? fbufs.ff[0][plane(t, j) * nf0 + f] : 0.0;
        xrb[t][j] = PDE_USE_B ? fbufs.fb[0][plane(t, j) * nf0 + f] : 0.0;
      }
#if PDE_USE_F || PDE_USE_B
    // faces c + 1 - e_d (c on the right), d ascending, over t then j
#pragma unroll
    for (int t = 0; t < N; t++)
#pragma unroll
      for (int j = 0; j < ND1; j++) uc = uc - coef(0, t, j) * (xlb[t][j] - xlf[t][j]);
    if constexpr (D == 2) {
      const long nf1 = g.nf[1];
      const long fb0 = (long)cx * (ny + 1) + cy;
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int j = 0; j < ND1; j++) {
          const double f = PDE_USE_F ? fbufs.ff[1][plane(t, j) * nf1 + fb0] : 0.0;
          const double b = PDE_USE_B ? fbufs.fb[1][plane(t, j) * nf1 + fb0] : 0.0;
          uc = uc - coef(1, t, j) * (b - f);
        }
      // face-cell c + 1 (c on the left): over t, then d, then j
#pragma unroll
      for (int t = 0; t < N; t++) {
#pragma unroll
        for (int j = 0; j < ND1; j++) uc = uc - coef(0, t, j) * (xrb[t][j] + xrf[t][j]);
#pragma unroll
        for (int j = 0; j < ND1; j++) {
          const double f = PDE_USE_F ? fbufs.ff[1][plane(t, j) * nf1 + fb0 + 1] : 0.0;
          const double b = PDE_USE_B ? fbufs.fb[1][plane(t, j) * nf1 + fb0 + 1] : 0.0;
          uc = uc - coef(1, t, j) * (b + f);
        }
      }
    } else {
#pragma unroll
      for (int t = 0; t < N; t++)
#pragma unroll
        for (int j = 0; j < ND1; j++) uc = uc - coef(0, t, j) * (xrb[t][j] + xrf[t][j]);
    }
#endif
    u[(long)v * g.ncell + c] = uc;
    if (uc != uc) bad = true;
    if (snap) ret[((long)push * g.ncell + c) * V + v] = uc;
#pragma unroll
    for (int t = 0; t < N; t++)
#pragma unroll
      for (int j = 0; j < ND1; j++) {
        xlf[t][j] = xrf[t][j];
        xlb[t][j] = xrb[t][j];
      }
  }
  if (bad) atomicOr(&st->nan_flag, 1u);
}

#define PYPDE_UPDATE_ROWS(R_)                                                                  \
  extern "C" __global__ void __launch_bounds__(256)                                           \
  k_update_r##R_(double* __restrict__ u, const double* __restrict__ vol, FaceBufs fbufs,      \
                 double* __restrict__ ret, StepState* st, Geom g) {                            \
    if (st->done) return;                                                                      \
    const Blk b = hw_blk();                                                                    \
    update_rows<R_>(u, vol, fbufs, ret, st, g, b);                                             \
    if (last_workgroup(&st->step_ticket, b.n)) advance_step(st, g.ndt);                        \
  }
PYPDE_UPDATE_ROWS(1)
PYPDE_UPDATE_ROWS(2)
PYPDE_UPDATE_ROWS(4)
PYPDE_UPDATE_ROWS(8)

// t += dt; count++; snapshot row; NaN and end-of-run checks (iterator.cpp:129-147).
// Run by the first wavefront of the corrector's last workgroup, which also
// folds the predictor's spread statistics into the totals (lane i takes slot
// i, then sums and a max across the wavefront) and clears the slots.
DEV void advance_step(StepState* st, int ndt) {
  if (threadIdx.x >= 64) return;
  {
    unsigned long long* ps = st->pstat[threadIdx.x];
    unsigned long long a0 = wave_sum_u64(ps[0]), a1 = wave_sum_u64(ps[1]);
    unsigned long long a2 = wave_sum_u64(ps[2]), a3 = wave_sum_u64(ps[3]);
    unsigned long long a5 = wave_sum_u64(ps[5]), a6 = wave_sum_u64(ps[6]);
    unsigned long long a7 = wave_sum_u64(ps[7]);
    unsigned long long m4 = ps[4];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      unsigned long long y = __shfl_xor(m4, o, 64);
      m4 = m4 > y ? m4 : y;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) ps[k] = 0ull;
    if (threadIdx.x == 0) {
      st->pred_iters += a0;
      st->pred_cells += a1;
      st->pred_capped += a2;
      st->pred_fallbacks += a3;
      st->pred_nfull += a5;
      st->pred_nshort += a6;
      st->pred_iters32 += a7;
      if ((unsigned int)m4 > st->pred_maxit) st->pred_maxit = (unsigned int)m4;
    }
  }
  if (threadIdx.x != 0) return;
  st->t = st->t + st->dt;
  st->count += 1;
  if (st->push < ndt && st->t >= (double)(st->push + 1) / (double)ndt * st->tf) st->push += 1;
  // raised by the corrector's other workgroups during this launch: read at
  // the device's coherence point
  if (atomicAdd(&st->nan_flag, 0u)) {
    st->done = 1;
    st->nan_step = (int)st->count;
  }
  if (!(st->dt > 0.0)) st->done = 1;  // no progress possible (infinite wave speed)
  if (!(st->t < st->tf) || (st->max_steps > 0 && st->count >= st->max_steps)) st->done = 1;
}

// Interleaved [cell][v] <-> structure-of-arrays [v][cell].
extern "C" __global__ void k_aos_to_soa(const double* __restrict__ a, double* __restrict__ s, long ncell) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= ncell * V) return;
  long c = i / V;
  int v = (int)(i % V);
  s[(long)v * ncell + c] = a[i];
}
extern "C" __global__ void k_soa_to_aos(const double* __restrict__ s, double* __restrict__ a, long ncell) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= ncell * V) return;
  long c = i / V;
  int v = (int)(i % V);
  a[i] = s[(long)v * ncell + c];
}

// Reconstruction planes [node * V + v][cell] -> the reference's weno_solver
// layout [cell][node][v] (api.cpp:32-48).
extern "C" __global__ void k_planes_to_rows(const double* __restrict__ s, double* __restrict__ a,
                                            long ncell) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= ncell * ND * V) return;
  long c = i / (ND * V);
  int p = (int)(i % (ND * V));
  a[i] = s[(long)p * ncell + c];
}

// ------------------------------------------------ learned Newton guess (MLP)
// Training pairs: features x = (wh of the cell (N^D * V values), dt), targets
// y = (qh - w broadcast in time) / dt (N^(D+1) * V values), written as fp32
// columns col0 + cell of the caller's feature-major arrays.
constexpr int MLP_IN = ND * V + 1;
constexpr int MLP_OUT = NT * V;

extern "C" __global__ void __launch_bounds__(256)
k_harvest(const double* __restrict__ wh, const double* __restrict__ qh, const StepState* st,
          float* __restrict__ X, float* __restrict__ Y, long ldx, long col0, long nwc) {
  const long c = (long)blockIdx.x * 256 + threadIdx.x;
  if (c >= nwc) return;
  const double dt = st->dt;
  for (int k = 0; k < ND * V; k++) X[(long)k * ldx + col0 + c] = (float)wh[(long)k * nwc + c];
  X[(long)(ND * V) * ldx + col0 + c] = (float)dt;
  for (int o = 0; o < MLP_OUT; o++) {
    const int p = o % (ND * V);  // (node, v) of output o = (t, node, v)
    Y[(long)o * ldx + col0 + c] = (float)((qh[(long)o * nwc + c] - wh[(long)p * nwc + c]) / dt);
  }
}

#if PDE_GUESS
#ifndef MLP_HID
#define MLP_HID 128
#endif
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct MLPW {
  const float *W1t, *b1, *W2t, *b2, *W3t, *b3;  // Wt: out x in, row-major
  const float *in_mean, *in_rstd, *out_mean, *out_std;
};

DEV short to_bf16(float x) {
  unsigned int b = __float_as_uint(x);
  b += 0x7FFFu + ((b >> 16) & 1u);
  return (short)(b >> 16);
}
DEV float silu(float z) { return z / (1.0f + __expf(-z)); }

constexpr int KB_IN = (MLP_IN + 7) / 8;          // k-blocks of the input layer
constexpr int RB_HID = MLP_HID / 32;             // 32-row blocks of a hidden layer
constexpr int RB_OUT = (MLP_OUT + 31) / 32;
static_assert(MLP_HID % 32 == 0, "hidden width must be a multiple of 32");

// A fragment of Wt (rows rb*32.., k-block kb) from fp32 memory
DEV bf16x4 wfrag(const float* Wt, int ld, int rows, int cols, int rb, int kb, int lane) {
  bf16x4 a;
  const int r = rb * 32 + (lane & 31);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int k = kb * 8 + 4 * (lane >> 5) + i;
    a[i] = to_bf16((r < rows && k < cols) ? Wt[(long)r * ld + k] : 0.0f);
  }
  return a;
}

// Weight fragments of the three layers, packed once per attached model by
// k_mlp_pack in the order k_mlp_guess reads them: W1 [RB_HID][KB_IN], W2
// [RB_HID][KH], W3 [RB_OUT][KH], each one bf16x4 per lane; then the FP32
// vectors in_mean, in_rstd, b1, b2, b3, out_mean, out_std.
constexpr int KH = MLP_HID / 8;                                    // k-blocks of a hidden layer
constexpr int NFRAG = RB_HID * KB_IN + RB_HID * KH + RB_OUT * KH;  // fragments per model
constexpr int AUX_IN_MEAN = 0, AUX_IN_RSTD = MLP_IN, AUX_B1 = 2 * MLP_IN,
              AUX_B2 = AUX_B1 + MLP_HID, AUX_B3 = AUX_B2 + MLP_HID, AUX_OUT_MEAN = AUX_B3 + MLP_OUT,
              AUX_OUT_STD = AUX_OUT_MEAN + MLP_OUT, NAUX = AUX_OUT_STD + MLP_OUT;

DEV const float* aux_src(const MLPW& m, int i, int* k) {
  if (i < AUX_IN_RSTD) { *k = i; return m.in_mean; }
  if (i < AUX_B1) { *k = i - AUX_IN_RSTD; return m.in_rstd; }
  if (i < AUX_B2) { *k = i - AUX_B1; return m.b1; }
  if (i < AUX_B3) { *k = i - AUX_B2; return m.b2; }
  if (i < AUX_OUT_MEAN) { *k = i - AUX_B3; return m.b3; }
  if (i < AUX_OUT_STD) { *k = i - AUX_OUT_MEAN; return m.out_mean; }
  *k = i - AUX_OUT_STD;
  return m.out_std;
}

extern "C" __global__ void __launch_bounds__(256)
k_mlp_pack(MLPW m, bf16x4* __restrict__ pack, float* __restrict__ aux) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < (long)NFRAG * 64) {
    const int lane = (int)(i % 64), f = (int)(i / 64);
    if (f < RB_HID * KB_IN)
      pack[i] = wfrag(m.W1t, MLP_IN, MLP_HID, MLP_IN, f / KB_IN, f % KB_IN, lane);
    else if (f < RB_HID * KB_IN + RB_HID * KH) {
      const int h = f - RB_HID * KB_IN;
      pack[i] = wfrag(m.W2t, MLP_HID, MLP_HID, MLP_HID, h / KH, h % KH, lane);
    } else {
      const int h = f - RB_HID * KB_IN - RB_HID * KH;
      pack[i] = wfrag(m.W3t, MLP_HID, MLP_OUT, MLP_HID, h / KH, h % KH, lane);
    }
  }
  if (i < NAUX) {
    int k;
    const float* src = aux_src(m, (int)i, &k);
    aux[i] = src[k];
  }
}

// Fused inference: one wavefront infers 32 cells (lane % 32 = cell), four
// wavefronts per workgroup; the packed BF16 fragments (84 KB at 46-128-128-135)
// are read from global memory, where every wavefront finds them in the L2.
// The transposed form H^T = W^T X^T makes each layer's accumulators the next
// layer's B operands; each hidden activation (bias + SiLU, rounded to BF16)
// is formed once.
extern "C" __global__ void __launch_bounds__(256)
k_mlp_guess(const double* __restrict__ wh, const StepState* st, const bf16x4* __restrict__ pack,
            const float* __restrict__ aux, double* __restrict__ guess, long nwc) {
  if (st->done) return;
  const int lane = threadIdx.x & 63;
  const bf16x4* w1 = pack;
  const bf16x4* w2 = pack + RB_HID * KB_IN * 64;
  const bf16x4* w3 = w2 + RB_HID * KH * 64;
  const float* sa = aux;
  const double dt = st->dt;
  {
    const long cb = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long cell = cb * 32 + (lane & 31);
    const bool ok = cell < nwc;
    bf16x4 xb[KB_IN];
#pragma unroll
    for (int kb = 0; kb < KB_IN; kb++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int k = kb * 8 + 4 * (lane >> 5) + i;
        float x = 0.0f;
        if (ok && k < MLP_IN) {
          const double raw = (k < ND * V) ? wh[(long)k * nwc + cell] : dt;
          x = ((float)raw - sa[AUX_IN_MEAN + k]) * sa[AUX_IN_RSTD + k];
        }
        xb[kb][i] = to_bf16(x);
      }
    f32x16 h[RB_HID];
#pragma unroll
    for (int rb = 0; rb < RB_HID; rb++) {
#pragma unroll
      for (int j = 0; j < 16; j++) h[rb][j] = 0.0f;
#pragma unroll
      for (int kb = 0; kb < KB_IN; kb++)
        h[rb] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(w1[(rb * KB_IN + kb) * 64 + lane],
                                                         xb[kb], h[rb], 0, 0, 0);
    }
    // k-block kb of the next layer's B operand = rows 8 kb .. 8 kb + 7 =
    // accumulator block kb / 4, group kb % 4
    bf16x4 a[KH];
    auto activate = [&](const f32x16* hh, int boff) {
#pragma unroll
      for (int kb = 0; kb < KH; kb++) {
        const int rb = kb / 4, gi = kb % 4;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int row = rb * 32 + 8 * gi + 4 * (lane >> 5) + r;
          a[kb][r] = to_bf16(silu(hh[rb][4 * gi + r] + sa[boff + row]));
        }
      }
    };
    activate(h, AUX_B1);
#pragma unroll
    for (int rb = 0; rb < RB_HID; rb++) {
#pragma unroll
      for (int j = 0; j < 16; j++) h[rb][j] = 0.0f;
#pragma unroll
      for (int kb = 0; kb < KH; kb++)
        h[rb] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(w2[(rb * KH + kb) * 64 + lane], a[kb],
                                                         h[rb], 0, 0, 0);
    }
    activate(h, AUX_B2);
#pragma unroll 1
    for (int rb = 0; rb < RB_OUT; rb++) {
      f32x16 y;
#pragma unroll
      for (int j = 0; j < 16; j++) y[j] = 0.0f;
#pragma unroll
      for (int kb = 0; kb < KH; kb++)
        y = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(w3[(rb * KH + kb) * 64 + lane], a[kb], y, 0,
                                                     0, 0);
      if (ok) {
#pragma unroll
        for (int j = 0; j < 16; j++) {
          const int o = rb * 32 + 8 * (j / 4) + 4 * (lane >> 5) + (j % 4);
          if (o < MLP_OUT) {
            const float val = (y[j] + sa[AUX_B3 + o]) * sa[AUX_OUT_STD + o] + sa[AUX_OUT_MEAN + o];
            const int p = o % (ND * V);
            guess[(long)o * nwc + cell] = wh[(long)p * nwc + cell] + dt * (double)val;
          }
        }
      }
    }
  }
}
#endif

#ifndef PDE_BATCH
#define PDE_BATCH 0
#endif
#if PDE_BATCH
// ------------------------------------------------------- batched ensembles
// Every member of an ensemble (grids of one system, order and variable
// count, each with its own size, spacing, walls and state) advances in one
// launch per stage.  Member m owns the workgroups [map[m].start,
// map[m + 1].start) of a launch, laid out as its own single-grid launch of
// gx x gy x gz workgroups, and runs the single-grid kernel body on its own
// buffers, so each member computes exactly what it computes alone.  Both
// tables are read through the constant address space (nothing writes them
// during a launch), so their uniform loads go to scalar registers.
struct Member {
  Geom g;
  double *u, *wh, *tr, *vol, *trd;
  double *ff[2], *fb[2];
  StepState* st;
  unsigned int* hot;
};
struct BlkMap {
  unsigned int start, gx, gy, gz;
};
#define PDE_CONST_AS __attribute__((address_space(4)))

// The member owning workgroup blk: map[nm].start is the launch's total.
DEV int batch_member(const PDE_CONST_AS BlkMap* map, int nm, unsigned int blk) {
  int lo = 0, hi = nm;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (map[mid].start <= blk) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Resolves the launch's workgroup to (member, its coordinates in the
// member's own launch); declares M (the member) and b (the coordinates).
#define PYPDE_BATCH_MEMBER                                                                     \
  const PDE_CONST_AS BlkMap* cmap = (const PDE_CONST_AS BlkMap*)map;                           \
  const int m_ = batch_member(cmap, nm, blockIdx.x);                                           \
  const unsigned int loc_ = blockIdx.x - cmap[m_].start;                                       \
  const unsigned int gx_ = cmap[m_].gx, gy_ = cmap[m_].gy;                                     \
  const Blk b = {loc_ % gx_, (loc_ / gx_) % gy_, loc_ / (gx_ * gy_), gx_ * gy_ * cmap[m_].gz}; \
  const Member& M = *(const Member*)&((const PDE_CONST_AS Member*)ms)[m_];

#define PYPDE_WENO_BATCH(NAME, TX, TY)                                                         \
  extern "C" __global__ void __launch_bounds__(256)                                            \
  NAME(const Member* ms, const BlkMap* map, int nm) {                                          \
    PYPDE_BATCH_MEMBER                                                                         \
    if (M.st->done) return;                                                                    \
    if constexpr (D == 1)                                                                      \
      weno_1d(M.u, M.wh, M.g, 1, 0, M.g.nwc, b);                                               \
    else                                                                                       \
      weno_2d<TX, TY>(M.u, M.wh, M.g, 1, 0, 0, M.g.nwc, b);                                    \
  }
PYPDE_WENO_BATCH(k_weno_b, 8, 32)
PYPDE_WENO_BATCH(k_weno_4x64_b, 4, 64)
PYPDE_WENO_BATCH(k_weno_16x16_b, 16, 16)
PYPDE_WENO_BATCH(k_weno_16x32_b, 16, 32)

extern "C" __global__ void __launch_bounds__(256)
k_dt_b(const Member* ms, const BlkMap* map, int nm) {
  PYPDE_BATCH_MEMBER
  if (M.st->done) return;
  dt_body(M.wh, M.st, M.g, b);
}

extern "C" __global__ void __launch_bounds__(64)
k_predictor_static_b(const Member* ms, const BlkMap* map, int nm, double newton_tol,
                     int newton_maxit) {
  __shared__ PredLDS sm;
  PYPDE_BATCH_MEMBER
  predictor_static_body(sm, M.wh, nullptr, M.tr, M.vol, nullptr, M.st, M.g, newton_tol,
                        newton_maxit, M.trd, b);
}
extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_predictor_static2_b(const Member* ms, const BlkMap* map, int nm, double newton_tol,
                      int newton_maxit) {
  __shared__ PredLDS sm;
  PYPDE_BATCH_MEMBER
  predictor_static_body(sm, M.wh, nullptr, M.tr, M.vol, nullptr, M.st, M.g, newton_tol,
                        newton_maxit, M.trd, b);
}
#if PDE_STIFF && !PDE_GUESS
extern "C" __global__ void __launch_bounds__(64)
k_predictor_cold_b(const Member* ms, const BlkMap* map, int nm, double newton_tol,
                   int newton_maxit) {
  __shared__ PredLDS sm;
  PYPDE_BATCH_MEMBER
  predictor_static_body<1>(sm, M.wh, nullptr, M.tr, M.vol, nullptr, M.st, M.g, newton_tol,
                           newton_maxit, nullptr, b, M.hot);
}
extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_predictor_cold3_b(const Member* ms, const BlkMap* map, int nm, double newton_tol,
                    int newton_maxit) {
  __shared__ PredLDS sm;
  PYPDE_BATCH_MEMBER
  predictor_static_body<1>(sm, M.wh, nullptr, M.tr, M.vol, nullptr, M.st, M.g, newton_tol,
                           newton_maxit, nullptr, b, M.hot);
}
extern "C" __global__ void __launch_bounds__(64)
k_predictor_hot_b(const Member* ms, const BlkMap* map, int nm, double newton_tol,
                  int newton_maxit) {
  __shared__ PredLDS sm;
  PYPDE_BATCH_MEMBER
  predictor_static_body<2>(sm, M.wh, nullptr, M.tr, M.vol, nullptr, M.st, M.g, newton_tol,
                           newton_maxit, nullptr, b, M.hot);
}

#endif

#define PYPDE_FLUX_BATCH(NAME, DIR, BS)                                                        \
  extern "C" __global__ void __launch_bounds__(BS)                                             \
  NAME(const Member* ms, const BlkMap* map, int nm) {                                          \
    PYPDE_BATCH_MEMBER                                                                         \
    if (M.st->done) return;                                                                    \
    flux_dir<DIR, BS>(M.tr, M.ff[DIR], M.fb[DIR], M.st, M.g, M.trd, b);                        \
  }
PYPDE_FLUX_BATCH(k_flux0_b, 0, 256)
PYPDE_FLUX_BATCH(k_flux0_128_b, 0, 128)
PYPDE_FLUX_BATCH(k_flux0_64_b, 0, 64)
#if PDE_D >= 2
PYPDE_FLUX_BATCH(k_flux1_b, 1, 256)
PYPDE_FLUX_BATCH(k_flux1_128_b, 1, 128)
PYPDE_FLUX_BATCH(k_flux1_64_b, 1, 64)
#endif

extern "C" __global__ void __launch_bounds__(256)
k_update_b(const Member* ms, const BlkMap* map, int nm) {
  PYPDE_BATCH_MEMBER
  if (M.st->done) return;
  const FaceBufs fbufs = {{M.ff[0], M.ff[1]}, {M.fb[0], M.fb[1]}};
  update_cell(M.u, M.vol, fbufs, nullptr, M.st, M.g, b);
  if (last_workgroup(&M.st->step_ticket, b.n)) advance_step(M.st, M.g.ndt);
}
#define PYPDE_UPDATE_ROWS_BATCH(R_)                                                            \
  extern "C" __global__ void __launch_bounds__(256)                                           \
  k_update_r##R_##_b(const Member* ms, const BlkMap* map, int nm) {                            \
    PYPDE_BATCH_MEMBER                                                                         \
    if (M.st->done) return;                                                                    \
    const FaceBufs fbufs = {{M.ff[0], M.ff[1]}, {M.fb[0], M.fb[1]}};                           \
    update_rows<R_>(M.u, M.vol, fbufs, nullptr, M.st, M.g, b);                                 \
    if (last_workgroup(&M.st->step_ticket, b.n)) advance_step(M.st, M.g.ndt);                  \
  }
PYPDE_UPDATE_ROWS_BATCH(1)
PYPDE_UPDATE_ROWS_BATCH(2)
PYPDE_UPDATE_ROWS_BATCH(4)
PYPDE_UPDATE_ROWS_BATCH(8)
#endif
