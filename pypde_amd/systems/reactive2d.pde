/* 2D reactive Euler with single-step Arrhenius kinetics.  Conserved
 * (rho, rho*u, rho*v, E, rho*lambda), lambda = reactant mass fraction,
 * p = (g-1)(E - rho|v|^2/2 - Qc*rho*lambda), rate = K exp(-Ea rho / p).
 * The reference's reactive example (pypde/tests/reactive_euler/system.py:13-61)
 * uses discrete-ignition kinetics in 1D; this system keeps its structure
 * (F advects every variable, S acts on rho*lambda only) with Arrhenius rates.
 * M is the analytic Jacobian dS/dQ used by the stiff Newton predictor.
 * F and M form 1 / rho once per state and multiply it in; S needs none. */
#define PDE_HAS_F 1
#define PDE_HAS_S 1
#define PDE_HAS_M 1
#ifndef PDE_GAMMA
#define PDE_GAMMA 1.2
#endif
#ifndef PDE_QC
#define PDE_QC 25.0
#endif
#ifndef PDE_EA
#define PDE_EA 20.0
#endif
#ifndef PDE_K
#define PDE_K 1.0e4
#endif
/* pressure, given 1 / rho (one division per state) */
PDE_FN double reactive_pressure_ri(const double *Q, double ri) {
  double ru = Q[1], rv = Q[2], E = Q[3], rl = Q[4];
  double ke = 0.5 * ((ru * ru + rv * rv) * ri);
  return (PDE_GAMMA - 1.0) * (E - ke - PDE_QC * rl);
}
PDE_FN void F(double *out, const double *Q, const double *dQ, int d) {
  double ru = Q[1], rv = Q[2], E = Q[3], rl = Q[4];
  double ri = 1.0 / Q[0];
  double p = reactive_pressure_ri(Q, ri);
  double vn = ((d == 0) ? ru : rv) * ri;
  out[0] = (d == 0) ? ru : rv;
  out[1] = ru * vn + ((d == 0) ? p : 0.0);
  out[2] = rv * vn + ((d == 1) ? p : 0.0);
  out[3] = (E + p) * vn;
  out[4] = rl * vn;
}
/* S from X = rho p / (g-1) = rho (E - Qc rho lambda) - |rho v|^2 / 2, which
 * needs no 1 / rho: p > 0 is X > 0 (rho > 0), and the Arrhenius exponent
 * -Ea rho / p = -Ea rho^2 / ((g-1) X) takes one division. */
PDE_FN void S(double *out, const double *Q) {
  double r = Q[0], ru = Q[1], rv = Q[2], E = Q[3], rl = Q[4];
  double X = r * (E - PDE_QC * rl) - 0.5 * (ru * ru + rv * rv);
  out[0] = 0.0;
  out[1] = 0.0;
  out[2] = 0.0;
  out[3] = 0.0;
  out[4] = (X > 0.0 && r > 0.0)
               ? -PDE_K * rl * pde_exp(-PDE_EA * (r * r) / ((PDE_GAMMA - 1.0) * X))
               : 0.0;
}
PDE_FN void M(double *out, const double *Q) {
  for (int i = 0; i < 25; i++) out[i] = 0.0;
  double r = Q[0], ri = 1.0 / r, rl = Q[4];
  double p = reactive_pressure_ri(Q, ri);
  if (!(p > 0.0)) return;
  double u = Q[1] * ri, v = Q[2] * ri;
  double pi = 1.0 / p;
  double e = pde_exp(-PDE_EA * r * pi);
  double a = -PDE_K * rl * e;       /* S4 = a, dS4/dq = -K e dq(rl) + a * dq(-Ea r/p) */
  double c = PDE_EA * r * (pi * pi); /* d(-Ea r/p)/dq_j = -Ea/p [j==0] + c dp/dq_j */
  double gm = PDE_GAMMA - 1.0;
  out[20] = a * (c * (gm * 0.5 * (u * u + v * v)) - PDE_EA * pi);
  out[21] = a * (c * (-gm * u));
  out[22] = a * (c * (-gm * v));
  out[23] = a * (c * gm);
  out[24] = -PDE_K * e + a * (c * (-gm * PDE_QC));
}
