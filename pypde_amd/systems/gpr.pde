/* Godunov-Peshkov-Romenski (GPR) continuum model, 17 variables, the
 * reference's GPR example (pypde/tests/gpr/system.py:8-75 with
 * misc/state.py, misc/eos.py, misc/mg.py, misc/sources.py, params.py):
 * conserved Q = (rho, rho E, rho v[3], A[3][3] row-major, rho J[3]);
 * stiffened-gas Mie-Gruneisen EOS with p_inf = 0, gamma = 1.4, c_v = 2.5,
 * shear and heat wave speeds c_s = c_alpha = 5, mu = 2e-2, Pr = 0.75.
 *
 *   E = E1(rho, p) + c_s^2/4 |dev(A^T A)|^2 + c_a^2/2 |J|^2 + |v|^2/2
 *   sigma = -rho A^T psi, psi = dE/dA = c_s^2 A dev(A^T A), T = p / rho
 *   F_d:  rho v_d,  rho E v_d + p v_d - sigma_d . v + q_d,
 *         rho v_d v + p e_d - sigma_d,  (A v)_i at rows 5 + 3i + d,
 *         rho v_d J + T e_d,   q = c_a^2 J T
 *   S:    A rows: -psi 3 det(A)^(5/3) / (c_s^2 tau1),
 *         J rows: -rho c_a^2 J / (c_a^2 tau2 (rho/rho0)(T0/T))
 *   B_d:  v_d on the A diagonal, minus v along rows 5+d, 8+d, 11+d as in the
 *         reference (system.py:61-75).
 * det(A)^(5/3) = det * cbrt(det)^2 with the prelude's pde_cbrt (the reference
 * uses det**(5/3); they agree to a few ulps). */
#define PDE_HAS_F 1
#define PDE_HAS_B 1
#define PDE_HAS_S 1
#define GPR_GAMMA 1.4
#define GPR_CV 2.5
#define GPR_CS2 25.0
#define GPR_CA2 25.0
#define GPR_MU 2.0e-2
#define GPR_PR 0.75
#define GPR_RHO0 1.0
#define GPR_P0 (1.0 / GPR_GAMMA)
#define GPR_KAPPA (GPR_MU * GPR_GAMMA * GPR_CV / GPR_PR)
#define GPR_T0 (GPR_P0 / (GPR_RHO0 * (GPR_GAMMA - 1.0) * GPR_CV))
#define GPR_TAU1 (6.0 * GPR_MU / (GPR_RHO0 * GPR_CS2))
#define GPR_TAU2 (GPR_KAPPA * GPR_RHO0 / (GPR_T0 * GPR_CA2))

/* dev(G) of the Gram matrix G = A^T A */
PDE_FN void gpr_devgram(const double *A, double *dG) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      dG[i * 3 + j] = A[0 * 3 + i] * A[0 * 3 + j] + A[1 * 3 + i] * A[1 * 3 + j] +
                      A[2 * 3 + i] * A[2 * 3 + j];
  const double tr3 = (dG[0] + dG[4] + dG[8]) / 3.0;
  dG[0] = dG[0] - tr3;
  dG[4] = dG[4] - tr3;
  dG[8] = dG[8] - tr3;
}

/* psi = c_s^2 A dev(G) */
PDE_FN void gpr_psi(const double *A, const double *dG, double *psi) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      psi[i * 3 + j] =
          GPR_CS2 * (A[i * 3 + 0] * dG[0 * 3 + j] + A[i * 3 + 1] * dG[1 * 3 + j] +
                     A[i * 3 + 2] * dG[2 * 3 + j]);
}

/* pressure from the conserved state: p = (E - |v|^2/2 - E_2A - E_2J) rho (gamma - 1) */
PDE_FN double gpr_pressure(const double *Q, const double *v, const double *J,
                           const double *dG) {
  const double r = Q[0];
  const double E = Q[1] / r;
  const double e3 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / 2.0;
  double s = 0.0;
  for (int i = 0; i < 9; i++) s = s + dG[i] * dG[i];
  const double e2a = GPR_CS2 / 4.0 * s;
  const double e2j = GPR_CA2 / 2.0 * (J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
  const double e1 = E - e3 - e2a - e2j;
  return e1 * r * (GPR_GAMMA - 1.0);
}

PDE_FN void F(double *out, const double *Q, const double *dQ, int d) {
  (void)dQ;
  const double r = Q[0];
  const double E = Q[1] / r;
  double v[3], J[3], dG[9], psi[9], sig[9];
  for (int i = 0; i < 3; i++) {
    v[i] = Q[2 + i] / r;
    J[i] = Q[14 + i] / r;
  }
  const double *A = Q + 5;
  gpr_devgram(A, dG);
  gpr_psi(A, dG, psi);
  const double p = gpr_pressure(Q, v, J, dG);
  const double T = p / (r * (GPR_GAMMA - 1.0) * GPR_CV);
  /* sigma = -rho A^T psi */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      sig[i * 3 + j] = -r * (A[0 * 3 + i] * psi[0 * 3 + j] + A[1 * 3 + i] * psi[1 * 3 + j] +
                             A[2 * 3 + i] * psi[2 * 3 + j]);
  for (int i = 0; i < 17; i++) out[i] = 0.0;
  const double vd = v[d];
  const double rvd = r * vd;
  out[0] = rvd;
  out[1] = rvd * E + p * vd;
  for (int i = 0; i < 3; i++) out[2 + i] = rvd * v[i];
  out[2 + d] = out[2 + d] + p;
  const double *sd = sig + 3 * d;
  out[1] = out[1] - (sd[0] * v[0] + sd[1] * v[1] + sd[2] * v[2]);
  for (int i = 0; i < 3; i++) out[2 + i] = out[2 + i] - sd[i];
  for (int i = 0; i < 3; i++)
    out[5 + 3 * i + d] = A[i * 3 + 0] * v[0] + A[i * 3 + 1] * v[1] + A[i * 3 + 2] * v[2];
  out[1] = out[1] + GPR_CA2 * J[d] * T;
  for (int i = 0; i < 3; i++) out[14 + i] = rvd * J[i];
  out[14 + d] = out[14 + d] + T;
}

PDE_FN void B(double *out, const double *Q, int d) {
  const double r = Q[0];
  double v[3];
  for (int i = 0; i < 3; i++) v[i] = Q[2 + i] / r;
  for (int i = 0; i < 17 * 17; i++) out[i] = 0.0;
  for (int i = 5; i < 14; i++) out[i * 17 + i] = v[d];
  for (int b = 0; b < 3; b++)
    for (int j = 0; j < 3; j++) {
      const int row = 5 + 3 * b + d, col = 5 + 3 * b + d + j;
      out[row * 17 + col] = out[row * 17 + col] - v[j];
    }
}

PDE_FN void S(double *out, const double *Q) {
  const double r = Q[0];
  double v[3], J[3], dG[9], psi[9];
  for (int i = 0; i < 3; i++) {
    v[i] = Q[2 + i] / r;
    J[i] = Q[14 + i] / r;
  }
  const double *A = Q + 5;
  gpr_devgram(A, dG);
  gpr_psi(A, dG, psi);
  const double p = gpr_pressure(Q, v, J, dG);
  const double T = p / (r * (GPR_GAMMA - 1.0) * GPR_CV);
  const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                     A[2] * (A[3] * A[7] - A[4] * A[6]);
  const double cb = pde_cbrt(det);
  const double th1 = 3.0 * (det * (cb * cb)) / (GPR_CS2 * GPR_TAU1);
  const double th2 = 1.0 / (GPR_CA2 * GPR_TAU2 * (r / GPR_RHO0) * (GPR_T0 / T));
  for (int i = 0; i < 5; i++) out[i] = 0.0;
  for (int i = 0; i < 9; i++) out[5 + i] = -psi[i] * th1;
  for (int i = 0; i < 3; i++) out[14 + i] = -r * (GPR_CA2 * J[i]) * th2;
}
