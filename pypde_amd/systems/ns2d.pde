/* 2D compressible Navier-Stokes, conserved (rho, rho u, rho v, E), ideal gas
 * (gamma 1.4, R = 1), constant viscosity mu and Prandtl number 0.72:
 * F_d = Euler flux - (0, tau_d1, tau_d2, u tau_d1 + v tau_d2 + kappa dT/dx_d),
 * velocity and temperature gradients from the conserved gradients dQ[d V + i]
 * (the form of the reference's Navier-Stokes example, pypde/tests/navier_stokes). */
#define PDE_HAS_F 1
#define PDE_SECOND_ORDER 1
#ifndef PDE_GAMMA
#define PDE_GAMMA 1.4
#endif
#ifndef PDE_MU
#define PDE_MU 0.01
#endif
#ifndef PDE_PR
#define PDE_PR 0.72
#endif
PDE_FN void F(double *out, const double *Q, const double *dQ, int d) {
  double r = Q[0], ru = Q[1], rv = Q[2], E = Q[3];
  double u = ru / r, v = rv / r;
  double p = (PDE_GAMMA - 1.0) * (E - 0.5 * (ru * u + rv * v));
  double T = p / r;
  double ux[2], vx[2], Tx[2];
  for (int k = 0; k < 2; k++) {
    const double *g = dQ + 4 * k;
    ux[k] = (g[1] - u * g[0]) / r;
    vx[k] = (g[2] - v * g[0]) / r;
    double ke = u * g[1] + v * g[2] - 0.5 * (u * u + v * v) * g[0];
    double px = (PDE_GAMMA - 1.0) * (g[3] - ke);
    Tx[k] = (px - T * g[0]) / r;
  }
  double div = ux[0] + vx[1];
  double t11 = PDE_MU * (2.0 * ux[0] - (2.0 / 3.0) * div);
  double t22 = PDE_MU * (2.0 * vx[1] - (2.0 / 3.0) * div);
  double t12 = PDE_MU * (ux[1] + vx[0]);
  double kappa = PDE_MU * PDE_GAMMA / ((PDE_GAMMA - 1.0) * PDE_PR);
  double vn = (d == 0) ? u : v;
  double td1 = (d == 0) ? t11 : t12, td2 = (d == 0) ? t12 : t22;
  out[0] = r * vn;
  out[1] = ru * vn + ((d == 0) ? p : 0.0) - td1;
  out[2] = rv * vn + ((d == 1) ? p : 0.0) - td2;
  out[3] = (E + p) * vn - (u * td1 + v * td2 + kappa * Tx[d]);
}
