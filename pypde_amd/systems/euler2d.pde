/* 2D Euler equations, conserved (rho, rho*u, rho*v, E), ideal gas; the flux of
 * the reference's C++ smoke driver src/test.cpp:4-29 written in conserved form. */
#define PDE_HAS_F 1
#ifndef PDE_GAMMA
#define PDE_GAMMA 1.4
#endif
PDE_FN void F(double *out, const double *Q, const double *dQ, int d) {
  const double g = PDE_GAMMA;
  double r = Q[0], ru = Q[1], rv = Q[2], E = Q[3];
  double u = ru / r, v = rv / r;
  double p = (g - 1.0) * (E - 0.5 * (ru * u + rv * v));
  double vn = (d == 0) ? u : v;
  out[0] = (d == 0) ? ru : rv;
  out[1] = ru * vn + ((d == 0) ? p : 0.0);
  out[2] = rv * vn + ((d == 1) ? p : 0.0);
  out[3] = (E + p) * vn;
}
