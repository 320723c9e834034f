/* 1D Euler equations, conserved (rho, rho*u, E), ideal gas.  The physics of the
 * reference example pypde/tests/euler/system.py:10-21 (which orders the variables
 * (rho, rho*E, rho*v)); here momentum precedes energy. */
#define PDE_HAS_F 1
#ifndef PDE_GAMMA
#define PDE_GAMMA 1.4
#endif
PDE_FN void F(double *out, const double *Q, const double *dQ, int d) {
  const double g = PDE_GAMMA;
  double r = Q[0], ru = Q[1], E = Q[2];
  double u = ru / r;
  double p = (g - 1.0) * (E - 0.5 * (ru * u));
  out[0] = ru;
  out[1] = ru * u + p;
  out[2] = (E + p) * u;
}
