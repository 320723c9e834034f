/* Linear advection of one scalar with velocity (1, 0.5, 0.25) along axes 0..2.
 * Smooth-data checks: the predictor of a polynomial is exact, and the scheme
 * converges at the reference's order (docs/pages/background.rst:4-16). */
#define PDE_HAS_F 1
PDE_FN void F(double *out, const double *Q, const double *dQ, int d) {
  out[0] = ((d == 0) ? 1.0 : ((d == 1) ? 0.5 : 0.25)) * Q[0];
}
