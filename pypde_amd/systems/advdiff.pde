/* Linear advection-diffusion, one variable, any ndim:
 * dQ/dt + div(a Q - nu grad Q) = 0 with a = (1, 0.5), nu = 0.02.  F depends on
 * the gradient (the reference's three-argument F, cfuncs.py:11-27), so the
 * solver runs its viscous (second-order) path. */
#define PDE_HAS_F 1
#define PDE_SECOND_ORDER 1
#ifndef PDE_NU
#define PDE_NU 0.02
#endif
PDE_FN void F(double *out, const double *Q, const double *dQ, int d) {
  double a = (d == 0) ? 1.0 : 0.5;
  out[0] = a * Q[0] - PDE_NU * dQ[d];
}
