/* Two coupled scalars advected in non-conservative form, Q_t + B_d Q_x_d = S(Q),
 * B_d = [[a_d, 0.1], [0.1, a_d]] (real eigenvalues a_d +- 0.1) and linear decay
 * S = -0.5 Q.  Exercises the B path (rhs B.dQ, face path integral, volume
 * term) of the reference: src/solvers/dg/dg.cpp:108-111, fv/fv.cpp:231-235,
 * fv/fluxes.cpp:103-116. */
#define PDE_HAS_B 1
#define PDE_HAS_S 1
PDE_FN void B(double *out, const double *Q, int d) {
  double a = (d == 0) ? 1.0 : -0.5;
  out[0] = a;
  out[1] = 0.1;
  out[2] = 0.1;
  out[3] = a;
}
PDE_FN void S(double *out, const double *Q) {
  out[0] = -0.5 * Q[0];
  out[1] = -0.5 * Q[1];
}
