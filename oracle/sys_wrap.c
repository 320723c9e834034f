/* Gives the functions of one PDE system source the reference's C callback ABI
 * (pypde/cfuncs.py:6-8: F(out,Q,dQ,d), B(out,Q,d), S(out,Q)); test-only. */
#include "pde_prelude.h"
#include PDE_SYSTEM_FILE

#ifdef PDE_HAS_F
void sys_F(double *out, double *Q, double *dQ, int d) { F(out, Q, dQ, d); }
#endif
#ifdef PDE_HAS_B
void sys_B(double *out, double *Q, int d) { B(out, Q, d); }
#endif
#ifdef PDE_HAS_S
void sys_S(double *out, double *Q) { S(out, Q); }
#endif
#ifdef PDE_HAS_M
void sys_M(double *out, double *Q) { M(out, Q); }
#endif
double sys_exp(double x) { return pde_exp(x); }
double sys_cbrt(double x) { return pde_cbrt(x); }
