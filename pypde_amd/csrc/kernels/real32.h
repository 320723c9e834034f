// FP32 phase of the mixed-precision predictor: the system source compiled a
// second time, in namespace f32, with `double` standing for R, a float that
// converts from double constants.  Every operation of the physics then runs
// in float: R op R, with double literals and parameters rounded to float
// where they meet an R.  pde_exp / pde_cbrt map to the float library
// routines (about 1 ulp of float).
namespace f32 {
struct R {
  float v;
  DEV constexpr R() : v(0.0f) {}
  DEV constexpr R(float x) : v(x) {}
  DEV constexpr R(double x) : v((float)x) {}
  DEV constexpr R(int x) : v((float)x) {}
  DEV R& operator+=(R b) { v = v + b.v; return *this; }
  DEV R& operator-=(R b) { v = v - b.v; return *this; }
  DEV R& operator*=(R b) { v = v * b.v; return *this; }
  DEV R& operator/=(R b) { v = v / b.v; return *this; }
};
DEV R operator+(R a) { return a; }
DEV R operator-(R a) { return R(-a.v); }
DEV R operator+(R a, R b) { return R(a.v + b.v); }
DEV R operator-(R a, R b) { return R(a.v - b.v); }
DEV R operator*(R a, R b) { return R(a.v * b.v); }
DEV R operator/(R a, R b) { return R(a.v / b.v); }
DEV bool operator<(R a, R b) { return a.v < b.v; }
DEV bool operator>(R a, R b) { return a.v > b.v; }
DEV bool operator<=(R a, R b) { return a.v <= b.v; }
DEV bool operator>=(R a, R b) { return a.v >= b.v; }
DEV bool operator==(R a, R b) { return a.v == b.v; }
DEV bool operator!=(R a, R b) { return a.v != b.v; }
DEV R sqrt(R a) { return R(__builtin_sqrtf(a.v)); }
DEV R fabs(R a) { return R(__builtin_fabsf(a.v)); }
DEV R fmax(R a, R b) { return R(__builtin_fmaxf(a.v, b.v)); }
DEV R fmin(R a, R b) { return R(__builtin_fminf(a.v, b.v)); }
DEV R exp(R a) { return R(expf(a.v)); }
DEV R log(R a) { return R(logf(a.v)); }
DEV R cbrt(R a) { return R(cbrtf(a.v)); }
DEV R pow(R a, R b) { return R(powf(a.v, b.v)); }
DEV R pde_exp(R a) { return R(expf(a.v)); }
DEV R pde_cbrt(R a) { return R(cbrtf(a.v)); }
}  // namespace f32
