// Host-compiled check of gpe::fast_sincos against long-double sin/cos.
// Prints the largest error in units of the field precision's epsilon over
// a sweep of arguments: small phases, moderate ones, and the reduction's
// boundaries.
#include <cmath>
#include <cstdio>
#include <random>
#include "../../generalizedgrosspitaevskii.jl_amd/csrc/gpe_math.hpp"

template <typename R>
static double sweep(double lim, double eps, long n, unsigned seed) {
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  double worst = 0;
  for (long i = 0; i < n; ++i) {
    double xd = (i % 3 == 0) ? u(g) * 1e-3 : (i % 3 == 1 ? u(g) * 10.0 : u(g) * lim);
    R x = (R)xd;
    R s, c;
    gpe::fast_sincos(x, s, c);
    long double xl = (long double)x;
    long double es = std::fabs((long double)s - sinl(xl)), ec = std::fabs((long double)c - cosl(xl));
    double e = (double)std::fmax(es, ec) / eps;  // absolute error in eps (|sin|,|cos| <= 1)
    if (e > worst) worst = e;
  }
  return worst;
}

// fast_sincos_n must equal the scalar routine value for value, with narrow,
// wide and non-finite arguments mixed in one call.
template <typename R>
static long vector_mismatches(unsigned seed) {
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  const R wide[] = {(R)3e9, (R)-7e12, (R)INFINITY, (R)-INFINITY, (R)NAN, (R)2e5};
  long bad = 0;
  for (int it = 0; it < 200000; ++it) {
    R x[8], s[8], c[8];
    for (int m = 0; m < 8; ++m) x[m] = (R)(u(g) * (m % 2 ? 1e3 : 3.0));
    if (it % 5 == 0) x[it % 8] = wide[(it / 5) % 6];
    gpe::fast_sincos_n<R, 8>(x, s, c);
    for (int m = 0; m < 8; ++m) {
      R s1, c1;
      gpe::fast_sincos(x[m], s1, c1);
      const bool same = (s1 == s[m] || (s1 != s1 && s[m] != s[m])) && (c1 == c[m] || (c1 != c1 && c[m] != c[m]));
      bad += !same;
    }
  }
  return bad;
}

int main() {
  double w64 = sweep<double>(1e6, 2.220446049250313e-16, 2000000, 1);
  double w32 = sweep<float>(1e4, 1.1920928955078125e-07, 2000000, 2);
  double s64, c64;
  gpe::fast_sincos(1e12, s64, c64);  // wide path
  printf("fp64_max_abs_err_eps %.3f\nfp32_max_abs_err_eps %.3f\nwide %.17g %.17g\n", w64, w32, s64 - std::sin(1e12),
         c64 - std::cos(1e12));
  printf("vector_mismatch_fp64 %ld\nvector_mismatch_fp32 %ld\n", vector_mismatches<double>(3), vector_mismatches<float>(4));
  return 0;
}
