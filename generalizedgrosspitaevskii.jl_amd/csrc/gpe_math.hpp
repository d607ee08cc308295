// Compact sine/cosine for the real-space half-step.
//
// Method: Cody-Waite reduction x = k pi/2 + r with k = rint(2x/pi) and a
// three-term split of pi/2 applied with fused multiply-adds, then the Taylor
// series of sin and cos on |r| <= pi/4 (to r^17 / r^18 in FP64, r^11 / r^12
// in FP32: truncation below 1e-19 and 2e-9 relative), and the quadrant
// rotation by k mod 4.  Arguments outside the range where the split is exact
// (|x| >= 1e9 in FP64, 1e5 in FP32) and non-finite ones take the library
// routine out of line.
//
// fast_sincos_n evaluates E arguments at once: the polynomial path runs for
// all of them without branches (so the E evaluations interleave), and one
// fix-up region, entered only when some argument is wide, replaces those
// results.  A per-value branch would split the code into E regions that the
// compiler cannot interleave, leaving each value's ~20 dependent operations
// exposed.
#pragma once
#include <hip/hip_runtime.h>

namespace gpe {

// The out-of-line routine returns the pair by value: results passed through
// pointers would live in scratch memory and be reloaded on every call site's
// fast path too.
template <typename R>
struct SinCos {
  R s, c;
};
__host__ __device__ __noinline__ inline SinCos<double> sincos_wide(double x) { return {sin(x), cos(x)}; }
__host__ __device__ __noinline__ inline SinCos<float> sincos_wide(float x) { return {sinf(x), cosf(x)}; }

__host__ __device__ __forceinline__ bool sincos_is_wide(double x) { return !(fabs(x) < 1e9); }
__host__ __device__ __forceinline__ bool sincos_is_wide(float x) { return !(fabsf(x) < 1e5f); }

// Rotation by the quadrant q = k mod 4, formed in floating point (exact for
// the integers k of the narrow range; any value for wide arguments, whose
// results are replaced).
__host__ __device__ __forceinline__ void quadrant(double k, double sr, double cr, double& s, double& c) {
  const double q = fma(-4.0, floor(k * 0.25), k);
  s = (q == 0.0) ? sr : (q == 1.0) ? cr : (q == 2.0) ? -sr : -cr;
  c = (q == 0.0) ? cr : (q == 1.0) ? -sr : (q == 2.0) ? -cr : sr;
}
__host__ __device__ __forceinline__ void quadrant(float k, float sr, float cr, float& s, float& c) {
  const float q = fmaf(-4.0f, floorf(k * 0.25f), k);
  s = (q == 0.0f) ? sr : (q == 1.0f) ? cr : (q == 2.0f) ? -sr : -cr;
  c = (q == 0.0f) ? cr : (q == 1.0f) ? -sr : (q == 2.0f) ? -cr : sr;
}

// The narrow-range evaluation, branch-free.
__host__ __device__ __forceinline__ void sincos_poly(double x, double& s, double& c) {
  const double k = rint(x * 0.6366197723675814);
  double r = fma(-k, 1.5707963267948966, x);
  r = fma(-k, 6.123233995736766e-17, r);
  r = fma(-k, -1.4973849048591698e-33, r);
  const double r2 = r * r;
  double p = 2.8114572543455206e-15;        //  1/17!
  p = fma(p, r2, -7.647163731819816e-13);   // -1/15!
  p = fma(p, r2, 1.6059043836821613e-10);   //  1/13!
  p = fma(p, r2, -2.505210838544172e-08);   // -1/11!
  p = fma(p, r2, 2.7557319223985893e-06);   //  1/9!
  p = fma(p, r2, -0.0001984126984126984);   // -1/7!
  p = fma(p, r2, 0.008333333333333333);     //  1/5!
  p = fma(p, r2, -0.16666666666666666);     // -1/3!
  const double sr = fma(p * r2, r, r);
  double q = -1.5619206968586225e-16;       // -1/18!
  q = fma(q, r2, 4.779477332387385e-14);    //  1/16!
  q = fma(q, r2, -1.1470745597729725e-11);  // -1/14!
  q = fma(q, r2, 2.08767569878681e-09);     //  1/12!
  q = fma(q, r2, -2.755731922398589e-07);   // -1/10!
  q = fma(q, r2, 2.48015873015873e-05);     //  1/8!
  q = fma(q, r2, -0.001388888888888889);    // -1/6!
  q = fma(q, r2, 0.041666666666666664);     //  1/4!
  q = fma(q, r2, -0.5);                     // -1/2!
  const double cr = fma(q, r2, 1.0);
  quadrant(k, sr, cr, s, c);
}

__host__ __device__ __forceinline__ void sincos_poly(float x, float& s, float& c) {
  const float k = rintf(x * 0.63661975f);
  float r = fmaf(-k, 1.5707963705062866f, x);
  r = fmaf(-k, -4.371138828673793e-08f, r);
  r = fmaf(-k, -1.7151245100058819e-15f, r);
  const float r2 = r * r;
  float p = -2.5052107943679403e-08f;       // -1/11!
  p = fmaf(p, r2, 2.7557318844628753e-06f); //  1/9!
  p = fmaf(p, r2, -0.00019841270113829523f);// -1/7!
  p = fmaf(p, r2, 0.008333333767950535f);   //  1/5!
  p = fmaf(p, r2, -0.1666666716337204f);    // -1/3!
  const float sr = fmaf(p * r2, r, r);
  float q = 2.0876755879584152e-09f;        //  1/12!
  q = fmaf(q, r2, -2.755731998149713e-07f); // -1/10!
  q = fmaf(q, r2, 2.4801587642286904e-05f); //  1/8!
  q = fmaf(q, r2, -0.0013888889225199819f); // -1/6!
  q = fmaf(q, r2, 0.0416666679084301f);     //  1/4!
  q = fmaf(q, r2, -0.5f);                   // -1/2!
  const float cr = fmaf(q, r2, 1.0f);
  quadrant(k, sr, cr, s, c);
}

template <typename R>
__host__ __device__ __forceinline__ void fast_sincos(R x, R& s, R& c) {
  sincos_poly(x, s, c);
  if (sincos_is_wide(x)) {
    const SinCos<R> w = sincos_wide(x);
    s = w.s;
    c = w.c;
  }
}

template <typename R, int E>
__host__ __device__ __forceinline__ void fast_sincos_n(const R (&x)[E], R (&s)[E], R (&c)[E]) {
  bool wide = false;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    sincos_poly(x[m], s[m], c[m]);
    wide = wide || sincos_is_wide(x[m]);
  }
  if (wide) {
#pragma unroll
    for (int m = 0; m < E; ++m)
      if (sincos_is_wide(x[m])) {
        const SinCos<R> w = sincos_wide(x[m]);
        s[m] = w.s;
        c[m] = w.c;
      }
  }
}

}  // namespace gpe
