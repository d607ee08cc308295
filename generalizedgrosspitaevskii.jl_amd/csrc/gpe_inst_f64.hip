// FP64 (ComplexF64 field) instantiations of the fused kernels.  Two FP64
// components at N = 4096 do not fit the LDS and are rejected by gpe_create.
#include "gpe_inst.hpp"

#define GPE_XCASE_f64(n)                                        \
  case n * 8 + 1: return gpe::launch_x<double, n, 1>(a, nlines, w, s, d); \
  case n * 8 + 2: if (n > 2048) return hipErrorInvalidValue; return gpe::launch_x<double, (n > 2048 ? 2048 : n), 2>(a, nlines, w, s, d);
#define GPE_YCASE_f64(n)                                   \
  case n * 8 + 1: return gpe::launch_y<double, n, 1>(a, c, s, d); \
  case n * 8 + 2: if (n > 2048) return hipErrorInvalidValue; return gpe::launch_y<double, (n > 2048 ? 2048 : n), 2>(a, c, s, d);
#define GPE_1CASE_f64(n)                                    \
  case n * 8 + 1: return gpe::launch_1d<double, n, 1>(a, s); \
  case n * 8 + 2: if (n > 2048) return hipErrorInvalidValue; return gpe::launch_1d<double, (n > 2048 ? 2048 : n), 2>(a, s);

#define GPE_XCASE34_f64(n)                                       \
  case n * 8 + 3: return gpe::launch_x<double, n, 3>(a, nlines, w, s, d); \
  case n * 8 + 4: return gpe::launch_x<double, n, 4>(a, nlines, w, s, d);
#define GPE_YCASE34_f64(n)                                  \
  case n * 8 + 3: return gpe::launch_y<double, n, 3>(a, c, s, d); \
  case n * 8 + 4: return gpe::launch_y<double, n, 4>(a, c, s, d);
#define GPE_1CASE34_f64(n)                                   \
  case n * 8 + 3: return gpe::launch_1d<double, n, 3>(a, s); \
  case n * 8 + 4: return gpe::launch_1d<double, n, 4>(a, s);

#define GPE_KCASE_f64(n) \
  case n: return gpe::launch_kabs2<double, n>(a, s);

#define GPE_WCASE_f64(n) \
  case n: return gpe::launch_kwin<double, n>(a, s);

#define GPE_RCASE_f64(n) \
  case n: return gpe::launch_rowcorr<double, n>(a, nblk, s);

#define GPE_SCASE_f64(n) \
  case n: return gpe::launch_sweep<double, n>(a, s);

GPE_DEFINE_RUNNERS(double, f64)
