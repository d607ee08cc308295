// FP32 (ComplexF32 field) instantiations of the fused kernels.
#include "gpe_inst.hpp"

#define GPE_XCASE_f32(n)                                       \
  case n * 8 + 1: return gpe::launch_x<float, n, 1>(a, nlines, w, s, d); \
  case n * 8 + 2: return gpe::launch_x<float, n, 2>(a, nlines, w, s, d);
#define GPE_YCASE_f32(n)                                  \
  case n * 8 + 1: return gpe::launch_y<float, n, 1>(a, c, s, d); \
  case n * 8 + 2: return gpe::launch_y<float, n, 2>(a, c, s, d);
#define GPE_1CASE_f32(n)                                   \
  case n * 8 + 1: return gpe::launch_1d<float, n, 1>(a, s); \
  case n * 8 + 2: return gpe::launch_1d<float, n, 2>(a, s);

#define GPE_XCASE34_f32(n)                                       \
  case n * 8 + 3: return gpe::launch_x<float, n, 3>(a, nlines, w, s, d); \
  case n * 8 + 4: return gpe::launch_x<float, n, 4>(a, nlines, w, s, d);
#define GPE_YCASE34_f32(n)                                  \
  case n * 8 + 3: return gpe::launch_y<float, n, 3>(a, c, s, d); \
  case n * 8 + 4: return gpe::launch_y<float, n, 4>(a, c, s, d);
#define GPE_1CASE34_f32(n)                                   \
  case n * 8 + 3: return gpe::launch_1d<float, n, 3>(a, s); \
  case n * 8 + 4: return gpe::launch_1d<float, n, 4>(a, s);

#define GPE_KCASE_f32(n) \
  case n: return gpe::launch_kabs2<float, n>(a, s);

#define GPE_WCASE_f32(n) \
  case n: return gpe::launch_kwin<float, n>(a, s);

#define GPE_RCASE_f32(n) \
  case n: return gpe::launch_rowcorr<float, n>(a, nblk, s);

#define GPE_SCASE_f32(n) \
  case n: return gpe::launch_sweep<float, n>(a, s);

GPE_DEFINE_RUNNERS(float, f32)

int gpe_rowcorr_blocks(int64_t N, int64_t nlines) { return gpe::rowcorr_blocks(N, nlines); }
