// Template launchers shared by the per-precision instantiation units.
#pragma once
#include "gpe_launch.hpp"

namespace gpe {

template <typename K>
static hipError_t set_lds(K k, size_t lds) {
  if (lds > 65536) return hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  return hipSuccess;
}

template <typename R, int N, int M>
static hipError_t launch_x(const XPassArgs& a, int64_t nlines, hipStream_t s) {
  constexpr int W = x_lines_per_wg(N);
  constexpr int T = N / kE;
  const size_t lds = xlds_bytes<R, N, M, W>();
  auto k = a.pt.noise_kind ? x_pass_kernel<R, N, M, W, true> : x_pass_kernel<R, N, M, W, false>;
  hipError_t e = set_lds(k, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3((unsigned)((nlines + W - 1) / W)), dim3(M * W * T), lds, s, a);
  return hipGetLastError();
}

template <typename R, int N, int M, int C>
static hipError_t launch_y_cols(const YPassArgs& a, hipStream_t s) {
  constexpr int T = N / kE;
  const size_t lds = ylds_bytes<R, N, M, C>();
  auto k = y_pass_kernel<R, N, M, C>;
  hipError_t e = set_lds(k, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3((unsigned)a.nblocks), dim3(M * C * T), lds, s, a);
  return hipGetLastError();
}

// Columns per workgroup of the y-pass, chosen per launch from two shapes:
// C = y_cols_per_wg(N), which keeps enough workgroups for a single grid, and
// 2C when that still leaves >= kWideMinBlocks workgroups (8 per CU).  Measured
// on 1024 trajectories of 256x256 (FP32, noise): 280 us per launch at C = 4,
// 250 us at C = 8, 352 us at C = 2; a single 512x512 grid is fastest at C = 2
// (6.9 us wave span against 8.5 at C = 4 and 14.8 at C = 8).
constexpr int64_t kWideMinBlocks = 2048;

template <typename R, int N, int M>
static hipError_t launch_y(const YPassArgs& a, hipStream_t s) {
  constexpr int T = N / kE;
  constexpr int C = y_cols_per_wg(N);
  constexpr int C2 = 2 * C;
  constexpr bool wide_ok = C2 <= 16 && M * C2 * T <= 1024 && ylds_bytes<R, N, M, C2>() <= 160 * 1024;
  if constexpr (wide_ok) {
    if (a.nx % C2 == 0) {
      YPassArgs b = a;
      b.ncolblk = a.nx / C2;
      b.nblocks = b.ncolblk * (a.nblocks / a.ncolblk);
      b.xcd_remap = (b.nblocks % 8 == 0) ? 1 : 0;
      if (b.nblocks >= kWideMinBlocks) return launch_y_cols<R, N, M, C2>(b, s);
    }
  }
  return launch_y_cols<R, N, M, C>(a, s);
}

template <typename R, int N, int M>
static hipError_t launch_1d(const Step1DArgs& a, hipStream_t s) {
  constexpr int W = x_lines_per_wg(N);
  constexpr int T = N / kE;
  const size_t lds = xlds_bytes<R, N, M, W>();
  auto k = a.pt.noise_kind ? step1d_kernel<R, N, M, W, true> : step1d_kernel<R, N, M, W, false>;
  hipError_t e = set_lds(k, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3((unsigned)((a.nlines + W - 1) / W)), dim3(M * W * T), lds, s, a);
  return hipGetLastError();
}

template <typename R, int N>
static hipError_t launch_sweep(const SweepArgs& a, hipStream_t s) {
  constexpr int W = x_lines_per_wg(N);
  constexpr int T = N / kE;
  const size_t lds = xlds_bytes<R, N, 1, W>();
  auto k = a.noise_kind ? sweep1d_kernel<R, N, W, true> : sweep1d_kernel<R, N, W, false>;
  hipError_t e = set_lds(k, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3((unsigned)((a.ninst + W - 1) / W)), dim3(W * T), lds, s, a);
  return hipGetLastError();
}

}  // namespace gpe

#define GPE_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)

#define GPE_DEFINE_RUNNERS(R, SUFFIX)                                                                  \
  hipError_t gpe_run_x_##SUFFIX(int M, int64_t N, const gpe::XPassArgs& a, int64_t nlines, hipStream_t s) { \
    switch (N * 4 + M) {                                                                               \
      GPE_SIZES(GPE_XCASE_##SUFFIX)                                                                    \
    }                                                                                                  \
    return hipErrorInvalidValue;                                                                       \
  }                                                                                                    \
  hipError_t gpe_run_y_##SUFFIX(int M, int64_t N, const gpe::YPassArgs& a, hipStream_t s) {           \
    switch (N * 4 + M) {                                                                               \
      GPE_SIZES(GPE_YCASE_##SUFFIX)                                                                    \
    }                                                                                                  \
    return hipErrorInvalidValue;                                                                       \
  }                                                                                                    \
  hipError_t gpe_run_sweep_##SUFFIX(int64_t N, const gpe::SweepArgs& a, hipStream_t s) {               \
    switch (N) {                                                                                       \
      GPE_SIZES(GPE_SCASE_##SUFFIX)                                                                    \
    }                                                                                                  \
    return hipErrorInvalidValue;                                                                       \
  }                                                                                                    \
  hipError_t gpe_run_1d_##SUFFIX(int M, int64_t N, const gpe::Step1DArgs& a, hipStream_t s) {         \
    switch (N * 4 + M) {                                                                               \
      GPE_SIZES(GPE_1CASE_##SUFFIX)                                                                    \
    }                                                                                                  \
    return hipErrorInvalidValue;                                                                       \
  }
