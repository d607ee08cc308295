// Template launchers shared by the per-precision instantiation units.
#pragma once
#include <cstring>

#include "gpe_launch.hpp"

namespace gpe {

template <typename K>
static hipError_t set_lds(K k, size_t lds) {
  if (lds > 65536) return hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  return hipSuccess;
}

template <typename K, typename A>
static hipError_t describe(LaunchDesc* d, K k, dim3 grid, dim3 block, size_t lds, const A& a) {
  static_assert(sizeof(A) <= sizeof(d->arg), "LaunchDesc::arg too small for the argument struct");
  d->func = (const void*)k;
  d->grid = grid;
  d->block = block;
  d->lds = (unsigned)lds;
  memcpy(d->arg, &a, sizeof(A));
  return hipSuccess;
}

template <typename R, int N, int M, int W>
static hipError_t launch_x_rows(const XPassArgs& a, int64_t nlines, hipStream_t s, LaunchDesc* d) {
  constexpr int T = N / kE;
  const size_t lds = xlds_bytes<R, N, M, W>();
  auto k = a.pt.noise_kind ? x_pass_kernel<R, N, M, W, true> : x_pass_kernel<R, N, M, W, false>;
  hipError_t e = set_lds(k, lds);
  if (e != hipSuccess) return e;
  const dim3 grid((unsigned)((nlines + W - 1) / W)), block(M * W * T);
  if (d) return describe(d, k, grid, block, lds, a);
  hipLaunchKernelGGL(k, grid, block, lds, s, a);
  return hipGetLastError();
}

// Rows per workgroup of the x-pass: w == 0 or x_lines_per_wg(N) is the
// built-in shape; a measured plan may also take twice as many
// (x_plan_valid).  Other values are rejected.
template <typename R, int N, int M>
static hipError_t launch_x(const XPassArgs& a, int64_t nlines, int w, hipStream_t s, LaunchDesc* d) {
  constexpr int W = x_lines_per_wg(N);
  if (w == 0 || w == W) return launch_x_rows<R, N, M, W>(a, nlines, s, d);
  if constexpr (x_plan_valid(N, M, sizeof(R) == 8, 2 * W))
    if (w == 2 * W) return launch_x_rows<R, N, M, 2 * W>(a, nlines, s, d);
  return hipErrorInvalidValue;
}

template <typename R, int N, int M, int C>
static hipError_t launch_y_cols(const YPassArgs& a, hipStream_t s, LaunchDesc* d = nullptr) {
  constexpr int T = N / kE;
  const size_t lds = ylds_bytes<R, N, M, C>();
  auto k = y_pass_kernel<R, N, M, C>;
  hipError_t e = set_lds(k, lds);
  if (e != hipSuccess) return e;
  const dim3 grid((unsigned)a.nblocks), block(M * C * T);
  if (d) return describe(d, k, grid, block, lds, a);
  hipLaunchKernelGGL(k, grid, block, lds, s, a);
  return hipGetLastError();
}

// The y-pass with C' columns per workgroup; `a` carries the block counts of
// y_cols_per_wg(N) columns and is re-derived for C'.
template <typename R, int N, int M, int CP>
static hipError_t launch_y_recount(const YPassArgs& a, hipStream_t s, LaunchDesc* d) {
  if (a.nx % CP != 0) return hipErrorInvalidValue;
  YPassArgs b = a;
  b.ncolblk = a.nx / CP;
  b.nblocks = b.ncolblk * (a.nblocks / a.ncolblk);
  b.xcd_remap = (b.nblocks % 8 == 0) ? 1 : 0;
  return launch_y_cols<R, N, M, CP>(b, s, d);
}

// c == 0: the built-in rule (y_rule_cols); otherwise exactly c columns, one
// of C/2, C, 2C where y_plan_valid allows it.
template <typename R, int N, int M>
static hipError_t launch_y(const YPassArgs& a, int c, hipStream_t s, LaunchDesc* d) {
  constexpr int C = y_cols_per_wg(N);
  constexpr int C2 = 2 * C;
  constexpr int CH = C > 1 ? C / 2 : 1;
  constexpr bool wide_ok = y_plan_valid(N, M, sizeof(R) == 8, C2);
  if (c == 0) c = y_rule_cols(N, M, sizeof(R) == 8, a.nx, a.nblocks / a.ncolblk);
  if (c == C) return launch_y_cols<R, N, M, C>(a, s, d);
  if constexpr (wide_ok)
    if (c == C2) return launch_y_recount<R, N, M, C2>(a, s, d);
  if constexpr (C > 1 && y_plan_valid(N, M, sizeof(R) == 8, CH))
    if (c == CH) return launch_y_recount<R, N, M, CH>(a, s, d);
  return hipErrorInvalidValue;
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

template <typename R, int N>
static hipError_t launch_kabs2(const KAbs2Args& a, hipStream_t s) {
  constexpr int W = x_lines_per_wg(N);
  constexpr int T = N / kE;
  const size_t lds = xlds_bytes<R, N, 1, W>();
  auto k = kspace_abs2_kernel<R, N, W>;
  hipError_t e = set_lds(k, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3((unsigned)((a.batch + W - 1) / W)), dim3(W * T), lds, s, a);
  return hipGetLastError();
}

template <typename R, int N>
static hipError_t launch_kwin(const KWinArgs& a, hipStream_t s) {
  constexpr int W = x_lines_per_wg(N);
  constexpr int T = N / kE;
  const size_t lds = xlds_bytes<R, N, 1, W>();
  auto k = kspace_windowed_kernel<R, N, W>;
  hipError_t e = set_lds(k, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3((unsigned)((a.batch + W - 1) / W)), dim3(W * T), lds, s, a);
  return hipGetLastError();
}

// Workgroups of row_autocorr_kernel: enough to fill the chip, each looping
// over its share of the lines.
constexpr int kRowCorrBlocks = 1024;

template <typename R, int N>
static hipError_t launch_rowcorr(const RowCorrArgs& a, int nblk, hipStream_t s) {
  constexpr int W = x_lines_per_wg(N);
  constexpr int T = N / kE;
  const size_t lds = xlds_bytes<R, N, 1, W>();
  auto k = row_autocorr_kernel<R, N, W>;
  hipError_t e = set_lds(k, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3((unsigned)nblk), dim3(W * T), lds, s, a);
  return hipGetLastError();
}

// Blocks row_autocorr_kernel uses for nlines lines of length N.
inline int rowcorr_blocks(int64_t N, int64_t nlines) {
  const int64_t W = x_lines_per_wg((int)N);
  const int64_t need = (nlines + W - 1) / W;
  return (int)(need < kRowCorrBlocks ? need : kRowCorrBlocks);
}

}  // namespace gpe

#define GPE_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)
// Three and four components: lines up to 2048 (a workgroup holds M lines of
// N/8 threads; 4096 would need 2048 threads).  The FP64 files stop at 1024
// (LDS), see GPE_SIZES34_f64.
#define GPE_SIZES34_f32(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048)
#define GPE_SIZES34_f64(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024)

#define GPE_DEFINE_RUNNERS(R, SUFFIX)                                                                  \
  hipError_t gpe_run_x_##SUFFIX(int M, int64_t N, const gpe::XPassArgs& a, int64_t nlines, int w, hipStream_t s,  \
                                gpe::LaunchDesc* d) {                                                  \
    switch (N * 8 + M) {                                                                               \
      GPE_SIZES(GPE_XCASE_##SUFFIX)                                                                    \
      GPE_SIZES34_##SUFFIX(GPE_XCASE34_##SUFFIX)                                                       \
    }                                                                                                  \
    return hipErrorInvalidValue;                                                                       \
  }                                                                                                    \
  hipError_t gpe_run_y_##SUFFIX(int M, int64_t N, const gpe::YPassArgs& a, int c, hipStream_t s,       \
                                gpe::LaunchDesc* d) {                                                  \
    switch (N * 8 + M) {                                                                               \
      GPE_SIZES(GPE_YCASE_##SUFFIX)                                                                    \
      GPE_SIZES34_##SUFFIX(GPE_YCASE34_##SUFFIX)                                                       \
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
    switch (N * 8 + M) {                                                                               \
      GPE_SIZES(GPE_1CASE_##SUFFIX)                                                                    \
      GPE_SIZES34_##SUFFIX(GPE_1CASE34_##SUFFIX)                                                       \
    }                                                                                                  \
    return hipErrorInvalidValue;                                                                       \
  } \
  hipError_t gpe_run_kabs2_##SUFFIX(int64_t N, const gpe::KAbs2Args& a, hipStream_t s) {               \
    switch (N) {                                                                                       \
      GPE_SIZES(GPE_KCASE_##SUFFIX)                                                                    \
    }                                                                                                  \
    return hipErrorInvalidValue;                                                                       \
  } \
  hipError_t gpe_run_kwin_##SUFFIX(int64_t N, const gpe::KWinArgs& a, hipStream_t s) {                 \
    switch (N) {                                                                                       \
      GPE_SIZES(GPE_WCASE_##SUFFIX)                                                                    \
    }                                                                                                  \
    return hipErrorInvalidValue;                                                                       \
  } \
  hipError_t gpe_run_rowcorr_##SUFFIX(int64_t N, const gpe::RowCorrArgs& a, int nblk, hipStream_t s) {  \
    switch (N) {                                                                                       \
      GPE_SIZES(GPE_RCASE_##SUFFIX)                                                                    \
    }                                                                                                  \
    return hipErrorInvalidValue;                                                                       \
  }
