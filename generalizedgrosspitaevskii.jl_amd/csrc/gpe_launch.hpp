// Launch shapes and the per-precision launch entry points of the fused
// kernels (instantiated in gpe_inst_f32.hip / gpe_inst_f64.hip).
#pragma once
#include "gpe_kernels.hpp"

namespace gpe {

// Lines per workgroup of the x-pass and of the 1D whole-step kernel: W*N/8
// threads per component, at least one wavefront, 64..512 threads.
__host__ __device__ constexpr int x_lines_per_wg(int N) { return N >= 512 ? 1 : 512 / N; }

// Columns per workgroup of the y-pass: C*N/8 threads per component, at least
// one wavefront, at most 16 columns.
__host__ __device__ constexpr int y_cols_per_wg(int N) {
  return (128 / (N / kE)) < 1 ? 1 : ((128 / (N / kE)) > 16 ? 16 : 128 / (N / kE));
}

// Bytes of the two LDS buffers of `lines` padded lines of N complex values.
__host__ __device__ constexpr size_t pass_lds_bytes(int N, int lines, bool f64) {
  return 2ull * lines * lds_line_stride(N) * (f64 ? 16 : 8);
}

// Shapes a plan may give the x-pass: x_lines_per_wg(N) rows per workgroup or
// twice as many, within 1024 threads and 160 KiB of LDS.
__host__ __device__ constexpr bool x_plan_valid(int N, int M, bool f64, int w) {
  return (w == x_lines_per_wg(N) || w == 2 * x_lines_per_wg(N)) && M * w * (N / kE) <= 1024 &&
         pass_lds_bytes(N, M * w, f64) <= 160 * 1024;
}

// Shapes a plan may give the y-pass: C = y_cols_per_wg(N) columns per
// workgroup, 2C, or C/2 while that still fills a wavefront per component; at
// most 16 columns, 1024 threads and 160 KiB of LDS.
__host__ __device__ constexpr bool y_plan_valid(int N, int M, bool f64, int c) {
  return (c == y_cols_per_wg(N) || c == 2 * y_cols_per_wg(N) ||
          (y_cols_per_wg(N) > 1 && c == y_cols_per_wg(N) / 2 && c * (N / kE) >= 64)) &&
         c >= 1 && c <= 16 && M * c * (N / kE) <= 1024 && pass_lds_bytes(N, M * c, f64) <= 160 * 1024;
}

// The built-in column count of a y-pass over nx columns of `batch` grids:
// C = y_cols_per_wg(N), which keeps enough workgroups for a single grid, and
// 2C when that still leaves >= kWideMinBlocks workgroups (8 per CU).
// Measured on 1024 trajectories of 256x256 (FP32, noise): 280 us per launch
// at C = 4, 250 us at C = 8, 352 us at C = 2; a single 512x512 grid is
// fastest at C = 2 (6.9 us wave span against 8.5 at C = 4 and 14.8 at C = 8).
constexpr int64_t kWideMinBlocks = 2048;
__host__ __device__ constexpr int y_rule_cols(int N, int M, bool f64, int64_t nx, int64_t batch) {
  return (y_plan_valid(N, M, f64, 2 * y_cols_per_wg(N)) && nx % (2 * y_cols_per_wg(N)) == 0 &&
          (nx / (2 * y_cols_per_wg(N))) * batch >= kWideMinBlocks)
             ? 2 * y_cols_per_wg(N)
             : y_cols_per_wg(N);
}

// A launch described instead of issued: kernel, shape and a copy of its
// argument struct, for adding the launch to a HIP graph as a kernel node.
struct LaunchDesc {
  const void* func = nullptr;
  dim3 grid, block;
  unsigned lds = 0;
  alignas(16) unsigned char arg[1024];
};

}  // namespace gpe

// Pass launchers: w / c rows / columns per workgroup, 0 for the built-in
// shape (x_lines_per_wg, y_rule_cols); other values must satisfy
// x_plan_valid / y_plan_valid (hipErrorInvalidValue otherwise).  With d
// non-null the x / y launchers fill *d and launch nothing.
hipError_t gpe_run_x_f32(int M, int64_t N, const gpe::XPassArgs& a, int64_t nlines, int w, hipStream_t s,
                         gpe::LaunchDesc* d = nullptr);
hipError_t gpe_run_x_f64(int M, int64_t N, const gpe::XPassArgs& a, int64_t nlines, int w, hipStream_t s,
                         gpe::LaunchDesc* d = nullptr);
hipError_t gpe_run_y_f32(int M, int64_t N, const gpe::YPassArgs& a, int c, hipStream_t s,
                         gpe::LaunchDesc* d = nullptr);
hipError_t gpe_run_y_f64(int M, int64_t N, const gpe::YPassArgs& a, int c, hipStream_t s,
                         gpe::LaunchDesc* d = nullptr);
hipError_t gpe_run_sweep_f32(int64_t N, const gpe::SweepArgs& a, hipStream_t s);
hipError_t gpe_run_sweep_f64(int64_t N, const gpe::SweepArgs& a, hipStream_t s);
hipError_t gpe_run_1d_f32(int M, int64_t N, const gpe::Step1DArgs& a, hipStream_t s);
hipError_t gpe_run_1d_f64(int M, int64_t N, const gpe::Step1DArgs& a, hipStream_t s);
hipError_t gpe_run_kabs2_f32(int64_t N, const gpe::KAbs2Args& a, hipStream_t s);
hipError_t gpe_run_kabs2_f64(int64_t N, const gpe::KAbs2Args& a, hipStream_t s);
hipError_t gpe_run_kwin_f32(int64_t N, const gpe::KWinArgs& a, hipStream_t s);
hipError_t gpe_run_kwin_f64(int64_t N, const gpe::KWinArgs& a, hipStream_t s);
hipError_t gpe_run_rowcorr_f32(int64_t N, const gpe::RowCorrArgs& a, int nblk, hipStream_t s);
hipError_t gpe_run_rowcorr_f64(int64_t N, const gpe::RowCorrArgs& a, int nblk, hipStream_t s);
int gpe_rowcorr_blocks(int64_t N, int64_t nlines);
