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

}  // namespace gpe

hipError_t gpe_run_x_f32(int M, int64_t N, const gpe::XPassArgs& a, int64_t nlines, hipStream_t s);
hipError_t gpe_run_x_f64(int M, int64_t N, const gpe::XPassArgs& a, int64_t nlines, hipStream_t s);
hipError_t gpe_run_y_f32(int M, int64_t N, const gpe::YPassArgs& a, hipStream_t s);
hipError_t gpe_run_y_f64(int M, int64_t N, const gpe::YPassArgs& a, hipStream_t s);
hipError_t gpe_run_sweep_f32(int64_t N, const gpe::SweepArgs& a, hipStream_t s);
hipError_t gpe_run_sweep_f64(int64_t N, const gpe::SweepArgs& a, hipStream_t s);
hipError_t gpe_run_1d_f32(int M, int64_t N, const gpe::Step1DArgs& a, hipStream_t s);
hipError_t gpe_run_1d_f64(int M, int64_t N, const gpe::Step1DArgs& a, hipStream_t s);
