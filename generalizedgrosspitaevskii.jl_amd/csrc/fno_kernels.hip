// Kernels of the Fourier-neural-operator warm start (no reference
// counterpart: the reference has no learned component).
//
//   gpe_fno_gemm     C = alpha A B + beta C, batched, FP32 storage with
//                    arbitrary element strides, operands rounded to BF16 and
//                    multiplied on the
//                    matrix cores (v_mfma_f32_16x16x16_bf16) with FP32
//                    accumulation.  Every dense product of the network and
//                    of its backward pass (truncated DFTs, spectral channel
//                    mixing, 1x1 convolutions, lifting, projection) is one.
//   gpe_fno_gemm_f32 the same product with FP32 operands
//                    (v_mfma_f32_16x16x4_f32): the band-limiting transforms
//                    of a prediction, where BF16-rounded basis functions
//                    would leave broadband error.
//   gpe_fno_bias_act bias add + GELU (tanh form) and its backward.
//   gpe_fno_adam     Adam update of FP32 parameters.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "../../include/gpe_hip.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even FP32 -> BF16 bits
__device__ __forceinline__ short to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u) return (short)(u >> 16);  // inf / nan
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

constexpr int TM = 64, TN = 64, TK = 32;

// Element (i, j) of an operand sits at base + i * rs + j * cs; batch b adds
// b * bs.  Transposed and strided views (channels, modes) are stride choices.
struct GemmArgs {
  const float* A;
  const float* B;
  float* C;
  int64_t m, n, k;
  int64_t ars, acs, brs, bcs, crs, ccs;
  int64_t sa, sb, sc;
  float alpha, beta;
};

// 256 threads = 4 wavefronts; the workgroup computes a 64 x 64 tile of C,
// each wavefront a 32 x 32 quarter as 2 x 2 MFMA 16x16 blocks.  The k loop
// runs over tiles of depth 32 that are double-buffered in LDS (BF16 or FP32
// per kBF16): while the matrix cores work on tile t, the next tile's values
// are already in flight into registers, and one barrier per tile separates
// the buffers.
//
// MFMA 16x16x16 BF16 operand layout (lane l of the wavefront):
//   A fragment: row l % 16, k = 4 (l / 16) .. 4 (l / 16) + 3
//   B fragment: k = 4 (l / 16) .. +3, column l % 16
// MFMA 16x16x4 FP32: A row l % 16, k = l / 16; B k = l / 16, column l % 16.
// Both: D rows 4 (l / 16) .. +3, column l % 16.
template <bool kBF16>
struct GemmLds {
  using T = typename std::conditional<kBF16, short, float>::type;
  T As[2][TM][TK + 4];
  T Bs[2][TK][TN + 4];
};

template <bool kBF16>
__device__ __forceinline__ typename GemmLds<kBF16>::T lds_value(float v) {
  if constexpr (kBF16)
    return to_bf16(v);
  else
    return v;
}

constexpr int kPerThread = TM * TK / 256;  // A (and B) values each thread stages per tile
static_assert(TM * TK == TK * TN, "square staging");

template <bool kBF16>
__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
  __shared__ GemmLds<kBF16> L;
  const int64_t bz = blockIdx.z;
  const float* A = g.A + bz * g.sa;
  const float* B = g.B + bz * g.sb;
  float* C = g.C + bz * g.sc;
  const int64_t m0 = (int64_t)blockIdx.y * TM, n0 = (int64_t)blockIdx.x * TN;
  const int tid = threadIdx.x;
  const int wave = tid / 64, lane = tid % 64;
  const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;
  floatx4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  // staging map: A value e -> (row e / TK, k e % TK), B value e -> (k e / TN, col e % TN)
  float ra[kPerThread], rb[kPerThread];
  auto fetch = [&](int64_t k0) {
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
      const int e = tid + q * 256;
      const int64_t gi = m0 + e / TK, gka = k0 + e % TK;
      ra[q] = (gi < g.m && gka < g.k) ? A[gi * g.ars + gka * g.acs] : 0.f;
      const int64_t gkb = k0 + e / TN, gj = n0 + e % TN;
      rb[q] = (gkb < g.k && gj < g.n) ? B[gkb * g.brs + gj * g.bcs] : 0.f;
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int q = 0; q < kPerThread; ++q) {
      const int e = tid + q * 256;
      L.As[buf][e / TK][e % TK] = lds_value<kBF16>(ra[q]);
      L.Bs[buf][e / TN][e % TN] = lds_value<kBF16>(rb[q]);
    }
  };
  const int64_t nt = (g.k + TK - 1) / TK;
  fetch(0);
  stage(0);
  __syncthreads();
  for (int64_t kt = 0; kt < nt; ++kt) {
    const int buf = (int)(kt & 1);
    if (kt + 1 < nt) fetch((kt + 1) * TK);  // in flight while the tile below is consumed
    if constexpr (kBF16) {
#pragma unroll
      for (int ks = 0; ks < TK; ks += 16) {
        const int kb = ks + 4 * (lane / 16);
        for (int i = 0; i < 2; ++i) {
          bf16x4 a;
          const int r = wm + 16 * i + lane % 16;
          a[0] = L.As[buf][r][kb];
          a[1] = L.As[buf][r][kb + 1];
          a[2] = L.As[buf][r][kb + 2];
          a[3] = L.As[buf][r][kb + 3];
          for (int j = 0; j < 2; ++j) {
            bf16x4 b;
            const int c = wn + 16 * j + lane % 16;
            b[0] = L.Bs[buf][kb][c];
            b[1] = L.Bs[buf][kb + 1][c];
            b[2] = L.Bs[buf][kb + 2][c];
            b[3] = L.Bs[buf][kb + 3][c];
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
          }
        }
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < TK; ks += 4) {
        const int kk = ks + lane / 16;
        for (int i = 0; i < 2; ++i) {
          const float a = L.As[buf][wm + 16 * i + lane % 16][kk];
          for (int j = 0; j < 2; ++j) {
            const float b = L.Bs[buf][kk][wn + 16 * j + lane % 16];
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i][j], 0, 0, 0);
          }
        }
      }
    }
    // the other buffer was last read in the previous tile, before the
    // barrier that ended it
    if (kt + 1 < nt) stage(buf ^ 1);
    __syncthreads();
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int q = 0; q < 4; ++q) {
        const int64_t r = m0 + wm + 16 * i + 4 * (lane / 16) + q;
        const int64_t c = n0 + wn + 16 * j + lane % 16;
        if (r < g.m && c < g.n) {
          float* dst = C + r * g.crs + c * g.ccs;
          const float prev = g.beta != 0.f ? *dst : 0.f;
          *dst = g.alpha * acc[i][j][q] + g.beta * prev;
        }
      }
}

__device__ __forceinline__ float gelu_tanh(float x, float* dgelu) {
  const float c = 0.7978845608028654f, a = 0.044715f;
  const float u = c * (x + a * x * x * x);
  const float t = tanhf(u);
  if (dgelu) *dgelu = 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * c * (1.f + 3.f * a * x * x);
  return 0.5f * x * (1.f + t);
}

// y = act(x + bias[channel]); rows of `cols` values, channel = (row % C) when
// per_row, else the column index (channels last).
struct BiasActArgs {
  const float* x;
  const float* bias;
  float* pre;  // optional copy of x + bias (for the backward)
  float* y;
  int64_t rows, cols, C;
  int channels_last, act;
};

__global__ void bias_act_kernel(BiasActArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.rows * a.cols) return;
  const int64_t ch = a.channels_last ? (i % a.cols) : ((i / a.cols) % a.C);
  const float v = a.x[i] + (a.bias ? a.bias[ch] : 0.f);
  if (a.pre) a.pre[i] = v;
  a.y[i] = a.act ? gelu_tanh(v, nullptr) : v;
}

// dx = dy * gelu'(pre)
__global__ void act_backward_kernel(const float* pre, const float* dy, float* dx, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float d;
  gelu_tanh(pre[i], &d);
  dx[i] = dy[i] * d;
}

struct AdamArgs {
  float* p;
  const float* grad;
  float* m;
  float* v;
  int64_t n;
  float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
};

__global__ void adam_kernel(AdamArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const float g = a.grad[i];
  const float m = a.b1 * a.m[i] + (1.f - a.b1) * g;
  const float v = a.b2 * a.v[i] + (1.f - a.b2) * g * g;
  a.m[i] = m;
  a.v[i] = v;
  a.p[i] -= a.lr * (m / a.c1) / (sqrtf(v / a.c2) + a.eps);
}

// y = a x + b y
__global__ void axpby_kernel(const float* x, float* y, float a, float b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] = a * x[i] + b * y[i];
}

}  // namespace

extern "C" int gpe_fno_axpby(const float* x, float* y, float a, float b, int64_t n, void* stream) {
  if (!x || !y || n < 1) return GPE_ERR_INVALID;
  hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, x, y, a, b,
                     n);
  return hipGetLastError() == hipSuccess ? GPE_OK : GPE_ERR_DEVICE;
}

template <bool kBF16>
static int launch_gemm(const float* A, const float* B, float* C, int64_t m, int64_t n, int64_t k,
                       const int64_t* strides, int64_t batch, float alpha, float beta, void* stream) {
  if (!A || !B || !C || !strides || m < 1 || n < 1 || k < 1 || batch < 1 || batch > 65535) return GPE_ERR_INVALID;
  if ((m + TM - 1) / TM > 65535) return GPE_ERR_INVALID;  // grid.y limit
  GemmArgs g{A, B, C, m, n, k, strides[0], strides[1], strides[2], strides[3], strides[4], strides[5],
             strides[6], strides[7], strides[8], alpha, beta};
  const dim3 grid((unsigned)((n + TN - 1) / TN), (unsigned)((m + TM - 1) / TM), (unsigned)batch);
  hipLaunchKernelGGL(gemm_kernel<kBF16>, grid, dim3(256), 0, (hipStream_t)stream, g);
  return hipGetLastError() == hipSuccess ? GPE_OK : GPE_ERR_DEVICE;
}

extern "C" int gpe_fno_gemm(const float* A, const float* B, float* C, int64_t m, int64_t n, int64_t k,
                            const int64_t* strides, int64_t batch, float alpha, float beta, void* stream) {
  return launch_gemm<true>(A, B, C, m, n, k, strides, batch, alpha, beta, stream);
}

extern "C" int gpe_fno_gemm_f32(const float* A, const float* B, float* C, int64_t m, int64_t n, int64_t k,
                                const int64_t* strides, int64_t batch, float alpha, float beta, void* stream) {
  return launch_gemm<false>(A, B, C, m, n, k, strides, batch, alpha, beta, stream);
}

extern "C" int gpe_fno_bias_act(const float* x, const float* bias, float* pre, float* y, int64_t rows, int64_t cols,
                                int64_t C, int32_t channels_last, int32_t act, void* stream) {
  if (!x || !y || rows < 1 || cols < 1) return GPE_ERR_INVALID;
  BiasActArgs a{x, bias, pre, y, rows, cols, C < 1 ? 1 : C, channels_last, act};
  const int64_t n = rows * cols;
  hipLaunchKernelGGL(bias_act_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a);
  return hipGetLastError() == hipSuccess ? GPE_OK : GPE_ERR_DEVICE;
}

extern "C" int gpe_fno_act_backward(const float* pre, const float* dy, float* dx, int64_t n, void* stream) {
  if (!pre || !dy || !dx || n < 1) return GPE_ERR_INVALID;
  hipLaunchKernelGGL(act_backward_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, pre,
                     dy, dx, n);
  return hipGetLastError() == hipSuccess ? GPE_OK : GPE_ERR_DEVICE;
}

extern "C" int gpe_fno_adam(float* p, const float* grad, float* m, float* v, int64_t n, float lr, float b1, float b2,
                            float eps, int64_t step, void* stream) {
  if (!p || !grad || !m || !v || n < 1 || step < 1) return GPE_ERR_INVALID;
  AdamArgs a{p, grad, m, v, n, lr, b1, b2, eps, 1.f - powf(b1, (float)step), 1.f - powf(b2, (float)step)};
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a);
  return hipGetLastError() == hipSuccess ? GPE_OK : GPE_ERR_DEVICE;
}
