// Fused Strang-splitting kernels for gfx950.
//
// One Strang step of the reference (src/strang_splitting.jl:86-90)
//     half(t+dt/2) ; FFT ; exp(-i dt D(k)) ; IFFT ; half(t+dt)
// is regrouped into streaming passes over the field:
//
//   2D:  x-pass   [IFFT_x] -> half-step(s) -> [frame write] -> [FFT_x]
//        y-pass   FFT_y -> exp(-i dt D(kx,ky)) / (Nx Ny) -> IFFT_y
//        The real-space half-steps that end step s and start step s+1 are
//        pointwise and adjacent, so one x-pass applies both.
//   1D:  whole-step kernel: every line stays in registers for a block of
//        steps (half, FFT, exp(-i dt D), IFFT, half, repeated).
//
// Every pass updates the field in place: a workgroup reads exactly the lines
// it writes.  The real-space half-step is muladd_kernel! of
// src/kernels.jl:37-54 for the descriptor forms
//     G_c(u) = c_c + sum_j g_cj |u_j|^2,  eta_c constant,
//     F_c(r, tau) = profile_c(r) * envelope(tau).
#pragma once
#include "gpe_device.hpp"

namespace gpe {

constexpr int kMaxM = 2;

// Parameters of one real-space half-step (src/strang_splitting.jl:78-84).
struct HalfStep {
  double delta;     // the half-step's dt (= dt/2)
  double fnow[2];   // (delta/2) * envelope(previous pump time)   (complex)
  double fnext[2];  // (delta/2) * envelope(this pump time)       (complex)
  uint64_t call;    // Philox draw counter of this half-step
};

// Per-point terms shared by every real-space half-step.
struct PointTerms {
  int has_g;  // nonlinearity G_c = c_c + sum_j g[c][j] |u_j|^2 (complex)
  double g[kMaxM][kMaxM][2];
  double c[kMaxM][2];
  int pump_kind;                 // 0 none, 1 table, 2 separable
  const void* pump_a[kMaxM];     // table: plane of the spatial grid; separable: dim-0 factor
  const void* pump_b[kMaxM];     // separable: dim-1 factor (2D), else null
  int pot_kind;                  // 0 none, else exp(-i dt/2 V) planes per component
  const void* pot[kMaxM];
  int noise_kind;                // 0 none, 1 complex, 2 real
  double eta[kMaxM];
  uint32_t seed_lo, seed_hi;
};

template <typename R>
__device__ __forceinline__ Cx<R> ldc(const void* p, int64_t i) {
  return reinterpret_cast<const Cx<R>*>(p)[i];
}
template <typename R>
__device__ __forceinline__ Cx<R> cxd(const double (&a)[2]) {
  return mk<R>((R)a[0], (R)a[1]);
}

// The Wiener increments -i sqrt(delta) eta_c xi_c of a thread's M x E values.
// Element m and m + E/2 (x and x + N/2 on the line) share one Philox draw:
// ctr = (P mod 2^32, c + 16 (P >> 32), call mod 2^32, call >> 32) with
// P = line * N/2 + x, key = seed; words (x,y) feed x, (z,w) feed x + N/2.
template <typename R, int N, int M>
__device__ __forceinline__ void noise_terms(Cx<R> (&nz)[M][kE], const HalfStep& hs, const PointTerms& pt, int t,
                                            int T, int64_t line) {
  const R sq = (R)sqrt(hs.delta);
#pragma unroll
  for (int m = 0; m < kE / 2; ++m) {
    const uint64_t P = (uint64_t)line * (uint64_t)(N / 2) + (uint64_t)(t + m * T);
#pragma unroll
    for (int c = 0; c < M; ++c) {
      U4 ctr;
      ctr.x = (uint32_t)P;
      ctr.y = (uint32_t)c + 16u * (uint32_t)(P >> 32);
      ctr.z = (uint32_t)hs.call;
      ctr.w = (uint32_t)(hs.call >> 32);
      const U4 w = philox4x32_10(ctr, pt.seed_lo, pt.seed_hi);
      Cx<R> z0 = box_muller<R>(w.x, w.y);
      Cx<R> z1 = box_muller<R>(w.z, w.w);
      if (pt.noise_kind == 2) {
        const R s2 = (R)1.41421356237309504880168872420969808;
        z0 = mk<R>(z0.re * s2, (R)0);
        z1 = mk<R>(z1.re * s2, (R)0);
      }
      const R a = sq * (R)pt.eta[c];
      nz[c][m] = mk<R>(a * z0.im, -a * z0.re);
      nz[c][m + kE / 2] = mk<R>(a * z1.im, -a * z1.re);
    }
  }
}

// Real-space half-step (muladd_kernel!, src/kernels.jl:37-54) on the M x E
// values a thread holds.  Element m sits at dim-0 index x = t + m*T of global
// line `line`; `sp_off` is the spatial flat index of x = 0 on that line and
// `y` its dim-1 index (2D; 0 in 1D).
template <typename R, int N, int M>
__device__ __forceinline__ void half_step(Cx<R> (&v)[M][kE], const HalfStep& hs, const PointTerms& pt, int t,
                                          int T, int64_t line, int64_t y, int64_t sp_off) {
  const R delta = (R)hs.delta;
  Cx<R> nz[M][kE];
  if (pt.noise_kind) noise_terms<R, N, M>(nz, hs, pt, t, T, line);
  Cx<R> pb[M];
  if (pt.pump_kind == 2) {
#pragma unroll
    for (int c = 0; c < M; ++c) pb[c] = pt.pump_b[c] ? ldc<R>(pt.pump_b[c], y) : mk<R>((R)1, (R)0);
  }
  const Cx<R> fnow = cxd<R>(hs.fnow), fnext = cxd<R>(hs.fnext);
#pragma unroll
  for (int m = 0; m < kE; ++m) {
    const int x = t + m * T;
    const int64_t pidx = sp_off + x;
    Cx<R> u[M], e[M];
#pragma unroll
    for (int c = 0; c < M; ++c) {
      u[c] = v[c][m];
      e[c] = mk<R>((R)1, (R)0);
    }
    // cis(-delta G(u)) from the pre-update fields (src/kernels.jl:40-44)
    if (pt.has_g) {
      R a2[M];
#pragma unroll
      for (int j = 0; j < M; ++j) a2[j] = abs2(u[j]);
#pragma unroll
      for (int c = 0; c < M; ++c) {
        R gr = (R)pt.c[c][0], gi = (R)pt.c[c][1];
#pragma unroll
        for (int j = 0; j < M; ++j) {
          gr += (R)pt.g[c][j][0] * a2[j];
          gi += (R)pt.g[c][j][1] * a2[j];
        }
        R s, co;
        if constexpr (sizeof(R) == 8) {
          sincos(delta * gr, &s, &co);
        } else {
          sincosf(delta * gr, &s, &co);
        }
        // exp(-i delta (gr + i gi)) = exp(delta gi) (cos(delta gr) - i sin(delta gr))
        const R mag = (gi == (R)0) ? (R)1 : (R)exp(delta * gi);
        e[c] = mk<R>(mag * co, -mag * s);
      }
    }
    if (pt.pot_kind) {
#pragma unroll
      for (int c = 0; c < M; ++c) e[c] = e[c] * ldc<R>(pt.pot[c], pidx);
    }
    // e (u + dt/2 F_now) + dt/2 F_next + noise   (src/kernels.jl:47-48)
#pragma unroll
    for (int c = 0; c < M; ++c) {
      Cx<R> w = u[c];
      Cx<R> fp;
      const bool pumped = pt.pump_kind != 0 && pt.pump_a[c] != nullptr;
      if (pumped) {
        fp = (pt.pump_kind == 1) ? ldc<R>(pt.pump_a[c], pidx) : ldc<R>(pt.pump_a[c], x) * pb[c];
        w = w + fp * fnow;
      }
      w = e[c] * w;
      if (pumped) w = w + fp * fnext;
      if (pt.noise_kind) w = w + nz[c][m];
      v[c][m] = w;
    }
  }
}

// exp(-i dt D(k)) (scaled by 1/prod(N)) applied at flat k index kidx.
// form 0: scale only (no dispersion), 1: scalar, 2: diagonal, 3: full MxM
// column-major.
template <typename R, int M>
__device__ __forceinline__ void apply_ed(Cx<R> (&w)[M], int form, const void* ed, int64_t kidx, R scl) {
  if (form == 0) {
#pragma unroll
    for (int c = 0; c < M; ++c) w[c] = scale(w[c], scl);
  } else if (form == 1) {
    const Cx<R> e = ldc<R>(ed, kidx);
#pragma unroll
    for (int c = 0; c < M; ++c) w[c] = e * w[c];
  } else if (form == 2) {
#pragma unroll
    for (int c = 0; c < M; ++c) w[c] = ldc<R>(ed, kidx * M + c) * w[c];
  } else {
    Cx<R> o[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
      o[i] = ldc<R>(ed, kidx * M * M + i) * w[0];
#pragma unroll
      for (int j = 1; j < M; ++j) o[i] = o[i] + ldc<R>(ed, kidx * M * M + i + M * j) * w[j];
    }
#pragma unroll
    for (int c = 0; c < M; ++c) w[c] = o[c];
  }
}

// ------------------------------------------------------------------ x-pass
enum : int { kInK = 1, kEnd = 2, kStart = 4, kOutK = 8 };

struct XPassArgs {
  void* u[kMaxM];      // field planes, nlines * N values each
  void* frame[kMaxM];  // optional real-space copy written after the end half-step
  int64_t nlines;      // lines along dim 0: prod(n[1:]) * batch
  int64_t ny;          // dim-1 extent (1 for 1D)
  int flags;
  HalfStep h_end, h_start;
  const void* tw;
  int tws;
  PointTerms pt;
};

template <typename R, int N, int M, int W>
__global__ void __launch_bounds__(W * (N / kE)) x_pass_kernel(XPassArgs a) {
  constexpr int T = N / kE;
  constexpr int LS = lds_line_stride(N);
  extern __shared__ unsigned char smem_raw[];
  Cx<R>* lds = reinterpret_cast<Cx<R>*>(smem_raw);
  const int tid = threadIdx.x;
  const int l = tid / T, t = tid % T;
  const int64_t line_raw = (int64_t)blockIdx.x * W + l;
  const bool valid = line_raw < a.nlines;
  const int64_t line = valid ? line_raw : a.nlines - 1;
  const int64_t y = line % a.ny;
  const int64_t sp_off = y * N;
  Cx<R> v[M][kE];
#pragma unroll
  for (int c = 0; c < M; ++c) {
    const Cx<R>* src = reinterpret_cast<const Cx<R>*>(a.u[c]) + line * N;
#pragma unroll
    for (int m = 0; m < kE; ++m) v[c][m] = src[t + m * T];
  }
  Cx<R>* my = lds + l * LS;
  const int cstride = W * LS, boff = M * W * LS;
  const Cx<R>* tw = reinterpret_cast<const Cx<R>*>(a.tw);
  if (a.flags & kInK) {
    fft_lines<R, N, T, M, true>(v, my, cstride, boff, t, tw, a.tws);
    __syncthreads();
  }
  if (a.flags & kEnd) half_step<R, N, M>(v, a.h_end, a.pt, t, T, line, y, sp_off);
  if (a.frame[0] != nullptr && valid) {
#pragma unroll
    for (int c = 0; c < M; ++c) {
      Cx<R>* dst = reinterpret_cast<Cx<R>*>(a.frame[c]) + line * N;
#pragma unroll
      for (int m = 0; m < kE; ++m) dst[t + m * T] = v[c][m];
    }
  }
  if (a.flags & kStart) half_step<R, N, M>(v, a.h_start, a.pt, t, T, line, y, sp_off);
  if (a.flags & kOutK) fft_lines<R, N, T, M, false>(v, my, cstride, boff, t, tw, a.tws);
  if (valid) {
#pragma unroll
    for (int c = 0; c < M; ++c) {
      Cx<R>* dst = reinterpret_cast<Cx<R>*>(a.u[c]) + line * N;
#pragma unroll
      for (int m = 0; m < kE; ++m) dst[t + m * T] = v[c][m];
    }
  }
}

// ------------------------------------------------------------------ y-pass
struct YPassArgs {
  void* u[kMaxM];
  int64_t nx;         // dim-0 extent (columns)
  int64_t ncolblk;    // nx / C
  int64_t nblocks;    // ncolblk * batch
  int xcd_remap;      // 1: spread consecutive column blocks over one XCD
  int ed_form;
  const void* ed;     // exp(-i dt D) / (nx ny), nx*ny entries of 1, M or M*M values
  double scale;       // 1/(nx ny) when ed_form == 0
  const void* tw;
  int tws;
};

template <typename R, int N, int M, int C>
__global__ void __launch_bounds__(C * (N / kE)) y_pass_kernel(YPassArgs a) {
  constexpr int T = N / kE;
  constexpr int LS = lds_line_stride(N);
  extern __shared__ unsigned char smem_raw[];
  Cx<R>* lds = reinterpret_cast<Cx<R>*>(smem_raw);
  int64_t blk = blockIdx.x;
  if (a.xcd_remap) {
    // Workgroups are dispatched round-robin over 8 XCDs; give each XCD a
    // contiguous range of column blocks so neighbouring columns that share
    // cache lines meet in one L2.
    const int64_t per = a.nblocks / 8;
    blk = (blk % 8) * per + blk / 8;
  }
  const int64_t b = blk / a.ncolblk;
  const int64_t x0 = (blk % a.ncolblk) * C;
  const int tid = threadIdx.x;
  const int cx = tid % C, t = tid / C;
  const int64_t nx = a.nx;
  const int64_t plane = nx * N;
  const int64_t off = b * plane + x0 + cx;
  Cx<R> v[M][kE];
#pragma unroll
  for (int c = 0; c < M; ++c) {
    const Cx<R>* src = reinterpret_cast<const Cx<R>*>(a.u[c]) + off;
#pragma unroll
    for (int m = 0; m < kE; ++m) v[c][m] = src[(int64_t)(t + m * T) * nx];
  }
  Cx<R>* my = lds + cx * LS;
  const int cstride = C * LS, boff = M * C * LS;
  const Cx<R>* tw = reinterpret_cast<const Cx<R>*>(a.tw);
  fft_lines<R, N, T, M, false>(v, my, cstride, boff, t, tw, a.tws);
#pragma unroll
  for (int m = 0; m < kE; ++m) {
    Cx<R> w[M];
#pragma unroll
    for (int c = 0; c < M; ++c) w[c] = v[c][m];
    apply_ed<R, M>(w, a.ed_form, a.ed, (int64_t)(t + m * T) * nx + x0 + cx, (R)a.scale);
#pragma unroll
    for (int c = 0; c < M; ++c) v[c][m] = w[c];
  }
  __syncthreads();
  fft_lines<R, N, T, M, true>(v, my, cstride, boff, t, tw, a.tws);
#pragma unroll
  for (int c = 0; c < M; ++c) {
    Cx<R>* dst = reinterpret_cast<Cx<R>*>(a.u[c]) + off;
#pragma unroll
    for (int m = 0; m < kE; ++m) dst[(int64_t)(t + m * T) * nx] = v[c][m];
  }
}

// ------------------------------------------------------------ 1D whole step
struct Step1DArgs {
  void* u[kMaxM];
  void* frame[kMaxM];   // optional copy of the field after the last step
  int64_t nlines;       // batch
  int64_t nsteps;       // Strang steps in this launch
  const double* env;    // complex envelope samples or null (envelope == 1)
  int64_t env_base;     // sample index of F_now for the first half-step
  double delta;         // dt/2
  uint64_t call0;       // Philox counter of the first half-step
  int ed_form;
  const void* ed;
  double scale;
  const void* tw;
  int tws;
  PointTerms pt;
};

template <typename R, int N, int M, int W>
__global__ void __launch_bounds__(W * (N / kE)) step1d_kernel(Step1DArgs a) {
  constexpr int T = N / kE;
  constexpr int LS = lds_line_stride(N);
  extern __shared__ unsigned char smem_raw[];
  Cx<R>* lds = reinterpret_cast<Cx<R>*>(smem_raw);
  const int tid = threadIdx.x;
  const int l = tid / T, t = tid % T;
  const int64_t line_raw = (int64_t)blockIdx.x * W + l;
  const bool valid = line_raw < a.nlines;
  const int64_t line = valid ? line_raw : a.nlines - 1;
  Cx<R> v[M][kE];
#pragma unroll
  for (int c = 0; c < M; ++c) {
    const Cx<R>* src = reinterpret_cast<const Cx<R>*>(a.u[c]) + line * N;
#pragma unroll
    for (int m = 0; m < kE; ++m) v[c][m] = src[t + m * T];
  }
  Cx<R>* my = lds + l * LS;
  const int cstride = W * LS, boff = M * W * LS;
  const Cx<R>* tw = reinterpret_cast<const Cx<R>*>(a.tw);
  const double q = 0.5 * a.delta;  // (dt/2)/2 multiplies F
  for (int64_t s = 0; s < a.nsteps; ++s) {
    HalfStep h;
    h.delta = a.delta;
    const int64_t e0 = a.env_base + 2 * s;
    double en[3][2] = {{1, 0}, {1, 0}, {1, 0}};
    if (a.env) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        en[i][0] = a.env[2 * (e0 + i)];
        en[i][1] = a.env[2 * (e0 + i) + 1];
      }
    }
    h.fnow[0] = q * en[0][0];
    h.fnow[1] = q * en[0][1];
    h.fnext[0] = q * en[1][0];
    h.fnext[1] = q * en[1][1];
    h.call = a.call0 + 2 * (uint64_t)s;
    half_step<R, N, M>(v, h, a.pt, t, T, line, 0, 0);
    fft_lines<R, N, T, M, false>(v, my, cstride, boff, t, tw, a.tws);
#pragma unroll
    for (int m = 0; m < kE; ++m) {
      Cx<R> w[M];
#pragma unroll
      for (int c = 0; c < M; ++c) w[c] = v[c][m];
      apply_ed<R, M>(w, a.ed_form, a.ed, t + m * T, (R)a.scale);
#pragma unroll
      for (int c = 0; c < M; ++c) v[c][m] = w[c];
    }
    __syncthreads();
    fft_lines<R, N, T, M, true>(v, my, cstride, boff, t, tw, a.tws);
    h.fnow[0] = q * en[1][0];
    h.fnow[1] = q * en[1][1];
    h.fnext[0] = q * en[2][0];
    h.fnext[1] = q * en[2][1];
    h.call += 1;
    half_step<R, N, M>(v, h, a.pt, t, T, line, 0, 0);
    __syncthreads();
  }
  if (valid) {
#pragma unroll
    for (int c = 0; c < M; ++c) {
      Cx<R>* dst = reinterpret_cast<Cx<R>*>(a.u[c]) + line * N;
#pragma unroll
      for (int m = 0; m < kE; ++m) dst[t + m * T] = v[c][m];
      if (a.frame[0] != nullptr) {
        Cx<R>* fr = reinterpret_cast<Cx<R>*>(a.frame[c]) + line * N;
#pragma unroll
        for (int m = 0; m < kE; ++m) fr[t + m * T] = v[c][m];
      }
    }
  }
}

// ---------------------------------------------------- dispersion exponential
// E(k) = scale * exp(-i dt D(k)) for every k point (get_exponential,
// src/misc.jl:12-20), FP64 arithmetic, stored in the field precision.  For
// the 2x2 form: exp(A) = e^tau [cosh(s) I + sinh(s)/s (A - tau I)] with
// A = -i dt D, tau = tr(A)/2, s^2 = ((a-d)/2)^2 + bc; Taylor series for |s|
// small.
struct EdArgs {
  const double* D;  // complex FP64 table, npts * per entries
  void* out;
  int64_t npts;
  int form;         // 1 scalar, 2 diag, 3 full
  int M;
  double dt;
  double scale;
};

__device__ __forceinline__ Cx<double> cexp_d(Cx<double> z) {
  double s, c;
  sincos(z.im, &s, &c);
  const double e = exp(z.re);
  return mk<double>(e * c, e * s);
}

__device__ __forceinline__ Cx<double> csqrt_d(Cx<double> z) {
  // principal square root, branch cut on the negative real axis
  const double r = hypot(z.re, z.im);
  if (r == 0.0) return mk<double>(0.0, 0.0);
  if (z.re >= 0.0) {
    const double a = sqrt(0.5 * (r + z.re));
    return mk<double>(a, z.im / (2.0 * a));
  }
  const double b = sqrt(0.5 * (r - z.re));
  const double bb = (z.im >= 0.0) ? b : -b;
  return mk<double>(z.im / (2.0 * bb), bb);
}

__device__ __forceinline__ Cx<double> cdiv_d(Cx<double> a, Cx<double> b) {
  const double d = b.re * b.re + b.im * b.im;
  return mk<double>((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}

__device__ __forceinline__ void expm2x2_d(const Cx<double> (&A)[4], Cx<double> (&E)[4]) {
  // A column-major: A[0]=a (0,0), A[1]=c (1,0), A[2]=b (0,1), A[3]=d (1,1)
  const Cx<double> a = A[0], c = A[1], b = A[2], d = A[3];
  const Cx<double> tau = scale(a + d, 0.5), h = scale(a - d, 0.5);
  const Cx<double> s2 = h * h + b * c;
  const Cx<double> s = csqrt_d(s2);
  Cx<double> ch, shs;
  if (hypot(s.re, s.im) < 1e-3) {
    // cosh = sum s2^k/(2k)!, sinh(s)/s = sum s2^k/(2k+1)!
    const double cc[6] = {1.0, 1.0 / 2, 1.0 / 24, 1.0 / 720, 1.0 / 40320, 1.0 / 3628800};
    const double sc[6] = {1.0, 1.0 / 6, 1.0 / 120, 1.0 / 5040, 1.0 / 362880, 1.0 / 39916800};
    ch = mk<double>(cc[5], 0.0);
    shs = mk<double>(sc[5], 0.0);
    for (int k = 4; k >= 0; --k) {
      ch = ch * s2 + mk<double>(cc[k], 0.0);
      shs = shs * s2 + mk<double>(sc[k], 0.0);
    }
  } else {
    const Cx<double> ep = cexp_d(s), em = cexp_d(mk<double>(-s.re, -s.im));
    ch = scale(ep + em, 0.5);
    shs = cdiv_d(scale(ep - em, 0.5), s);
  }
  const Cx<double> et = cexp_d(tau);
  E[0] = et * (ch + shs * h);
  E[3] = et * (ch - shs * h);
  E[2] = et * (shs * b);
  E[1] = et * (shs * c);
}

template <typename R>
__global__ void ed_table_kernel(EdArgs a) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.npts) return;
  Cx<R>* out = reinterpret_cast<Cx<R>*>(a.out);
  const Cx<double>* D = reinterpret_cast<const Cx<double>*>(a.D);
  if (a.form == 1 || a.form == 2) {
    const int per = (a.form == 1) ? 1 : a.M;
    for (int i = 0; i < per; ++i) {
      const Cx<double> z = D[k * per + i];
      // exp(-i dt z)
      const Cx<double> e = cexp_d(mk<double>(a.dt * z.im, -a.dt * z.re));
      out[k * per + i] = mk<R>((R)(e.re * a.scale), (R)(e.im * a.scale));
    }
  } else if (a.M == 1) {
    const Cx<double> z = D[k];
    const Cx<double> e = cexp_d(mk<double>(a.dt * z.im, -a.dt * z.re));
    out[k] = mk<R>((R)(e.re * a.scale), (R)(e.im * a.scale));
  } else {
    Cx<double> A[4], E[4];
    for (int i = 0; i < 4; ++i) {
      const Cx<double> z = D[k * 4 + i];
      A[i] = mk<double>(a.dt * z.im, -a.dt * z.re);
    }
    expm2x2_d(A, E);
    for (int i = 0; i < 4; ++i) out[k * 4 + i] = mk<R>((R)(E[i].re * a.scale), (R)(E[i].im * a.scale));
  }
}

// exp(-i tau V) planes (potential, get_exponential with dt/2), per component.
struct EvArgs {
  const double* V;  // complex table
  void* out[kMaxM];
  int64_t npts;
  int form;         // 1 scalar (shared by all components), 2 diag
  int M;
  double tau;
};

template <typename R>
__global__ void ev_table_kernel(EvArgs a) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.npts) return;
  const Cx<double>* V = reinterpret_cast<const Cx<double>*>(a.V);
  for (int c = 0; c < a.M; ++c) {
    const Cx<double> z = (a.form == 1) ? V[k] : V[k * a.M + c];
    const Cx<double> e = cexp_d(mk<double>(a.tau * z.im, -a.tau * z.re));
    reinterpret_cast<Cx<R>*>(a.out[c])[k] = mk<R>((R)e.re, (R)e.im);
  }
}

}  // namespace gpe
