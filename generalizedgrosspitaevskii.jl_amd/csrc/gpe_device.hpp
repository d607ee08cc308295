// Device building blocks for the gfx950 Strang-splitting kernels:
//   * Cx<R>        interleaved complex value (R = float or double)
//   * dft2/4/8     radix butterflies
//   * stockham     an LDS-staged Stockham autosort FFT of M lines of length N,
//                  each line spread over T threads that hold E = N/T
//                  values in registers at positions t + m*T (cyclic layout)
//   * philox4x32_10 and the Box-Muller map used for the Wiener increments
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpe {

template <typename R>
struct alignas(2 * sizeof(R)) Cx {
  R re, im;
};

template <typename R>
__device__ __forceinline__ Cx<R> mk(R a, R b) {
  Cx<R> z;
  z.re = a;
  z.im = b;
  return z;
}
template <typename R>
__device__ __forceinline__ Cx<R> operator+(Cx<R> a, Cx<R> b) {
  return mk<R>(a.re + b.re, a.im + b.im);
}
template <typename R>
__device__ __forceinline__ Cx<R> operator-(Cx<R> a, Cx<R> b) {
  return mk<R>(a.re - b.re, a.im - b.im);
}
// Products of two terms are written with an explicit fused form.  Left to
// floating-point contraction, x*y - z*w may fuse either product, and the
// compiler can pick differently for the E values a thread holds once their
// code is interleaved; the rounding of equal inputs then depends on the
// value's position, which breaks exact symmetries of the field (a
// homogeneous state stays homogeneous in exact arithmetic and must stay so
// here, or split-step parametric modes near dt k^2/2 = n pi get a seed).
template <typename R>
__device__ __forceinline__ R fmx(R a, R b, R c) {
  return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fmx<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
template <typename R>
__device__ __forceinline__ Cx<R> operator*(Cx<R> a, Cx<R> b) {
  return mk<R>(fmx<R>(a.re, b.re, -(a.im * b.im)), fmx<R>(a.re, b.im, a.im * b.re));
}
template <typename R>
__device__ __forceinline__ Cx<R> scale(Cx<R> a, R s) {
  return mk<R>(a.re * s, a.im * s);
}
template <typename R>
__device__ __forceinline__ Cx<R> conjx(Cx<R> a) {
  return mk<R>(a.re, -a.im);
}
// multiply by -i (forward) or +i (inverse)
template <typename R, bool INV>
__device__ __forceinline__ Cx<R> mul_mi(Cx<R> a) {
  return INV ? mk<R>(-a.im, a.re) : mk<R>(a.im, -a.re);
}
template <typename R>
__device__ __forceinline__ R abs2(Cx<R> a) {
  return fmx<R>(a.re, a.re, a.im * a.im);
}

// ---------------------------------------------------------------- butterflies
template <typename R, bool INV>
__device__ __forceinline__ void dft2(Cx<R>& a0, Cx<R>& a1) {
  Cx<R> t = a0;
  a0 = t + a1;
  a1 = t - a1;
}

template <typename R, bool INV>
__device__ __forceinline__ void dft4(Cx<R>& a0, Cx<R>& a1, Cx<R>& a2, Cx<R>& a3) {
  Cx<R> s02 = a0 + a2, d02 = a0 - a2, s13 = a1 + a3, d13 = mul_mi<R, INV>(a1 - a3);
  a0 = s02 + s13;
  a2 = s02 - s13;
  a1 = d02 + d13;
  a3 = d02 - d13;
}

// exp(-+ i pi/4 * k) products used by the radix-8 butterfly
template <typename R, bool INV>
__device__ __forceinline__ Cx<R> w8_1(Cx<R> a) {
  const R h = (R)0.70710678118654752440084436210484903928;
  return INV ? mk<R>((a.re - a.im) * h, (a.re + a.im) * h) : mk<R>((a.re + a.im) * h, (a.im - a.re) * h);
}
template <typename R, bool INV>
__device__ __forceinline__ Cx<R> w8_3(Cx<R> a) {
  const R h = (R)0.70710678118654752440084436210484903928;
  return INV ? mk<R>((-a.re - a.im) * h, (a.re - a.im) * h) : mk<R>((a.im - a.re) * h, (-a.re - a.im) * h);
}

template <typename R, bool INV>
__device__ __forceinline__ void dft8(Cx<R> (&a)[8]) {
  Cx<R> b0 = a[0] + a[4], b1 = a[1] + a[5], b2 = a[2] + a[6], b3 = a[3] + a[7];
  Cx<R> c0 = a[0] - a[4];
  Cx<R> c1 = w8_1<R, INV>(a[1] - a[5]);
  Cx<R> c2 = mul_mi<R, INV>(a[2] - a[6]);
  Cx<R> c3 = w8_3<R, INV>(a[3] - a[7]);
  dft4<R, INV>(b0, b1, b2, b3);
  dft4<R, INV>(c0, c1, c2, c3);
  a[0] = b0;
  a[1] = c0;
  a[2] = b1;
  a[3] = c1;
  a[4] = b2;
  a[5] = c2;
  a[6] = b3;
  a[7] = c3;
}

template <typename R, bool INV, int RADIX>
__device__ __forceinline__ void dft(Cx<R> (&a)[RADIX]) {
  if constexpr (RADIX == 8) {
    dft8<R, INV>(a);
  } else if constexpr (RADIX == 4) {
    dft4<R, INV>(a[0], a[1], a[2], a[3]);
  } else if constexpr (RADIX == 2) {
    dft2<R, INV>(a[0], a[1]);
  } else {
    static_assert(RADIX == 8 || RADIX == 4 || RADIX == 2, "radix");
  }
}

// ------------------------------------------------------------ Stockham engine
// Values per thread; every supported N >= 16 uses E = 8.
constexpr int kE = 8;

// LDS row stride (in elements) of one line of length N: one pad slot per 8
// elements plus one, so that strided butterfly writes spread over banks.
__host__ __device__ constexpr int lds_line_stride(int N) { return N + N / 8 + 1; }
__device__ __forceinline__ int lds_pad(int p) { return p + (p >> 3); }

template <int N, int NS>
struct PassRadix {
  static constexpr int rem = N / NS;
  static constexpr int value = (rem % 8 == 0) ? 8 : rem;
};

// Twiddles a thread needs over all passes of a length-N transform (passes
// with NS > 1 carry R-1 twiddles per butterfly), and the offset of pass NS.
__host__ __device__ constexpr int tw_count_before(int N, int NS) {
  int ns = 1, c = 0;
  while (ns < N && ns < NS) {
    const int rem = N / ns;
    const int r = (rem % 8 == 0) ? 8 : rem;
    if (ns > 1) c += (kE / r) * (r - 1);
    ns *= r;
  }
  return c;
}
__host__ __device__ constexpr int tw_count(int N) { return tw_count_before(N, N) > 0 ? tw_count_before(N, N) : 1; }

// LDS exchanges of one length-N transform (passes - 1).  Pass p writes buffer
// (BUF0 + p) mod 2, so the last exchange uses buffer (BUF0 + exchanges - 1) mod 2.
__host__ __device__ constexpr int fft_exchanges(int N) {
  int ns = 1, p = 0;
  while (ns < N) {
    const int rem = N / ns;
    ns *= (rem % 8 == 0) ? 8 : rem;
    ++p;
  }
  return p - 1;
}

// Per-thread twiddle registers of one transform length, loaded once (at
// kernel start, together with the field) from the FP64-accurate table
// tw[m] = exp(-2 pi i m / NT), NT = N * tws.  The inverse transform uses the
// conjugates.
template <typename R, int N>
struct TwReg {
  Cx<R> w[tw_count(N)];
};

template <typename R, int N, int T, int NS>
__device__ __forceinline__ void load_twiddles_pass(TwReg<R, N>& tr, int t, const Cx<R>* __restrict__ tw, int tws) {
  constexpr int RADIX = PassRadix<N, NS>::value;
  constexpr int B = kE / RADIX;
  if constexpr (NS > 1) {
    constexpr int off = tw_count_before(N, NS);
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int k = (t + b * T) & (NS - 1);
      const int base = k * (N / (NS * RADIX)) * tws;
#pragma unroll
      for (int r = 1; r < RADIX; ++r) tr.w[off + b * (RADIX - 1) + r - 1] = tw[r * base];
    }
  }
  if constexpr (NS * RADIX < N) load_twiddles_pass<R, N, T, NS * RADIX>(tr, t, tw, tws);
}

template <typename R, int N, int T>
__device__ __forceinline__ void load_twiddles(TwReg<R, N>& tr, int t, const Cx<R>* __restrict__ tw, int tws) {
  load_twiddles_pass<R, N, T, 1>(tr, t, tw, tws);
}

// Twiddles read from the table at each pass instead of held in registers.
// Kernels of more than 512 threads per workgroup run with at most 128 VGPRs;
// there the ~40 registers of a long transform's twiddles made the field spill
// to scratch memory.  Same table entries and index as load_twiddles_pass, so
// the arithmetic is identical.
template <typename R, int N>
struct TwLazy {
  const Cx<R>* __restrict__ w;
  int tws;
};

template <typename R, int N, int T>
__device__ __forceinline__ void load_twiddles(TwLazy<R, N>& tr, int, const Cx<R>* __restrict__ tw, int tws) {
  tr.w = tw;
  tr.tws = tws;
}

// Twiddle r (1..RADIX-1) of butterfly b in the pass of span NS.
template <typename R, int N, int T, int NS, int RADIX>
__device__ __forceinline__ Cx<R> tw_at(const TwReg<R, N>& tr, int, int b, int r) {
  return tr.w[tw_count_before(N, NS) + b * (RADIX - 1) + r - 1];
}
template <typename R, int N, int T, int NS, int RADIX>
__device__ __forceinline__ Cx<R> tw_at(const TwLazy<R, N>& tr, int t, int b, int r) {
  const int k = (t + b * T) & (NS - 1);
  return tr.w[r * (k * (N / (NS * RADIX)) * tr.tws)];
}

// Twiddle storage of a kernel of `threads` threads per workgroup.
template <bool LAZY, typename R, int N>
struct TwPick {
  using type = TwReg<R, N>;
};
template <typename R, int N>
struct TwPick<true, R, N> {
  using type = TwLazy<R, N>;
};
template <typename R, int N, int THREADS>
using TwFor = typename TwPick<(THREADS > 512), R, N>::type;

// Orders the LDS writes of a line's threads before their reads.  When the T
// threads of every line lie inside one wavefront (LOCAL), a wavefront-scope
// fence around a wave barrier suffices: a wavefront's LDS instructions execute
// in order, so only the compiler must be kept from moving reads above writes.
// Otherwise the whole workgroup synchronises.
template <bool LOCAL>
__device__ __forceinline__ void line_sync() {
  if constexpr (LOCAL) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One Stockham pass on M bundled lines followed (unless last) by an LDS
// exchange into the cyclic layout.  `lds` points at this thread-group's lines;
// line c of the bundle lives at lds + c*lstride (+ lds_pad(p)).  Two LDS
// buffers (lds, lds + buf_off) alternate between passes so one barrier per
// pass suffices.
template <typename R, int N, int T, int M, bool INV, int NS, int BUF, bool LOCAL, typename TW>
__device__ __forceinline__ void stockham_pass(Cx<R> (&v)[M][kE], Cx<R>* lds, int lstride, int buf_off, int t,
                                              const TW& tr) {
  constexpr int RADIX = PassRadix<N, NS>::value;
  constexpr int B = kE / RADIX;
  constexpr bool LAST = (NS * RADIX == N);
  static_assert(kE % RADIX == 0, "radix must divide the per-thread count");
  Cx<R>* buf = lds + BUF * buf_off;
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int j = t + b * T;
    const int k = j & (NS - 1);
#pragma unroll
    for (int c = 0; c < M; ++c) {
      Cx<R> a[RADIX];
#pragma unroll
      for (int r = 0; r < RADIX; ++r) a[r] = v[c][b + r * B];
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < RADIX; ++r) {
          const Cx<R> w = tw_at<R, N, T, NS, RADIX>(tr, t, b, r);
          a[r] = a[r] * (INV ? conjx(w) : w);
        }
      }
      dft<R, INV, RADIX>(a);
      if constexpr (LAST) {
#pragma unroll
        for (int r = 0; r < RADIX; ++r) v[c][b + r * B] = a[r];
      } else {
        const int idxD = (j / NS) * NS * RADIX + k;
#pragma unroll
        for (int r = 0; r < RADIX; ++r) buf[c * lstride + lds_pad(idxD + r * NS)] = a[r];
      }
    }
  }
  if constexpr (!LAST) {
    line_sync<LOCAL>();
#pragma unroll
    for (int c = 0; c < M; ++c)
#pragma unroll
      for (int m = 0; m < kE; ++m) v[c][m] = buf[c * lstride + lds_pad(t + m * T)];
    stockham_pass<R, N, T, M, INV, NS * RADIX, 1 - BUF, LOCAL, TW>(v, lds, lstride, buf_off, t, tr);
  }
}

// Unnormalised DFT (INV=false: exp(-2 pi i jk/N); INV=true: exp(+2 pi i jk/N))
// of M bundled lines, input and output in the cyclic register layout.  LOCAL:
// the caller guarantees each line's threads share one wavefront.
template <typename R, int N, int T, int M, bool INV, bool LOCAL = false, int BUF0 = 0, typename TW = TwReg<R, N>>
__device__ __forceinline__ void fft_lines(Cx<R> (&v)[M][kE], Cx<R>* lds, int lstride, int buf_off, int t,
                                          const TW& tr) {
  static_assert(N / T == kE, "E = N/T must be 8");
  static_assert(!LOCAL || (T <= 64 && 64 % T == 0), "a wave-local line must fit in one wavefront");
  stockham_pass<R, N, T, M, INV, 1, BUF0, LOCAL, TW>(v, lds, lstride, buf_off, t, tr);
}

// ------------------------------------------------------------------- Philox
struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    // one 32x32 -> 64-bit product per multiplier (a single v_mad_u64_u32)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    U4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Box-Muller: (wa, wb) -> complex normal with E|z|^2 = 1:
// z = sqrt(-log u_a) exp(2 pi i u_b), u = (w + 1/2) 2^-32.  The FP64 form is
// kept out of line (library log and sincospi); the FP32 form uses the
// hardware log2 and the revolution-based sin/cos of the ALU.
template <typename R>
__device__ Cx<R> box_muller(uint32_t wa, uint32_t wb);

template <>
__device__ __noinline__ Cx<double> box_muller<double>(uint32_t wa, uint32_t wb) {
  const double ua = ((double)wa + 0.5) * 2.3283064365386962890625e-10;
  const double ub = ((double)wb + 0.5) * 2.3283064365386962890625e-10;
  const double r = sqrt(-log(ua));
  double s, c;
  sincospi(2.0 * ub, &s, &c);
  return mk<double>(r * c, r * s);
}

template <>
__device__ __forceinline__ Cx<float> box_muller<float>(uint32_t wa, uint32_t wb) {
  const float ua = ((float)(wa >> 8) + 0.5f) * 5.9604644775390625e-08f;
  const float ub = ((float)(wb >> 8) + 0.5f) * 5.9604644775390625e-08f;
  const float r = sqrtf(-__logf(ua));
  return mk<float>(r * __builtin_amdgcn_cosf(ub), r * __builtin_amdgcn_sinf(ub));
}

}  // namespace gpe
