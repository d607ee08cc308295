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
//   3D:  x-pass as in 2D; y-pass FFT_y only; z-pass FFT_z ->
//        exp(-i dt D(kx,ky,kz)) / (Nx Ny Nz) -> IFFT_z; y-pass IFFT_y only.
//   1D:  whole-step kernel: every line stays in registers for a block of
//        steps (half, FFT, exp(-i dt D), IFFT, half, repeated).
//
// Every pass updates the field in place: a workgroup reads exactly the lines
// it writes.  The real-space half-step is muladd_kernel! of
// src/kernels.jl:37-54 for the descriptor forms
//     G_c(u) = c_c + sum_j g_cj |u_j|^2,  eta_c constant,
//     F_c(r, tau) = profile_c(r) * envelope(tau), or F_c sampled on the grid
//     at every pump time when it does not factor.
#pragma once
#include "gpe_device.hpp"
#include "gpe_math.hpp"

namespace gpe {

constexpr int kMaxM = 4;

// Parameters of one real-space half-step (src/strang_splitting.jl:78-84).
struct HalfStep {
  double delta;     // the half-step's dt (= dt/2)
  double fnow[2];   // (delta/2) * envelope(previous pump time)   (complex)
  double fnext[2];  // (delta/2) * envelope(this pump time)       (complex)
  uint64_t call;    // Philox draw counter of this half-step
  int64_t jnow, jnext;  // pump-time sample indices of F_now / F_next (sampled pump)
};

// Per-point terms shared by every real-space half-step, for up to MM
// components.  Kernels take PointTerms (kMaxM) as an argument; a kernel that
// builds the terms in registers (the sweep, one component) uses MM = 1 so the
// struct stays small enough to live in registers.
template <int MM>
struct PointTermsT {
  int gmask;   // bit c: component c has a nonlinearity G_c = c_c + sum_j g[c][j] |u_j|^2
  int g_cplx;  // some coefficient of G has an imaginary part (gain/loss)
  int g_coupled;  // some G_c depends on |u_j|^2 with j != c
  double g[MM][MM][2];
  double c[MM][2];
  int pump_kind;                 // 0 none, 1 table, 2 separable, 3 sampled at every pump time
  const void* pump_a[MM];        // table: plane of the spatial grid; separable: dim-0 factor;
                                 // sampled: plane j of the spatial grid at pump_stride * j
  const void* pump_b[MM];        // separable: dim-1 factor (2D), else null
  int64_t pump_stride;           // sampled: points per sample plane
  int pot_kind;                  // 0 none, else exp(-i dt/2 V) planes per component
  const void* pot[MM];
  int noise_kind;                // 0 none, 1 complex, 2 real
  double eta[MM];
  uint32_t seed_lo, seed_hi;
};
using PointTerms = PointTermsT<kMaxM>;

template <typename R>
__device__ __forceinline__ Cx<R> ldc(const void* p, int64_t i) {
  return reinterpret_cast<const Cx<R>*>(p)[i];
}
template <typename R>
__device__ __forceinline__ Cx<R> cxd(const double (&a)[2]) {
  return mk<R>((R)a[0], (R)a[1]);
}

// Each thread works on ONE component: the M components of a line (or a
// column) are held by M groups of threads of the same workgroup, and values
// cross between components through LDS only where the arithmetic couples
// them (|u_j|^2 in G_c for j != c; the off-diagonal of exp(-i dt D)).
// Workgroups are laid out component-major, so every wavefront belongs to a
// single component and component-dependent branches are wave-uniform.

// Entry c (< M) of a per-component argument array.  c is uniform over a
// wavefront but not known to the compiler as such; a select chain keeps the
// kernel arguments out of private memory.
template <int M, typename T, int K>
__device__ __forceinline__ T pick(const T (&a)[K], int c) {
  static_assert(M <= K, "more components than entries");
  T r = a[0];
#pragma unroll
  for (int j = 1; j < M; ++j) r = (c == j) ? a[j] : r;
  return r;
}

// The Wiener increment -i sqrt(delta) eta_c xi_c of the thread's E values.
// Element m and m + E/2 (x and x + N/2 on the line) share one Philox draw:
// ctr = (P mod 2^32, c + 16 (P >> 32), call mod 2^32, call >> 32) with
// P = pbase + x, key = seed; words (x,y) feed x, (z,w) feed x + N/2.
// pbase is the flat index of the line's first point over 2 (line * N/2 for
// equal lines; offset/2 for concatenated instances of uneven length).
template <typename R, int N, typename PT>
__device__ __forceinline__ void noise_terms(Cx<R> (&nz)[kE], const HalfStep& hs, const PT& pt, int c, R eta,
                                            int t, int T, int64_t pbase) {
  const R a = (R)sqrt(hs.delta) * eta;
#pragma unroll
  for (int m = 0; m < kE / 2; ++m) {
    const uint64_t P = (uint64_t)pbase + (uint64_t)(t + m * T);
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
    nz[m] = mk<R>(a * z0.im, -a * z0.re);
    nz[m + kE / 2] = mk<R>(a * z1.im, -a * z1.re);
  }
}

// Per-thread constants of the real-space half-step for component c: the
// pump profile and exp(-i dt/2 V) at the thread's points (loaded at kernel
// start, so their latency overlaps the field loads) and the nonlinearity
// coefficients of row c in the field precision.
//
// LAZY (kernels of more than 512 threads per workgroup, which run with at most
// 128 VGPRs): the pump profile and exp(-i dt/2 V) are read where the half-step
// uses them instead of held in 16 registers per value pair (same loads and
// products, so the same values).
template <typename R, int M, bool LAZY = false>
struct PointCacheT {
  Cx<R> pump[LAZY ? 1 : kE];
  Cx<R> pot[LAZY ? 1 : kE];
  uint32_t yline;     // the line's index over the spatial grid (sampled pump: plane offset yline * N)
  bool pumped, has_g;
  R cr, ci, gr[M], gi[M], eta;  // row c of g
  // g_cc for M > 2, taken from the kernel arguments: selecting it from gr[] at
  // use let the compiler index that array by c, which put the whole cache in
  // scratch memory (for M <= 2 the select stays a select and costs no registers)
  R gdr, gdi;
};

// g_cc (real or imaginary part) of the row a thread holds
template <typename R, int M, bool IM, bool LZ>
__device__ __forceinline__ R gdiag(const PointCacheT<R, M, LZ>& pc, int c) {
  if constexpr (M > 2) {
    return IM ? pc.gdi : pc.gdr;
  } else {
    const R(&row)[M] = IM ? pc.gi : pc.gr;
    return (M == 2 && c == 1) ? row[M - 1] : row[0];
  }
}

template <typename R, int M>
using PointCache = PointCacheT<R, M, false>;

// Pump profile and exp(-i dt/2 V) at value m of a thread of component c (the
// line's spatial offset is yline * N: y * N in the 2D/3D row pass, 0 in 1D).
template <typename R, int N, int M, bool LZ, typename PT>
__device__ __forceinline__ Cx<R> pump_of(const PT& pt, const PointCacheT<R, M, LZ>& pc, int c, int m, int t, int T) {
  if constexpr (!LZ) {
    return pc.pump[m];
  } else {
    const int x = t + m * T;
    const void* pa = pick<M>(pt.pump_a, c);
    if (pt.pump_kind == 1) return ldc<R>(pa, (int64_t)pc.yline * N + x);
    const void* pbp = pick<M>(pt.pump_b, c);
    const Cx<R> pb = (pt.pump_kind == 2 && pbp) ? ldc<R>(pbp, pc.yline) : mk<R>((R)1, (R)0);
    return ldc<R>(pa, x) * pb;
  }
}
template <typename R, int N, int M, bool LZ, typename PT>
__device__ __forceinline__ Cx<R> pot_of(const PT& pt, const PointCacheT<R, M, LZ>& pc, int c, int m, int t, int T) {
  if constexpr (!LZ)
    return pc.pot[m];
  else
    return ldc<R>(pick<M>(pt.pot, c), (int64_t)pc.yline * N + t + m * T);
}

template <typename R, int M, typename PT, bool LZ>
__device__ __forceinline__ void load_point_cache(PointCacheT<R, M, LZ>& pc, const PT& pt, int c, int t, int T,
                                                 int64_t y, int64_t sp_off) {
  const void* pa = pick<M>(pt.pump_a, c);
  pc.pumped = pt.pump_kind != 0 && pa != nullptr;
  pc.yline = (uint32_t)y;
  if constexpr (!LZ) {
    if (pc.pumped && pt.pump_kind != 3) {
      const void* pbp = pick<M>(pt.pump_b, c);
      const Cx<R> pb = (pt.pump_kind == 2 && pbp) ? ldc<R>(pbp, y) : mk<R>((R)1, (R)0);
#pragma unroll
      for (int m = 0; m < kE; ++m) {
        const int x = t + m * T;
        pc.pump[m] = (pt.pump_kind == 1) ? ldc<R>(pa, sp_off + x) : ldc<R>(pa, x) * pb;
      }
    }
    if (pt.pot_kind) {
      const void* pp = pick<M>(pt.pot, c);
#pragma unroll
      for (int m = 0; m < kE; ++m) pc.pot[m] = ldc<R>(pp, sp_off + t + m * T);
    }
  }
  pc.has_g = (pt.gmask >> c) & 1;
  double cr = pt.c[0][0], ci = pt.c[0][1];
  double gr[M], gi[M];
#pragma unroll
  for (int j = 0; j < M; ++j) {
    gr[j] = pt.g[0][j][0];
    gi[j] = pt.g[0][j][1];
  }
#pragma unroll
  for (int i = 1; i < M; ++i) {
    const bool me = c == i;
    cr = me ? pt.c[i][0] : cr;
    ci = me ? pt.c[i][1] : ci;
#pragma unroll
    for (int j = 0; j < M; ++j) {
      gr[j] = me ? pt.g[i][j][0] : gr[j];
      gi[j] = me ? pt.g[i][j][1] : gi[j];
    }
  }
  pc.cr = (R)cr;
  pc.ci = (R)ci;
#pragma unroll
  for (int j = 0; j < M; ++j) {
    pc.gr[j] = (R)gr[j];
    pc.gi[j] = (R)gi[j];
  }
  if constexpr (M > 2) {
    double dr = pt.g[0][0][0], di = pt.g[0][0][1];
#pragma unroll
    for (int i = 1; i < M; ++i) {
      dr = (c == i) ? pt.g[i][i][0] : dr;
      di = (c == i) ? pt.g[i][i][1] : di;
    }
    pc.gdr = (R)dr;
    pc.gdi = (R)di;
  }
  pc.eta = (R)pick<M>(pt.eta, c);
}

// Sampled pump: the line's first point in sample plane 0 of component c
// (formed where it is used, so no pointer stays live in the kernels).
template <typename R, int N, int M, typename PT, bool LZ>
__device__ __forceinline__ const Cx<R>* sampled_line(const PT& pt, const PointCacheT<R, M, LZ>& pc, int c) {
  return reinterpret_cast<const Cx<R>*>(pick<M>(pt.pump_a, c)) + (int64_t)pc.yline * N;
}

// Real-space half-step (muladd_kernel!, src/kernels.jl:37-54) on the E values
// of component c a thread holds; element m sits at dim-0 index x = t + m*T of
// a line whose first Philox pair index is `pbase`.  When G couples the
// components, |u_j|^2 of the other components comes through the LDS slice
// `a2` (M rows of N values for this thread's line); every thread of the
// workgroup takes the same branch, so the barriers are uniform.
template <typename R, int N, int M, bool NOISE, typename PT, bool LZ>
__device__ __forceinline__ void half_step(Cx<R> (&v)[kE], int c, const HalfStep& hs, const PT& pt,
                                          const PointCacheT<R, M, LZ>& pc, int t, int T, int64_t pbase, R* a2) {
  const R delta = (R)hs.delta;
  R a2j[M][kE];
  if (pt.gmask) {
#pragma unroll
    for (int m = 0; m < kE; ++m) a2j[0][m] = abs2(v[m]);
    if (M > 1 && pt.g_coupled) {
      __syncthreads();
#pragma unroll
      for (int m = 0; m < kE; ++m) a2[c * N + t + m * T] = a2j[0][m];
      __syncthreads();
#pragma unroll
      for (int j = 0; j < M; ++j)
#pragma unroll
        for (int m = 0; m < kE; ++m) a2j[j][m] = a2[j * N + t + m * T];
      __syncthreads();
    }
  }
  if constexpr (NOISE) {
    // Stochastic kernels are bound by the Philox/Box-Muller arithmetic and
    // need their occupancy: one compact loop over the values (fewer live
    // registers and less code than the interleaved form below).
    Cx<R> nz[kE];
    noise_terms<R, N>(nz, hs, pt, c, pc.eta, t, T, pbase);
    const Cx<R> fnow = cxd<R>(hs.fnow), fnext = cxd<R>(hs.fnext);
    // a sampled pump (F at every pump time) is added around the loop below,
    // which keeps the profile forms' code and registers unchanged; the
    // envelope scalars are (delta/2) exactly for it
    const bool sampled = pc.pumped && pt.pump_kind == 3;
    const bool profile = pc.pumped && pt.pump_kind != 3;
    if (sampled) {
      const Cx<R>* ptab = sampled_line<R, N, M>(pt, pc, c);
#pragma unroll
      for (int m = 0; m < kE; ++m) v[m] = v[m] + ptab[hs.jnow * pt.pump_stride + t + m * T] * fnow;
    }
#pragma unroll
    for (int m = 0; m < kE; ++m) {
      Cx<R> e = mk<R>((R)1, (R)0);
      if (pc.has_g) {
        R gr = pc.cr, gi = pc.ci;
        if (M > 1 && pt.g_coupled) {
#pragma unroll
          for (int j = 0; j < M; ++j) {
            gr += pc.gr[j] * a2j[j][m];
            gi += pc.gi[j] * a2j[j][m];
          }
        } else {
          gr += gdiag<R, M, false>(pc, c) * a2j[0][m];
          gi += gdiag<R, M, true>(pc, c) * a2j[0][m];
        }
        R sn, co;
        fast_sincos(delta * gr, sn, co);
        const R mag = pt.g_cplx ? (R)exp(delta * gi) : (R)1;
        e = mk<R>(mag * co, -mag * sn);
      }
      if (pt.pot_kind) e = e * pot_of<R, N, M>(pt, pc, c, m, t, T);
      Cx<R> w = v[m];
      if (profile) w = w + pump_of<R, N, M>(pt, pc, c, m, t, T) * fnow;
      w = e * w;
      if (profile) w = w + pump_of<R, N, M>(pt, pc, c, m, t, T) * fnext;
      v[m] = sampled ? w : w + nz[m];
    }
    if (sampled) {
      const Cx<R>* ptab = sampled_line<R, N, M>(pt, pc, c);
#pragma unroll
      for (int m = 0; m < kE; ++m) v[m] = (v[m] + ptab[hs.jnext * pt.pump_stride + t + m * T] * fnext) + nz[m];
    }
    return;
  }
  // The branches below depend only on wave-uniform terms; each guards a
  // straight loop over the E values so their arithmetic interleaves.
  //
  // e = cis(-delta G(u)) exp(-i delta V) from the pre-update fields
  // (src/kernels.jl:40-44); has_e false means e == 1 exactly.
  Cx<R> e[kE];
  const bool has_e = pc.has_g || pt.pot_kind;
  if (pc.has_g) {
    R arg[kE], gim[kE];
#pragma unroll
    for (int m = 0; m < kE; ++m) {
      R gr = pc.cr, gi = pc.ci;
      if (M > 1 && pt.g_coupled) {
#pragma unroll
        for (int j = 0; j < M; ++j) {
          gr += pc.gr[j] * a2j[j][m];
          gi += pc.gi[j] * a2j[j][m];
        }
      } else {
        gr += gdiag<R, M, false>(pc, c) * a2j[0][m];
        gi += gdiag<R, M, true>(pc, c) * a2j[0][m];
      }
      arg[m] = delta * gr;
      gim[m] = gi;
    }
    R s[kE], co[kE];
    fast_sincos_n<R, kE>(arg, s, co);
    // exp(-i delta (gr + i gi)) = exp(delta gi) (cos(delta gr) - i sin(delta gr))
    if (pt.g_cplx) {
#pragma unroll
      for (int m = 0; m < kE; ++m) {
        const R mag = exp(delta * gim[m]);
        e[m] = mk<R>(mag * co[m], -mag * s[m]);
      }
    } else {
#pragma unroll
      for (int m = 0; m < kE; ++m) e[m] = mk<R>(co[m], -s[m]);
    }
    if (pt.pot_kind) {
#pragma unroll
      for (int m = 0; m < kE; ++m) e[m] = e[m] * pot_of<R, N, M>(pt, pc, c, m, t, T);
    }
  } else if (pt.pot_kind) {
#pragma unroll
    for (int m = 0; m < kE; ++m) e[m] = pot_of<R, N, M>(pt, pc, c, m, t, T);
  }
  // e (u + dt/2 F_now) + dt/2 F_next + noise   (src/kernels.jl:47-48)
  if (pc.pumped && pt.pump_kind == 3) {
    const Cx<R> fnow = cxd<R>(hs.fnow), fnext = cxd<R>(hs.fnext);
    const Cx<R>* ptab = sampled_line<R, N, M>(pt, pc, c);
#pragma unroll
    for (int m = 0; m < kE; ++m) {
      const Cx<R> w = v[m] + ptab[hs.jnow * pt.pump_stride + t + m * T] * fnow;
      v[m] = (has_e ? e[m] * w : w) + ptab[hs.jnext * pt.pump_stride + t + m * T] * fnext;
    }
  } else if (pc.pumped) {
    const Cx<R> fnow = cxd<R>(hs.fnow), fnext = cxd<R>(hs.fnext);
    if (has_e) {
#pragma unroll
      for (int m = 0; m < kE; ++m) {
        const Cx<R> f = pump_of<R, N, M>(pt, pc, c, m, t, T);
        v[m] = e[m] * (v[m] + f * fnow) + f * fnext;
      }
    } else {
#pragma unroll
      for (int m = 0; m < kE; ++m) {
        const Cx<R> f = pump_of<R, N, M>(pt, pc, c, m, t, T);
        v[m] = (v[m] + f * fnow) + f * fnext;
      }
    }
  } else if (has_e) {
#pragma unroll
    for (int m = 0; m < kE; ++m) v[m] = e[m] * v[m];
  }
}

// Two adjacent real-space half-steps with nothing between them (the one that
// ends step s and the one that starts step s+1).  For a component whose
// half-step only turns the phase -- no pump, noise or potential, G_c real and
// a function of |u_c| alone -- the first leaves |u_c| unchanged, so the pair is
// cis(-2 delta G(|u|^2)) u: one sine/cosine per value instead of two, equal to
// the two half-steps up to rounding.  The test is wave-uniform, and without
// coupling half_step has no barrier, so the components' wavefronts may take
// different paths.
template <typename R, int M, typename PT, bool LZ>
__device__ __forceinline__ bool phase_only(const PT& pt, const PointCacheT<R, M, LZ>& pc, const HalfStep& h1,
                                           const HalfStep& h2) {
  return pc.has_g && !pc.pumped && !pt.pot_kind && !pt.g_cplx && !(M > 1 && pt.g_coupled) && h1.delta == h2.delta;
}

template <typename R, int M, bool LZ>
__device__ __forceinline__ void phase_pair(Cx<R> (&v)[kE], int c, const HalfStep& h1,
                                           const PointCacheT<R, M, LZ>& pc) {
  const R twod = (R)2 * (R)h1.delta;  // exact doubling of the half-step's delta
  const R gcr = gdiag<R, M, false>(pc, c);
  R arg[kE], s[kE], co[kE];
#pragma unroll
  for (int m = 0; m < kE; ++m) arg[m] = twod * (pc.cr + gcr * abs2(v[m]));
  fast_sincos_n<R, kE>(arg, s, co);
#pragma unroll
  for (int m = 0; m < kE; ++m) v[m] = mk<R>(co[m], -s[m]) * v[m];
}

// exp(-i dt D(k)) / prod(N) of the k points a thread holds.  The table is
// indexed by (k0', k1') where k_d' = min(k_d, N_d - k_d) when D is even in
// k_d (fold bit d), giving (N_d/2+1) rows instead of N_d, and stored with k1'
// fastest: the column pass reads a column's entries contiguously.  Entry
// forms: 1 scalar, 2 diagonal (M values), and for two components a record of
// four values [E00, E01, E10, E11] (3 full 2x2; 4 symmetric, [a, b, b, d]):
// component 0 takes its pair (own, other) from slots 0-1 and component 1 its
// pair (other, own) from slots 2-3, each as one aligned load.
struct EdDesc {
  int form;       // 0: no dispersion, multiply by `scale` only
  int fold;       // bit 0: dim 0 folded, bit 1: dim 1 folded
  const void* tab;
  int64_t n0, n1; // grid extents
  int64_t e0, e1; // table extents along dim 0 and dim 1 (folded)
  double scale;
};

// Row c of the matrix at the thread's k points: own[m] = E_cc (or the scalar
// or diagonal entry), other[m] = E_c,1-c for the 2x2 forms.  The coupled
// M x M form (M > 2) keeps the k points instead, (a0 + s0 m, a1 + s1 m), and
// reads row c where it is applied.
template <typename R>
struct EdVals {
  Cx<R> own[kE];
  Cx<R> other[kE];
  int64_t a0, s0, a1, s1;
};

__device__ __forceinline__ int64_t ed_index(const EdDesc& d, int64_t k0, int64_t k1) {
  if (d.fold & 1) k0 = (k0 <= d.n0 - k0) ? k0 : d.n0 - k0;
  if (d.fold & 2) k1 = (k1 <= d.n1 - k1) ? k1 : d.n1 - k1;
  return k1 + d.e1 * k0;
}

template <typename R>
struct alignas(2 * sizeof(Cx<R>)) CxPair {
  Cx<R> a, b;
};

// k points (k0, k1) = (a0 + s0 m, a1 + s1 m), m = 0..E-1.
template <typename R, int M>
__device__ __forceinline__ void load_ed(EdVals<R>& ev, const EdDesc& d, int c, int64_t a0, int64_t s0, int64_t a1,
                                        int64_t s1) {
  if (d.form == 0) return;
  if (M > 2 && d.form == 5) {
    ev.a0 = a0;
    ev.s0 = s0;
    ev.a1 = a1;
    ev.s1 = s1;
    return;
  }
  if (M == 2 && d.form >= 3) {
    const CxPair<R>* rec = reinterpret_cast<const CxPair<R>*>(d.tab);
#pragma unroll
    for (int m = 0; m < kE; ++m) {
      const CxPair<R> p = rec[2 * ed_index(d, a0 + s0 * m, a1 + s1 * m) + c];
      ev.own[m] = c ? p.b : p.a;
      ev.other[m] = c ? p.a : p.b;
    }
    return;
  }
  const int per = (d.form == 2) ? M : 1, io = (d.form == 2) ? c : 0;
#pragma unroll
  for (int m = 0; m < kE; ++m) ev.own[m] = ldc<R>(d.tab, ed_index(d, a0 + s0 * m, a1 + s1 * m) * per + io);
}

// Apply row c of exp(-i dt D) / prod(N).  For the 2x2 forms the partner
// component's values at the same k points come through LDS: `mine` and
// `theirs` are this thread's and the partner's padded line slots, and the
// exchange is bracketed by workgroup barriers (uniform: the form is).
template <typename R, int M>
__device__ __forceinline__ void apply_ed(Cx<R> (&v)[kE], const EdDesc& d, const EdVals<R>& ev, int t, int T,
                                         Cx<R>* mine, const Cx<R>* theirs, int c = 0, int cstride = 0) {
  if (d.form == 0) {
#pragma unroll
    for (int m = 0; m < kE; ++m) v[m] = scale(v[m], (R)d.scale);
    return;
  }
  if (M > 2 && d.form == 5) {
    // coupled M x M: every component's line goes to LDS (component j of this
    // line at mine + (j - c) cstride), then row c of E(k) multiplies the M
    // values at each k; the table holds E row-major per k
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kE; ++m) mine[lds_pad(t + m * T)] = v[m];
    __syncthreads();
    const Cx<R>* base = mine - (int64_t)c * cstride;
#pragma unroll
    for (int m = 0; m < kE; ++m) {
      const int64_t row = (ed_index(d, ev.a0 + ev.s0 * m, ev.a1 + ev.s1 * m) * M + c) * M;
      Cx<R> acc = mk<R>((R)0, (R)0);
#pragma unroll
      for (int j = 0; j < M; ++j) acc = acc + ldc<R>(d.tab, row + j) * base[(int64_t)j * cstride + lds_pad(t + m * T)];
      v[m] = acc;
    }
    __syncthreads();
    return;
  }
  if (M == 2 && d.form >= 3) {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kE; ++m) mine[lds_pad(t + m * T)] = v[m];
    __syncthreads();
    Cx<R> p[kE];
#pragma unroll
    for (int m = 0; m < kE; ++m) p[m] = theirs[lds_pad(t + m * T)];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kE; ++m) v[m] = ev.own[m] * v[m] + ev.other[m] * p[m];
    return;
  }
#pragma unroll
  for (int m = 0; m < kE; ++m) v[m] = ev.own[m] * v[m];
}

template <typename R>
__device__ __forceinline__ Cx<R> (&as_bundle(Cx<R> (&v)[kE]))[1][kE] {
  return reinterpret_cast<Cx<R>(&)[1][kE]>(v);
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

// LDS of the x-pass / 1D kernel: two FFT buffers of M*W padded lines.  The
// |u_j|^2 exchange of coupled nonlinearities reuses the first buffer (it runs
// between transforms, fenced by barriers).
template <typename R, int N, int M, int W>
__host__ __device__ constexpr size_t xlds_bytes() {
  return 2ull * M * W * lds_line_stride(N) * sizeof(Cx<R>);
}

template <typename R, int N, int M, int W, bool NOISE>
__global__ void __launch_bounds__(M * W * (N / kE)) x_pass_kernel(XPassArgs a) {
  constexpr int T = N / kE;
  constexpr int LS = lds_line_stride(N);
  extern __shared__ unsigned char smem_raw[];
  Cx<R>* lds = reinterpret_cast<Cx<R>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c = tid / (W * T);
  const int l = (tid / T) % W, t = tid % T;
  const int64_t line_raw = (int64_t)blockIdx.x * W + l;
  const bool valid = line_raw < a.nlines;
  const int64_t line = valid ? line_raw : a.nlines - 1;
  const int64_t y = line % a.ny;
  Cx<R>* uc = reinterpret_cast<Cx<R>*>(pick<M>(a.u, c)) + line * N;
  Cx<R> v[kE];
#pragma unroll
  for (int m = 0; m < kE; ++m) v[m] = uc[t + m * T];
  TwFor<R, N, M * W * (N / kE)> tr;
  if (a.flags & (kInK | kOutK)) load_twiddles<R, N, T>(tr, t, reinterpret_cast<const Cx<R>*>(a.tw), a.tws);
  PointCacheT<R, M, (M * W * T > 512)> pc;
  if (a.flags & (kEnd | kStart)) load_point_cache<R, M>(pc, a.pt, c, t, T, y, y * N);
  Cx<R>* my = lds + (c * W + l) * LS;
  const int boff = M * W * LS;
  R* a2 = reinterpret_cast<R*>(lds) + (size_t)l * M * N;
  // Lines are T consecutive threads, so for T <= 64 a line is wave-local.
  constexpr bool LOCAL = T <= 64;
  if (a.flags & kInK) {
    fft_lines<R, N, T, 1, true, LOCAL>(as_bundle(v), my, 0, boff, t, tr);
    // the coupled half-step reuses buffer 0 for |u_j|^2 (it starts with a
    // workgroup barrier of its own); no other LDS use follows
    if (M > 1 && a.pt.g_coupled) __syncthreads();
  }
  void* fr = pick<M>(a.frame, c);
  // stochastic runs keep both half-steps: the noise acts on the components
  bool pair = false;
  if constexpr (!NOISE)
    pair = (a.flags & kEnd) && (a.flags & kStart) && fr == nullptr && phase_only<R, M>(a.pt, pc, a.h_end, a.h_start);
  if (pair) {
    phase_pair<R, M>(v, c, a.h_end, pc);
  } else {
    if (a.flags & kEnd) half_step<R, N, M, NOISE>(v, c, a.h_end, a.pt, pc, t, T, line * (N / 2), a2);
    if (fr != nullptr && valid) {
      Cx<R>* dst = reinterpret_cast<Cx<R>*>(fr) + line * N;
#pragma unroll
      for (int m = 0; m < kE; ++m) dst[t + m * T] = v[m];
    }
    if (a.flags & kStart) half_step<R, N, M, NOISE>(v, c, a.h_start, a.pt, pc, t, T, line * (N / 2), a2);
  }
  if (a.flags & kOutK) fft_lines<R, N, T, 1, false, LOCAL>(as_bundle(v), my, 0, boff, t, tr);
  if (valid) {
#pragma unroll
    for (int m = 0; m < kE; ++m) uc[t + m * T] = v[m];
  }
}

// ------------------------------------------------------------------ y-pass
struct YPassArgs {
  void* u[kMaxM];
  int64_t nx;         // columns per grid (dim-0 extent; dims 0 and 1 of a 3D grid's z-pass)
  int64_t ncolblk;    // nx / C
  int64_t nblocks;    // ncolblk * batch
  int xcd_remap;      // 1: give each XCD a contiguous range of column blocks
  int mode;           // 0: FFT, exp(-i dt D), IFFT; 1: FFT only; 2: IFFT only
  int64_t kx_n, ky_n; // z-pass of a 3D grid: column = kx + kx_n ky, table k1 = ky + ky_n kz (0 otherwise)
  EdDesc ed;
  const void* tw;
  int tws;
};

template <typename R, int N, int M, int C>
__host__ __device__ constexpr size_t ylds_bytes() {
  return 2ull * M * C * lds_line_stride(N) * sizeof(Cx<R>);
}

template <typename R, int N, int M, int C>
__global__ void __launch_bounds__(M * C * (N / kE)) y_pass_kernel(YPassArgs a) {
  constexpr int T = N / kE;
  constexpr int LS = lds_line_stride(N);
  extern __shared__ unsigned char smem_raw[];
  Cx<R>* lds = reinterpret_cast<Cx<R>*>(smem_raw);
  int64_t blk = blockIdx.x;
  if (a.xcd_remap) {
    // Workgroups are dispatched round-robin over the 8 XCDs; map them so each
    // XCD works on a contiguous range of column blocks and neighbouring
    // columns that share cache lines meet in one L2.
    const int64_t per = a.nblocks / 8;
    blk = (blk % 8) * per + blk / 8;
  }
  const int64_t b = blk / a.ncolblk;
  const int64_t x0 = (blk % a.ncolblk) * C;
  const int tid = threadIdx.x;
  const int c = tid / (C * T);
  const int r = tid % (C * T);
  const int cx = r % C, t = r / C;
  const int64_t nx = a.nx;
  Cx<R>* uc = reinterpret_cast<Cx<R>*>(pick<M>(a.u, c)) + b * nx * N + x0 + cx;
  Cx<R> v[kE];
#pragma unroll
  for (int m = 0; m < kE; ++m) v[m] = uc[(int64_t)(t + m * T) * nx];
  TwFor<R, N, M * C * (N / kE)> tr;
  load_twiddles<R, N, T>(tr, t, reinterpret_cast<const Cx<R>*>(a.tw), a.tws);
  EdVals<R> ev;
  if (a.kx_n) {
    // z-pass of a 3D grid: the table holds (kx, ky + ny kz), kz fastest
    const int64_t col = x0 + cx;
    load_ed<R, M>(ev, a.ed, c, col % a.kx_n, 0, col / a.kx_n + a.ky_n * t, a.ky_n * T);
  } else if (a.mode == 0) {
    load_ed<R, M>(ev, a.ed, c, x0 + cx, 0, t, T);
  }
  Cx<R>* my = lds + (c * C + cx) * LS;
  const Cx<R>* partner = lds + ((M - 1 - c) * C + cx) * LS;
  const int boff = M * C * LS;
  // a component's C columns interleave over C*T consecutive threads
  constexpr bool LOCAL = C * T <= 64;
  if (a.mode != 0) {
    // transform-only pass (the y direction of a 3D step, on either side of
    // the z-pass); the mode is uniform over the launch
    if (a.mode == 1)
      fft_lines<R, N, T, 1, false, LOCAL>(as_bundle(v), my, 0, boff, t, tr);
    else
      fft_lines<R, N, T, 1, true, LOCAL>(as_bundle(v), my, 0, boff, t, tr);
#pragma unroll
    for (int m = 0; m < kE; ++m) uc[(int64_t)(t + m * T) * nx] = v[m];
    return;
  }
  // buffer read by the forward transform's last exchange; the other one is
  // free once that transform ends, so the components exchange through it and
  // the inverse transform starts in this one (no barrier between them)
  constexpr int LAST = (fft_exchanges(N) - 1) & 1;
  fft_lines<R, N, T, 1, false, LOCAL>(as_bundle(v), my, 0, boff, t, tr);
  if (M == 2 && a.ed.form >= 3) {
    Cx<R>* mine = my + (1 - LAST) * boff;
    const Cx<R>* theirs = partner + (1 - LAST) * boff;
#pragma unroll
    for (int m = 0; m < kE; ++m) mine[lds_pad(t + m * T)] = v[m];
    __syncthreads();
    Cx<R> p[kE];
#pragma unroll
    for (int m = 0; m < kE; ++m) p[m] = theirs[lds_pad(t + m * T)];
    // a wave-local inverse transform would not wait for the partner's reads
    // before writing buffer 1 - LAST (its second pass)
    if constexpr (LOCAL) __syncthreads();
#pragma unroll
    for (int m = 0; m < kE; ++m) v[m] = ev.own[m] * v[m] + ev.other[m] * p[m];
  } else {
    apply_ed<R, M>(v, a.ed, ev, t, T, my, partner, c, C * LS);
    line_sync<LOCAL>();
  }
  fft_lines<R, N, T, 1, true, LOCAL, LAST>(as_bundle(v), my, 0, boff, t, tr);
#pragma unroll
  for (int m = 0; m < kE; ++m) uc[(int64_t)(t + m * T) * nx] = v[m];
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
  EdDesc ed;
  const void* tw;
  int tws;
  PointTerms pt;
};

template <typename R, int N, int M, int W, bool NOISE>
__global__ void __launch_bounds__(M * W * (N / kE)) step1d_kernel(Step1DArgs a) {
  constexpr int T = N / kE;
  constexpr bool LOCAL = T <= 64;  // lines are T consecutive threads
  constexpr int LS = lds_line_stride(N);
  extern __shared__ unsigned char smem_raw[];
  Cx<R>* lds = reinterpret_cast<Cx<R>*>(smem_raw);
  const int tid = threadIdx.x;
  const int c = tid / (W * T);
  const int l = (tid / T) % W, t = tid % T;
  const int64_t line_raw = (int64_t)blockIdx.x * W + l;
  const bool valid = line_raw < a.nlines;
  const int64_t line = valid ? line_raw : a.nlines - 1;
  Cx<R>* uc = reinterpret_cast<Cx<R>*>(pick<M>(a.u, c)) + line * N;
  Cx<R> v[kE];
#pragma unroll
  for (int m = 0; m < kE; ++m) v[m] = uc[t + m * T];
  TwFor<R, N, M * W * (N / kE)> tr;
  load_twiddles<R, N, T>(tr, t, reinterpret_cast<const Cx<R>*>(a.tw), a.tws);
  PointCacheT<R, M, (M * W * T > 512)> pc;
  load_point_cache<R, M>(pc, a.pt, c, t, T, 0, 0);
  EdVals<R> ev;
  load_ed<R, M>(ev, a.ed, c, t, T, 0, 0);
  Cx<R>* my = lds + (c * W + l) * LS;
  const Cx<R>* partner = lds + ((M - 1 - c) * W + l) * LS;
  const int boff = M * W * LS;
  R* a2 = reinterpret_cast<R*>(lds) + (size_t)l * M * N;
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
    h.jnow = e0;
    h.jnext = e0 + 1;
    half_step<R, N, M, NOISE>(v, c, h, a.pt, pc, t, T, line * (N / 2), a2);
    fft_lines<R, N, T, 1, false, LOCAL>(as_bundle(v), my, 0, boff, t, tr);
    apply_ed<R, M>(v, a.ed, ev, t, T, my, partner, c, W * LS);
    line_sync<LOCAL>();
    fft_lines<R, N, T, 1, true, LOCAL>(as_bundle(v), my, 0, boff, t, tr);
    h.fnow[0] = q * en[1][0];
    h.fnow[1] = q * en[1][1];
    h.fnext[0] = q * en[2][0];
    h.fnext[1] = q * en[2][1];
    h.call += 1;
    h.jnow = e0 + 1;
    h.jnext = e0 + 2;
    half_step<R, N, M, NOISE>(v, c, h, a.pt, pc, t, T, line * (N / 2), a2);
    line_sync<LOCAL>();
  }
  if (valid) {
#pragma unroll
    for (int m = 0; m < kE; ++m) uc[t + m * T] = v[m];
    void* fr = pick<M>(a.frame, c);
    if (fr != nullptr) {
      Cx<R>* dst = reinterpret_cast<Cx<R>*>(fr) + line * N;
#pragma unroll
      for (int m = 0; m < kE; ++m) dst[t + m * T] = v[m];
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
  const double* D;  // complex FP64 table, npts * per entries, k0 fastest (compact order)
  void* out;        // device table, k1 fastest (ed_index order)
  int64_t npts;
  int64_t e0, e1;   // compact extents: npts = e0 * e1
  int form;         // 1 scalar, 2 diag, 3 full 2x2, 4 symmetric 2x2 (a, b, d), 5 coupled M x M (M > 2)
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

// exp(A) of an M x M complex matrix (M <= kMaxM, row-major with row stride
// kMaxM) in FP64: scaling and squaring (Higham, SIAM J. Matrix Anal. Appl.
// 26 (2005)) with a degree-18 Taylor polynomial of A / 2^s, s chosen so that
// the max-row-sum norm of A / 2^s is <= 1/2 (truncation < 0.5^19 / 19! ~ 2e-23
// relative to the norm bound).
__device__ __noinline__ void expm_small_d(const Cx<double> (&A)[kMaxM * kMaxM], Cx<double> (&E)[kMaxM * kMaxM],
                                          int M) {
  double nrm = 0.0;
  for (int r = 0; r < M; ++r) {
    double row = 0.0;
    for (int c = 0; c < M; ++c) row += hypot(A[r * kMaxM + c].re, A[r * kMaxM + c].im);
    nrm = fmax(nrm, row);
  }
  int s = 0;
  while (nrm > 0.5 && s < 60) {
    nrm *= 0.5;
    ++s;
  }
  const double f = ldexp(1.0, -s);
  Cx<double> X[kMaxM * kMaxM], T[kMaxM * kMaxM], P[kMaxM * kMaxM];
  for (int i = 0; i < kMaxM * kMaxM; ++i) {
    X[i] = scale(A[i], f);
    E[i] = mk<double>(0.0, 0.0);
    T[i] = mk<double>(0.0, 0.0);
  }
  for (int r = 0; r < M; ++r) {
    E[r * kMaxM + r] = mk<double>(1.0, 0.0);
    T[r * kMaxM + r] = mk<double>(1.0, 0.0);
  }
  // E = sum_{n <= 18} X^n / n!, term T_n = T_{n-1} X / n
  for (int n = 1; n <= 18; ++n) {
    for (int r = 0; r < M; ++r)
      for (int c = 0; c < M; ++c) {
        Cx<double> acc = mk<double>(0.0, 0.0);
        for (int j = 0; j < M; ++j) acc = acc + T[r * kMaxM + j] * X[j * kMaxM + c];
        P[r * kMaxM + c] = scale(acc, 1.0 / n);
      }
    for (int r = 0; r < M; ++r)
      for (int c = 0; c < M; ++c) {
        T[r * kMaxM + c] = P[r * kMaxM + c];
        E[r * kMaxM + c] = E[r * kMaxM + c] + P[r * kMaxM + c];
      }
  }
  for (int q = 0; q < s; ++q) {
    for (int r = 0; r < M; ++r)
      for (int c = 0; c < M; ++c) {
        Cx<double> acc = mk<double>(0.0, 0.0);
        for (int j = 0; j < M; ++j) acc = acc + E[r * kMaxM + j] * E[j * kMaxM + c];
        P[r * kMaxM + c] = acc;
      }
    for (int r = 0; r < M; ++r)
      for (int c = 0; c < M; ++c) E[r * kMaxM + c] = P[r * kMaxM + c];
  }
}

template <typename R>
__global__ void ed_table_kernel(EdArgs a) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.npts) return;
  Cx<R>* out = reinterpret_cast<Cx<R>*>(a.out);
  const Cx<double>* D = reinterpret_cast<const Cx<double>*>(a.D);
  const int64_t o = k / a.e0 + a.e1 * (k % a.e0);  // transposed position
  if (a.form == 1 || a.form == 2) {
    const int per = (a.form == 1) ? 1 : a.M;
    for (int i = 0; i < per; ++i) {
      const Cx<double> z = D[k * per + i];
      // exp(-i dt z)
      const Cx<double> e = cexp_d(mk<double>(a.dt * z.im, -a.dt * z.re));
      out[o * per + i] = mk<R>((R)(e.re * a.scale), (R)(e.im * a.scale));
    }
  } else if (a.form == 5) {
    // coupled M x M (M = 3, 4), D column-major per k; E stored row-major
    Cx<double> A[kMaxM * kMaxM], E[kMaxM * kMaxM];
    const int M = a.M;
    for (int col = 0; col < M; ++col)
      for (int row = 0; row < M; ++row) {
        const Cx<double> z = D[k * M * M + row + M * col];
        A[row * kMaxM + col] = mk<double>(a.dt * z.im, -a.dt * z.re);
      }
    expm_small_d(A, E, M);
    for (int row = 0; row < M; ++row)
      for (int col = 0; col < M; ++col) {
        const Cx<double> e = E[row * kMaxM + col];
        out[(o * M + row) * M + col] = mk<R>((R)(e.re * a.scale), (R)(e.im * a.scale));
      }
  } else if (a.form == 4) {
    // symmetric 2x2 stored as (a, b, d); exp of a symmetric matrix is symmetric
    Cx<double> A[4], E[4];
    const int src[4] = {0, 1, 1, 2};
    for (int i = 0; i < 4; ++i) {
      const Cx<double> z = D[k * 3 + src[i]];
      A[i] = mk<double>(a.dt * z.im, -a.dt * z.re);
    }
    expm2x2_d(A, E);
    // record [E00, E01, E10, E11] = [a, b, b, d]
    const int src4[4] = {0, 2, 1, 3};
    for (int i = 0; i < 4; ++i)
      out[o * 4 + i] = mk<R>((R)(E[src4[i]].re * a.scale), (R)(E[src4[i]].im * a.scale));
  } else {
    Cx<double> A[4], E[4];
    for (int i = 0; i < 4; ++i) {
      const Cx<double> z = D[k * 4 + i];
      A[i] = mk<double>(a.dt * z.im, -a.dt * z.re);
    }
    expm2x2_d(A, E);
    // E column-major [E00, E10, E01, E11] -> record [E00, E01, E10, E11]
    const int src4[4] = {0, 2, 1, 3};
    for (int i = 0; i < 4; ++i)
      out[o * 4 + i] = mk<R>((R)(E[src4[i]].re * a.scale), (R)(E[src4[i]].im * a.scale));
  }
}

// exp(-i tau V) planes (potential, get_exponential with dt/2), per component.
// --------------------------------------- matrix nonlinearity half-steps
// The real-space half-step of src/kernels.jl:37-54 when nonlinearity(u) is a
// 2x2 SMatrix: cis(-delta G(u)) is then a per-point matrix exponential
// (LinearAlgebra's cis of a matrix) applied to the component vector,
//   u <- exp(-i delta G(u)) (u + delta/2 F_now) + delta/2 F_next,
// with G_ab(u) = c_ab + sum_j g_abj |u_j|^2 from the pre-update fields, in
// FP64.  A 2D step runs it between an x-pass that only inverse-transforms and
// one that only transforms.
struct GmatArgs {
  void* u[kMaxM];
  void* frame[kMaxM];   // real-space copy after the end half-step, or null
  int64_t npts;         // points per component, all trajectories
  int64_t n0, nsp;      // dim-0 extent, points per grid
  int flags;            // kEnd | kStart
  HalfStep h_end, h_start;
  double g[2][2][2][2]; // g[a][b][j] complex
  double c[2][2][2];    // c[a][b] complex
  int pump_kind;        // 0 none, 1 table, 2 separable, 3 sampled (plane j at nsp * j)
  const void* pump_a[kMaxM];
  const void* pump_b[kMaxM];
  int noise_kind;       // 0 none, 1 complex, 2 real (as PointTerms)
  double eta[kMaxM];
  uint32_t seed_lo, seed_hi;
  int pot_kind;         // 0 none, 1 scalar V, 2 SVector V
  const void* pot[kMaxM];  // exp(-i dt/2 V) planes
};

// The Wiener increment of component c at flat point p of a half-step, drawn
// exactly as noise_terms does for the fused kernels: line = p / n0, the pair
// index P = line * n0/2 + (x mod n0/2) and words (x, y) for x < n0/2, (z, w)
// for the other half of the line.
template <typename R>
__device__ __forceinline__ Cx<double> gmat_noise(const GmatArgs& a, const HalfStep& h, int c, int64_t p) {
  const int64_t half = a.n0 / 2;
  const int64_t x = p % a.n0, line = p / a.n0;
  const uint64_t P = (uint64_t)(line * half + (x % half));
  U4 ctr;
  ctr.x = (uint32_t)P;
  ctr.y = (uint32_t)c + 16u * (uint32_t)(P >> 32);
  ctr.z = (uint32_t)h.call;
  ctr.w = (uint32_t)(h.call >> 32);
  const U4 w = philox4x32_10(ctr, a.seed_lo, a.seed_hi);
  Cx<R> z = x < half ? box_muller<R>(w.x, w.y) : box_muller<R>(w.z, w.w);
  if (a.noise_kind == 2) z = mk<R>(z.re * (R)1.41421356237309504880168872420969808, (R)0);
  const R amp = (R)sqrt(h.delta) * (R)a.eta[c];
  return mk<double>((double)(amp * z.im), (double)(-amp * z.re));
}

template <typename R>
__device__ __forceinline__ Cx<double> gmat_pump(const GmatArgs& a, int comp, int64_t p, int64_t j = 0) {
  const void* pa = a.pump_a[comp];
  if (a.pump_kind == 0 || pa == nullptr) return mk<double>(0.0, 0.0);
  const int64_t sp = p % a.nsp;
  Cx<R> v;
  if (a.pump_kind == 3) {
    v = ldc<R>(pa, j * a.nsp + sp);
  } else if (a.pump_kind == 1) {
    v = ldc<R>(pa, sp);
  } else {
    const Cx<R> pb = a.pump_b[comp] ? ldc<R>(a.pump_b[comp], sp / a.n0) : mk<R>((R)1, (R)0);
    v = ldc<R>(pa, sp % a.n0) * pb;
  }
  return mk<double>((double)v.re, (double)v.im);
}

template <typename R>
__global__ void __launch_bounds__(256) gmat_half_kernel(GmatArgs a) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.npts) return;
  Cx<R>* u0 = reinterpret_cast<Cx<R>*>(a.u[0]);
  Cx<R>* u1 = reinterpret_cast<Cx<R>*>(a.u[1]);
  Cx<double> v[2] = {mk<double>((double)u0[p].re, (double)u0[p].im), mk<double>((double)u1[p].re, (double)u1[p].im)};
  const Cx<double> F[2] = {gmat_pump<R>(a, 0, p), gmat_pump<R>(a, 1, p)};
  auto half = [&](const HalfStep& h) {
    const double a2[2] = {abs2(v[0]), abs2(v[1])};
    Cx<double> A[4];  // -i delta G, column-major
    for (int col = 0; col < 2; ++col)
      for (int row = 0; row < 2; ++row) {
        double gr = a.c[row][col][0], gi = a.c[row][col][1];
        for (int j = 0; j < 2; ++j) {
          gr += a.g[row][col][j][0] * a2[j];
          gi += a.g[row][col][j][1] * a2[j];
        }
        A[row + 2 * col] = mk<double>(h.delta * gi, -h.delta * gr);
      }
    Cx<double> E[4];
    expm2x2_d(A, E);
    const Cx<double> fn = mk<double>(h.fnow[0], h.fnow[1]), fx = mk<double>(h.fnext[0], h.fnext[1]);
    // F at the previous and at this pump time (one profile unless sampled)
    const bool smp = a.pump_kind == 3;
    const Cx<double> Fn[2] = {smp ? gmat_pump<R>(a, 0, p, h.jnow) : F[0], smp ? gmat_pump<R>(a, 1, p, h.jnow) : F[1]};
    const Cx<double> Fx[2] = {smp ? gmat_pump<R>(a, 0, p, h.jnext) : F[0], smp ? gmat_pump<R>(a, 1, p, h.jnext) : F[1]};
    const Cx<double> w0 = v[0] + Fn[0] * fn, w1 = v[1] + Fn[1] * fn;
    // _mul(cis(-delta G), exp(-i delta V)) of src/kernels.jl:9-15,44: a
    // scalar V scales the matrix; an SVector V makes it the vector E pv, which
    // then multiplies the fields elementwise (SVector * SVector)
    const int64_t sp = p % a.nsp;
    if (a.pot_kind == 2) {
      const Cx<R> q0 = ldc<R>(a.pot[0], sp), q1 = ldc<R>(a.pot[1], sp);
      const Cx<double> p0 = mk<double>((double)q0.re, (double)q0.im), p1 = mk<double>((double)q1.re, (double)q1.im);
      const Cx<double> e0 = E[0] * p0 + E[2] * p1, e1 = E[1] * p0 + E[3] * p1;
      v[0] = e0 * w0 + Fx[0] * fx;
      v[1] = e1 * w1 + Fx[1] * fx;
    } else {
      if (a.pot_kind == 1) {
        const Cx<R> q = ldc<R>(a.pot[0], sp);
        const Cx<double> sv = mk<double>((double)q.re, (double)q.im);
        for (int i = 0; i < 4; ++i) E[i] = E[i] * sv;
      }
      v[0] = E[0] * w0 + E[2] * w1 + Fx[0] * fx;
      v[1] = E[1] * w0 + E[3] * w1 + Fx[1] * fx;
    }
    if (a.noise_kind) {
      v[0] = v[0] + gmat_noise<R>(a, h, 0, p);
      v[1] = v[1] + gmat_noise<R>(a, h, 1, p);
    }
  };
  if (a.flags & kEnd) half(a.h_end);
  if (a.frame[0]) {
    reinterpret_cast<Cx<R>*>(a.frame[0])[p] = mk<R>((R)v[0].re, (R)v[0].im);
    reinterpret_cast<Cx<R>*>(a.frame[1])[p] = mk<R>((R)v[1].re, (R)v[1].im);
  }
  if (a.flags & kStart) half(a.h_start);
  u0[p] = mk<R>((R)v[0].re, (R)v[0].im);
  u1[p] = mk<R>((R)v[1].re, (R)v[1].im);
}

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

// ------------------------------------------- uneven-size instance sweeps
// A sweep holds many independent 1D single-component instances of different
// lengths and parameters, concatenated in one buffer (the reference solves
// them one GrossPitaevskiiProblem at a time).  One launch per distinct length
// N processes that length's instances for a whole save block, each line
// resident in registers like step1d_kernel.
struct SweepInst {
  int64_t off;       // first point in the concatenated field / tables
  int64_t pump_off;  // first point of the pump profile, -1 if unpumped
  int64_t res_off;   // first value of this instance's result block (frames of N values)
  double g[2];       // G = c + g |u|^2 (complex)
  double c[2];
  double eta;        // constant position-noise amplitude
};

struct SweepArgs {
  void* u;                   // concatenated field
  void* res;                 // instance-major frames, or null
  int64_t frame;             // frame index written at the end of the launch
  const SweepInst* inst;     // all instances
  const int32_t* ids;        // instances of this launch's length
  int64_t ninst;             // number of ids
  int64_t nsteps;
  const double* env;         // shared complex envelope samples or null (== 1)
  int64_t env_base;
  double delta;              // dt/2
  uint64_t call0;
  const void* ed;            // exp(-i dt D_i(k)) per instance (scalar form, unscaled)
  const void* pump;          // concatenated pump profiles
  const void* tw;
  int tws;
  int noise_kind;
  uint32_t seed_lo, seed_hi;
  int g_cplx;
};

template <typename R, int N, int W, bool NOISE>
__global__ void __launch_bounds__(W * (N / kE)) sweep1d_kernel(SweepArgs a) {
  constexpr int T = N / kE;
  constexpr bool LOCAL = T <= 64;  // instances are T consecutive threads
  constexpr int LS = lds_line_stride(N);
  extern __shared__ unsigned char smem_raw[];
  Cx<R>* lds = reinterpret_cast<Cx<R>*>(smem_raw);
  const int tid = threadIdx.x;
  const int l = tid / T, t = tid % T;
  const int64_t j_raw = (int64_t)blockIdx.x * W + l;
  const bool valid = j_raw < a.ninst;
  const SweepInst in = a.inst[a.ids[valid ? j_raw : a.ninst - 1]];
  Cx<R>* uc = reinterpret_cast<Cx<R>*>(a.u) + in.off;
  Cx<R> v[kE];
#pragma unroll
  for (int m = 0; m < kE; ++m) v[m] = uc[t + m * T];
  TwReg<R, N> tr;
  load_twiddles<R, N, T>(tr, t, reinterpret_cast<const Cx<R>*>(a.tw), a.tws);
  // per-instance terms: exp(-i dt D)/N, pump profile, G and eta of row 0
  const R invn = (R)1 / (R)N;
  EdVals<R> ev;
#pragma unroll
  for (int m = 0; m < kE; ++m) ev.own[m] = scale(ldc<R>(a.ed, in.off + t + m * T), invn);
  PointTermsT<1> pt;
  memset(&pt, 0, sizeof pt);
  pt.gmask = (in.g[0] != 0.0 || in.g[1] != 0.0 || in.c[0] != 0.0 || in.c[1] != 0.0) ? 1 : 0;
  pt.g_cplx = a.g_cplx;
  pt.g[0][0][0] = in.g[0];
  pt.g[0][0][1] = in.g[1];
  pt.c[0][0] = in.c[0];
  pt.c[0][1] = in.c[1];
  pt.noise_kind = a.noise_kind;
  pt.eta[0] = in.eta;
  pt.seed_lo = a.seed_lo;
  pt.seed_hi = a.seed_hi;
  pt.pump_kind = (in.pump_off >= 0 && a.pump) ? 2 : 0;
  pt.pump_a[0] = pt.pump_kind ? (const void*)(reinterpret_cast<const Cx<R>*>(a.pump) + in.pump_off) : nullptr;
  PointCache<R, 1> pc;
  load_point_cache<R, 1>(pc, pt, 0, t, T, 0, 0);
  Cx<R>* my = lds + l * LS;
  const int boff = W * LS;
  EdDesc ed;
  memset(&ed, 0, sizeof ed);
  ed.form = 1;
  const double q = 0.5 * a.delta;
  const int64_t pbase = in.off / 2;
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
    half_step<R, N, 1, NOISE>(v, 0, h, pt, pc, t, T, pbase, nullptr);
    fft_lines<R, N, T, 1, false, LOCAL>(as_bundle(v), my, 0, boff, t, tr);
    apply_ed<R, 1>(v, ed, ev, t, T, my, my);
    line_sync<LOCAL>();
    fft_lines<R, N, T, 1, true, LOCAL>(as_bundle(v), my, 0, boff, t, tr);
    h.fnow[0] = q * en[1][0];
    h.fnow[1] = q * en[1][1];
    h.fnext[0] = q * en[2][0];
    h.fnext[1] = q * en[2][1];
    h.call += 1;
    half_step<R, N, 1, NOISE>(v, 0, h, pt, pc, t, T, pbase, nullptr);
    line_sync<LOCAL>();
  }
  if (valid) {
#pragma unroll
    for (int m = 0; m < kE; ++m) uc[t + m * T] = v[m];
    if (a.res) {
      Cx<R>* dst = reinterpret_cast<Cx<R>*>(a.res) + in.res_off + a.frame * N;
#pragma unroll
      for (int m = 0; m < kE; ++m) dst[t + m * T] = v[m];
    }
  }
}

// ------------------------------------------------- ensemble observables
// For every spatial point r and component c: the batch means of |u|^2 and
// |u|^4 of one saved frame, accumulated in FP64 with Neumaier compensation
// (examples/truncated_wigner.jl:110-165 computes these moments after the
// solve in user code).  One thread per point; the batch loop reads
// consecutive points at fixed trajectory, so loads are coalesced.
struct ObsArgs {
  const void* u[kMaxM];  // real-space frame, nsp * batch values per component
  double* n1[kMaxM];     // nsp outputs per component (this frame)
  double* n2[kMaxM];
  int64_t nsp, batch;
  int M;
};

__device__ __forceinline__ void neumaier_add(double& s, double& comp, double x) {
  const double t = s + x;
  comp += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
  s = t;
}

template <typename R>
__global__ void observables_kernel(ObsArgs a) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.nsp) return;
  const double inv = 1.0 / (double)a.batch;
  for (int c = 0; c < a.M; ++c) {
    const Cx<R>* u = reinterpret_cast<const Cx<R>*>(a.u[c]) + r;
    double s1 = 0, c1 = 0, s2 = 0, c2 = 0;
    int64_t b = 0;
    for (; b + 4 <= a.batch; b += 4) {
      Cx<R> v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = u[(b + i) * a.nsp];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const double x = (double)v[i].re * (double)v[i].re + (double)v[i].im * (double)v[i].im;
        neumaier_add(s1, c1, x);
        neumaier_add(s2, c2, x * x);
      }
    }
    for (; b < a.batch; ++b) {
      const Cx<R> v = u[b * a.nsp];
      const double x = (double)v.re * (double)v.re + (double)v.im * (double)v.im;
      neumaier_add(s1, c1, x);
      neumaier_add(s2, c2, x * x);
    }
    a.n1[c][r] = (s1 + c1) * inv;
    a.n2[c][r] = (s2 + c2) * inv;
  }
}

// ------------------------------------------------ steady-state residual
// sum |a - b|^2 and sum |a|^2 over all points and components in FP64 with
// Neumaier compensation: per-block partials, then one block folds them.
// ---------------------------------------- momentum-space pair correlations
// n(k) and G2(k, k') of a 1D truncated-Wigner ensemble, the observables of
// examples/truncated_wigner.jl:110-165:
//   ft = fft(u_b) / N per trajectory b,  a_b(k) = |ft(k)|^2,
//   nbar(k) = mean_b a_b(k),  n(k) = nbar(k) - c/2,
//   G2(k, k') = mean_b a_b(k) a_b(k') - (1 + delta_kk') c/2 (nbar(k) + nbar(k') - c/2)
// with c the commutator (1/L).  Everything is in FFT order; the example's
// fftshift is a permutation left to the caller.
struct KAbs2Args {
  const void* u;  // N x batch complex values (one component of one frame)
  double* a;      // batch x N, k fastest
  int64_t batch;
  const void* tw;
  int tws;
};

// a_b(k): W trajectories per workgroup, each line transformed by its T
// threads with the fused kernels' forward Stockham FFT.
template <typename R, int N, int W>
__global__ void __launch_bounds__(W * (N / kE)) kspace_abs2_kernel(KAbs2Args p) {
  constexpr int T = N / kE;
  constexpr int LS = lds_line_stride(N);
  constexpr bool LOCAL = T <= 64;
  extern __shared__ unsigned char smem_raw[];
  Cx<R>* lds = reinterpret_cast<Cx<R>*>(smem_raw);
  const int tid = threadIdx.x;
  const int l = tid / T, t = tid % T;
  const int64_t line_raw = (int64_t)blockIdx.x * W + l;
  const bool valid = line_raw < p.batch;
  const int64_t line = valid ? line_raw : p.batch - 1;
  const Cx<R>* uc = reinterpret_cast<const Cx<R>*>(p.u) + line * N;
  Cx<R> v[kE];
#pragma unroll
  for (int m = 0; m < kE; ++m) v[m] = uc[t + m * T];
  TwReg<R, N> tr;
  load_twiddles<R, N, T>(tr, t, reinterpret_cast<const Cx<R>*>(p.tw), p.tws);
  fft_lines<R, N, T, 1, false, LOCAL>(as_bundle(v), lds + l * LS, 0, W * LS, t, tr);
  if (valid) {
#pragma unroll
    for (int m = 0; m < kE; ++m) {
      const double re = (double)v[m].re / N, im = (double)v[m].im / N;
      p.a[line * N + t + m * T] = re * re + im * im;
    }
  }
}

// nbar(k) = mean_b a_b(k) (Neumaier-compensated) and n(k) = nbar(k) - c/2.
// (A template so that every unit including this header may instantiate it.)
template <int UNUSED = 0>
__global__ void __launch_bounds__(256) kspace_mean_kernel(const double* a, int64_t n, int64_t batch, double comm,
                                                          double* nbar, double* nk) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double s = 0, comp = 0;
  for (int64_t b = 0; b < batch; ++b) neumaier_add(s, comp, a[b * n + k]);
  const double m = (s + comp) / (double)batch;
  nbar[k] = m;
  nk[k] = m - comm / 2;
}

// G2 on 16 x 16 tiles of (k, k'): the trajectory sum of a_b(k) a_b(k') is
// staged through LDS 16 trajectories at a time and accumulated per entry with
// Neumaier compensation.
template <int UNUSED = 0>
__global__ void __launch_bounds__(256) kspace_g2_kernel(const double* a, const double* nbar, int64_t n, int64_t batch,
                                                        double comm, double* g2) {
  __shared__ double am[16][17], an[16][17];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int64_t m0 = (int64_t)blockIdx.x * 16, n0 = (int64_t)blockIdx.y * 16;
  double s = 0, comp = 0;
  for (int64_t b0 = 0; b0 < batch; b0 += 16) {
    const int64_t b = b0 + ty;
    am[ty][tx] = b < batch ? a[b * n + m0 + tx] : 0.0;
    an[ty][tx] = b < batch ? a[b * n + n0 + tx] : 0.0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; ++j) neumaier_add(s, comp, am[j][tx] * an[j][ty]);
    __syncthreads();
  }
  const int64_t km = m0 + tx, kn = n0 + ty;
  const double d = km == kn ? 2.0 : 1.0;
  g2[kn * n + km] = (s + comp) / (double)batch - d * (comm / 2) * (nbar[km] + nbar[kn] - comm / 2);
}

// Windowed first-order correlation of a 1D ensemble frame
// (test/windowed_ft.jl:31-49): X_j,b = fft(w_j u_b) for two windows w_1, w_2,
// G(k1, k2) = sum_b X_1,b(k1) conj(X_2,b(k2)) / (N batch), in FFT order (the
// reference's fftshift / ifftshift pair is a permutation with signs (-1)^k,
// applied by the caller).
struct KWinArgs {
  const void* u;       // N x batch complex values
  const double* w1;    // N complex FP64 window values
  const double* w2;
  double* x1;          // batch x N complex FP64, k fastest
  double* x2;
  int64_t batch;
  const void* tw;
  int tws;
};

template <typename R, int N, int W>
__global__ void __launch_bounds__(W * (N / kE)) kspace_windowed_kernel(KWinArgs p) {
  constexpr int T = N / kE;
  constexpr int LS = lds_line_stride(N);
  constexpr bool LOCAL = T <= 64;
  extern __shared__ unsigned char smem_raw[];
  Cx<R>* lds = reinterpret_cast<Cx<R>*>(smem_raw);
  const int tid = threadIdx.x;
  const int l = tid / T, t = tid % T;
  const int64_t line_raw = (int64_t)blockIdx.x * W + l;
  const bool valid = line_raw < p.batch;
  const int64_t line = valid ? line_raw : p.batch - 1;
  const Cx<R>* uc = reinterpret_cast<const Cx<R>*>(p.u) + line * N;
  Cx<R> u[kE];
#pragma unroll
  for (int m = 0; m < kE; ++m) u[m] = uc[t + m * T];
  TwReg<R, N> tr;
  load_twiddles<R, N, T>(tr, t, reinterpret_cast<const Cx<R>*>(p.tw), p.tws);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const double* w = j ? p.w2 : p.w1;
    double* x = j ? p.x2 : p.x1;
    Cx<R> v[kE];
#pragma unroll
    for (int m = 0; m < kE; ++m) {
      const int i = t + m * T;
      v[m] = u[m] * mk<R>((R)w[2 * i], (R)w[2 * i + 1]);
    }
    // the second transform reuses the LDS buffers of the first
    if (j) line_sync<LOCAL>();
    fft_lines<R, N, T, 1, false, LOCAL>(as_bundle(v), lds + l * LS, 0, W * LS, t, tr);
    if (valid) {
#pragma unroll
      for (int m = 0; m < kE; ++m) {
        const int64_t o = 2 * (line * N + t + m * T);
        x[o] = (double)v[m].re;
        x[o + 1] = (double)v[m].im;
      }
    }
  }
}

// G(k1, k2) on 16 x 16 tiles, trajectory sums staged through LDS and
// accumulated with Neumaier compensation (real and imaginary parts).
template <int UNUSED = 0>
__global__ void __launch_bounds__(256) kspace_g1_kernel(const double* x1, const double* x2, int64_t n, int64_t batch,
                                                        double* g1) {
  __shared__ double ar[16][17], ai[16][17], br[16][17], bi[16][17];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int64_t m0 = (int64_t)blockIdx.x * 16, n0 = (int64_t)blockIdx.y * 16;
  double sr = 0, cr = 0, si = 0, ci = 0;
  for (int64_t b0 = 0; b0 < batch; b0 += 16) {
    const int64_t b = b0 + ty;
    const bool in = b < batch;
    ar[ty][tx] = in ? x1[2 * (b * n + m0 + tx)] : 0.0;
    ai[ty][tx] = in ? x1[2 * (b * n + m0 + tx) + 1] : 0.0;
    br[ty][tx] = in ? x2[2 * (b * n + n0 + tx)] : 0.0;
    bi[ty][tx] = in ? x2[2 * (b * n + n0 + tx) + 1] : 0.0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      // (a + i b') (c - i d) = (a c + b' d) + i (b' c - a d)
      neumaier_add(sr, cr, ar[j][tx] * br[j][ty] + ai[j][tx] * bi[j][ty]);
      neumaier_add(si, ci, ai[j][tx] * br[j][ty] - ar[j][tx] * bi[j][ty]);
    }
    __syncthreads();
  }
  const double scale = 1.0 / ((double)n * (double)batch);
  const int64_t o = 2 * ((n0 + ty) * n + m0 + tx);
  g1[o] = (sr + cr) * scale;
  g1[o + 1] = (si + ci) * scale;
}

// Density-density correlation along dimension 0 of an ensemble frame:
// for every line (the other dimensions and the trajectories), with
// n(x) = |u(x)|^2, the circular autocorrelation A(r) = sum_x n(x) n(x + r)
// = IFFT(|FFT(n)|^2)(r) / N.  Each workgroup accumulates A(r) and sum_x n(x)
// over the lines it owns (Neumaier-compensated FP64) and writes one partial
// row of N + 1 values; row_autocorr_final_kernel sums the partial rows.
struct RowCorrArgs {
  const void* u;     // N x nlines complex values
  int64_t nlines;
  double* partial;   // gridDim.x rows of N + 1 values: A(r) sums, then sum n
  const void* tw;
  int tws;
};

template <typename R, int N, int W>
__global__ void __launch_bounds__(W * (N / kE)) row_autocorr_kernel(RowCorrArgs p) {
  constexpr int T = N / kE;
  constexpr int LS = lds_line_stride(N);
  constexpr bool LOCAL = T <= 64;
  static_assert((size_t)W * N + (size_t)W * T <= 2ull * W * LS * sizeof(Cx<R>) / sizeof(double),
                "the LDS of the transforms must hold the final reduction");
  extern __shared__ unsigned char smem_raw[];
  Cx<R>* lds = reinterpret_cast<Cx<R>*>(smem_raw);
  const int tid = threadIdx.x;
  const int l = tid / T, t = tid % T;
  TwReg<R, N> tr;
  load_twiddles<R, N, T>(tr, t, reinterpret_cast<const Cx<R>*>(p.tw), p.tws);
  double acc[kE], cmp[kE];
#pragma unroll
  for (int m = 0; m < kE; ++m) acc[m] = cmp[m] = 0.0;
  double s1 = 0.0, c1 = 0.0;
  const Cx<R>* u = reinterpret_cast<const Cx<R>*>(p.u);
  for (int64_t base = (int64_t)blockIdx.x * W; base < p.nlines; base += (int64_t)gridDim.x * W) {
    const int64_t line = base + l;
    const bool valid = line < p.nlines;
    const int64_t ln = valid ? line : p.nlines - 1;
    Cx<R> v[kE];
#pragma unroll
    for (int m = 0; m < kE; ++m) {
      const R n = abs2(u[ln * N + t + m * T]);
      if (valid) neumaier_add(s1, c1, (double)n);
      v[m] = mk<R>(n, (R)0);
    }
    fft_lines<R, N, T, 1, false, LOCAL>(as_bundle(v), lds + l * LS, 0, W * LS, t, tr);
#pragma unroll
    for (int m = 0; m < kE; ++m) v[m] = mk<R>(abs2(v[m]), (R)0);
    line_sync<LOCAL>();
    fft_lines<R, N, T, 1, true, LOCAL>(as_bundle(v), lds + l * LS, 0, W * LS, t, tr);
    if (valid) {
#pragma unroll
      for (int m = 0; m < kE; ++m) neumaier_add(acc[m], cmp[m], (double)v[m].re / N);
    }
    line_sync<LOCAL>();
  }
  __syncthreads();
  double* red = reinterpret_cast<double*>(smem_raw);
#pragma unroll
  for (int m = 0; m < kE; ++m) red[l * N + t + m * T] = acc[m] + cmp[m];
  red[W * N + tid] = s1 + c1;
  __syncthreads();
  double* out = p.partial + (int64_t)blockIdx.x * (N + 1);
  if (l == 0) {
#pragma unroll
    for (int m = 0; m < kE; ++m) {
      double s = 0.0, c = 0.0;
      for (int j = 0; j < W; ++j) neumaier_add(s, c, red[j * N + t + m * T]);
      out[t + m * T] = s + c;
    }
  }
  if (tid == 0) {
    double s = 0.0, c = 0.0;
    for (int j = 0; j < W * T; ++j) neumaier_add(s, c, red[W * N + j]);
    out[N] = s + c;
  }
}

// Sums the partial rows and forms, with npts = N * nlines and the mean
// Wigner density nw = (sum n) / npts,
//   g2r[r] = (sum A(r)) / npts - (1 + delta_r0) c/2 (2 nw - c/2),
//   nphys  = nw - c/2
// (the normal-ordered <n(x) n(x + r)> of examples/truncated_wigner.jl's
// ordering correction, averaged over x, the other dimensions and the
// trajectories).
template <int UNUSED = 0>
__global__ void __launch_bounds__(256) row_autocorr_final_kernel(const double* partial, int nblk, int64_t n,
                                                                 int64_t nlines, double comm, double* g2r,
                                                                 double* nphys) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > n) return;
  double s = 0.0, c = 0.0, s1 = 0.0, c1 = 0.0;
  for (int b = 0; b < nblk; ++b) {
    neumaier_add(s, c, partial[(int64_t)b * (n + 1) + r]);
    neumaier_add(s1, c1, partial[(int64_t)b * (n + 1) + n]);
  }
  const double npts = (double)n * (double)nlines;
  const double nw = (s1 + c1) / npts;
  if (r == n) {
    nphys[0] = nw - comm / 2;
    return;
  }
  const double d = r == 0 ? 2.0 : 1.0;
  g2r[r] = (s + c) / npts - d * (comm / 2) * (2 * nw - comm / 2);
}

struct ResidualArgs {
  const void* a[kMaxM];  // u_{n+1} (real space)
  const void* b[kMaxM];  // u_n
  int64_t npts;
  int M;
  double* partial;       // 4 doubles per block: diff sum/comp, norm sum/comp
};

template <typename R>
__global__ void __launch_bounds__(256) residual_partial_kernel(ResidualArgs r) {
  __shared__ double sh[4][256];
  double sd = 0, cd = 0, sn = 0, cn = 0;
  for (int c = 0; c < r.M; ++c) {
    const Cx<R>* a = reinterpret_cast<const Cx<R>*>(r.a[c]);
    const Cx<R>* b = reinterpret_cast<const Cx<R>*>(r.b[c]);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < r.npts;
         i += (int64_t)gridDim.x * blockDim.x) {
      const Cx<R> x = a[i], y = b[i];
      const double dr = (double)x.re - (double)y.re, di = (double)x.im - (double)y.im;
      neumaier_add(sd, cd, dr * dr + di * di);
      neumaier_add(sn, cn, (double)x.re * (double)x.re + (double)x.im * (double)x.im);
    }
  }
  sh[0][threadIdx.x] = sd;
  sh[1][threadIdx.x] = cd;
  sh[2][threadIdx.x] = sn;
  sh[3][threadIdx.x] = cn;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0, c0 = 0, s1 = 0, c1 = 0;
    for (int i = 0; i < (int)blockDim.x; ++i) {
      neumaier_add(s0, c0, sh[0][i]);
      c0 += sh[1][i];
      neumaier_add(s1, c1, sh[2][i]);
      c1 += sh[3][i];
    }
    double* o = r.partial + 4 * blockIdx.x;
    o[0] = s0;
    o[1] = c0;
    o[2] = s1;
    o[3] = c1;
  }
}

template <int UNUSED = 0>
__global__ void residual_final_kernel(const double* partial, int nblocks, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s0 = 0, c0 = 0, s1 = 0, c1 = 0;
  for (int i = 0; i < nblocks; ++i) {
    neumaier_add(s0, c0, partial[4 * i]);
    c0 += partial[4 * i + 1];
    neumaier_add(s1, c1, partial[4 * i + 2]);
    c1 += partial[4 * i + 3];
  }
  out[0] = s0 + c0;
  out[1] = s1 + c1;
}

}  // namespace gpe
