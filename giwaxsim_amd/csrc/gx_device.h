// Device-side building blocks shared by the gfx950 kernels of the
// forward-scattering path: exact FP64 floor division (the reference's numpy
// float `//`), order-preserving keys for FP64 min/max atomics, complex FP32
// helpers and an in-wave Stockham FFT (64 lanes per transform).
//
// Everything here is compiled with -ffp-contract=off so that the FP64 index
// arithmetic rounds after every operation exactly like numpy's elementwise
// evaluation; fused multiply-adds are written out explicitly where wanted.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/giwaxsim_amd.h"

#define GX_WAVE 64

namespace gx {

// ------------------------------------------------------------------------------------
// exact floor(a / b) for doubles (b > 0)
//
// numpy's float floor-divide (used for every voxel and bin index in the
// reference: tools/voxelgrids.py:328-329, 497-499; tools/detector.py:217-219)
// is fmod based and returns the exact floor of the real quotient a/b of the
// two doubles, e.g. 1.0 // 0.2 == 4.0 while floor(1.0/0.2) == 5.0.
// Here: q0 = floor(a * binv) is within one of the exact floor; the sign of
// the exact residual a - q*b is obtained from a single-rounding fma (rounding
// never flips the sign of a nonzero real), and q is corrected by one step.
// Valid while |a/b| < 2^50, far beyond any grid index here.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ double floor_div(double a, double b, double binv) {
  double q = floor(a * binv);
  if (fma(-q, b, a) < 0.0) {
    q -= 1.0;
  } else if (fma(-(q + 1.0), b, a) >= 0.0) {
    q += 1.0;
  }
  return q;
}

// numpy `.astype(int)` of a floor-div result (already integral): truncation.
__device__ __forceinline__ long long to_int(double q) { return (long long)q; }

// to_int(floor_div(a, b)) as an index when it lies in [lo, hi) (lo >= 0), else -1: the
// range test on the integral quotient, then one int32 conversion (exact in that range)
__device__ __forceinline__ int floor_div_in(double a, double b, double binv, int lo, int hi) {
  const double q = floor_div(a, b, binv);
  return (q >= (double)lo && q < (double)hi) ? (int)q : -1;
}

// ------------------------------------------------------------------------------------
// order-preserving uint64 keys for doubles (for atomicMin/atomicMax)
// ------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long dkey(double v) {
  unsigned long long u = (unsigned long long)__double_as_longlong(v);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double dkey_inv(unsigned long long k) {
  unsigned long long u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)u);
}

// ------------------------------------------------------------------------------------
// wave-level synchronisation for LDS exchanges between lanes of one wave
// ------------------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, GX_WAVE));
  return v;
}
__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, GX_WAVE));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, GX_WAVE));
  return v;
}

// ------------------------------------------------------------------------------------
// complex arithmetic on float2 (FP32 transforms) and double2 (FP64 transforms)
// ------------------------------------------------------------------------------------
template <typename C> struct Real;
template <> struct Real<float2> { using T = float; };
template <> struct Real<double2> { using T = double; };

__device__ __forceinline__ float2 mk(float a, float b, float2) { return make_float2(a, b); }
__device__ __forceinline__ double2 mk(double a, double b, double2) { return make_double2(a, b); }

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
// multiply by -i
template <typename C> __device__ __forceinline__ C cmul_mi(C a) { return C{a.y, -a.x}; }
// |a|^2
__device__ __forceinline__ float cnorm(float2 a) { return fmaf(a.x, a.x, a.y * a.y); }
__device__ __forceinline__ double cnorm(double2 a) { return fma(a.x, a.x, a.y * a.y); }

// ------------------------------------------------------------------------------------
// forward DFTs of small radix in registers, W_R = exp(-2 pi i / R), natural order
// ------------------------------------------------------------------------------------
template <int R, typename C> struct Dft;

template <typename C> struct Dft<2, C> {
  static __device__ __forceinline__ void run(C* v) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename C> struct Dft<4, C> {
  static __device__ __forceinline__ void run(C* v) {
    C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    C t2 = cadd(v[1], v[3]), t3 = cmul_mi(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <typename C> struct Dft<8, C> {
  static __device__ __forceinline__ void run(C* v) {
    using T = typename Real<C>::T;
    const T c = (T)0.70710678118654752440;
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, C>::run(e);
    Dft<4, C>::run(o);
    C w1 = C{c * (o[1].x + o[1].y), c * (o[1].y - o[1].x)};   // * (c, -c)
    C w2 = cmul_mi(o[2]);                                     // * -i
    C w3 = C{c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y)};  // * (-c, -c)
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);
    v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);
    v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);
    v[7] = csub(e[3], w3);
  }
};

template <typename C> struct Dft<16, C> {
  static __device__ __forceinline__ void run(C* v) {
    using T = typename Real<C>::T;
    // W16^k = (cos(2 pi k/16), -sin(2 pi k/16))
    const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173;
    const T c2 = (T)0.70710678118654752440;
    C e[8], o[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      e[i] = v[2 * i];
      o[i] = v[2 * i + 1];
    }
    Dft<8, C>::run(e);
    Dft<8, C>::run(o);
    const C w[8] = {C{(T)1, (T)0}, C{c1, -s1}, C{c2, -c2}, C{s1, -c1},
                    C{(T)0, (T)-1}, C{-s1, -c1}, C{-c2, -c2}, C{-c1, -s1}};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      C t = (k == 0) ? o[0] : (k == 4 ? cmul_mi(o[4]) : cmul(o[k], w[k]));
      v[k] = cadd(e[k], t);
      v[k + 8] = csub(e[k], t);
    }
  }
};

// ------------------------------------------------------------------------------------
// In-wave Stockham autosort FFT.  One transform of N points per wave; each of
// the 64 lanes holds P = N/64 points in registers.  A stage of radix R with
// Ns = product of the previous radices works on butterflies j = lane + 64 t:
//   inputs  x[j + r N/R]  (r < R), twiddled by W_{Ns R}^{r (j mod Ns)},
//   outputs y[(j / Ns) Ns R + (j mod Ns) + r Ns].
// Data moves between stages through a per-wave LDS buffer of N complex values.
// The first stage's inputs are supplied by the caller in registers
// (v[t R + r] = x[j + r N/R]); the last stage leaves its outputs in
// registers, v[t R + r] = X[j + r N/R] with j = lane + 64 t.
// tw[m] = exp(-2 pi i m / N), m < N, precomputed in FP64 (rounded for FP32).
// ------------------------------------------------------------------------------------
template <int N, int R, int NS, int P, typename C, int L = GX_WAVE>
__device__ __forceinline__ void fft_butterflies(C (&v)[P], const C* __restrict__ tw, int lane) {
  constexpr int NB = P / R;
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    if constexpr (NS > 1) {
      const int j = lane + L * t;
      const int k = j & (NS - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const int m = (r * k) * (N / (NS * R));
        v[t * R + r] = cmul(v[t * R + r], tw[m]);
      }
    }
    Dft<R, C>::run(&v[t * R]);
  }
}

template <int N, int R, int NS, int P, typename C, int L = GX_WAVE>
__device__ __forceinline__ void fft_store(C (&v)[P], C* lds, int lane) {
  constexpr int NB = P / R;
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const int j = lane + L * t;
    const int base = (j / NS) * NS * R + (j & (NS - 1));
#pragma unroll
    for (int r = 0; r < R; ++r) lds[base + r * NS] = v[t * R + r];
  }
}

template <int N, int R, int P, typename C, int L = GX_WAVE>
__device__ __forceinline__ void fft_load(C (&v)[P], const C* lds, int lane) {
  constexpr int NB = P / R;
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const int j = lane + L * t;
#pragma unroll
    for (int r = 0; r < R; ++r) v[t * R + r] = lds[j + r * (N / R)];
  }
}

// Radix plans (first stage radix is R0) of the in-wave transforms (64 lanes).  Plan PL
// = 0 is the default; N = 1024 and 2048 have two more orderings of their radices (the
// frame passes' autotuner picks one per pass and workload); N = 256 and 512 have one.
template <int N, int PL = 0> struct FftPlan;
template <int PL> struct FftPlan<256, PL> { static constexpr int R0 = 4, R1 = 4, R2 = 4, R3 = 4; static constexpr int NST = 4; };
template <int PL> struct FftPlan<512, PL> { static constexpr int R0 = 8, R1 = 8, R2 = 8, R3 = 1; static constexpr int NST = 3; };
template <> struct FftPlan<1024, 0> { static constexpr int R0 = 16, R1 = 16, R2 = 4, R3 = 1; static constexpr int NST = 3; };
template <> struct FftPlan<1024, 1> { static constexpr int R0 = 16, R1 = 8, R2 = 8, R3 = 1; static constexpr int NST = 3; };
template <> struct FftPlan<1024, 2> { static constexpr int R0 = 8, R1 = 16, R2 = 8, R3 = 1; static constexpr int NST = 3; };
template <> struct FftPlan<1024, 3> { static constexpr int R0 = 8, R1 = 8, R2 = 16, R3 = 1; static constexpr int NST = 3; };
template <> struct FftPlan<2048, 0> { static constexpr int R0 = 16, R1 = 16, R2 = 8, R3 = 1; static constexpr int NST = 3; };
template <> struct FftPlan<2048, 1> { static constexpr int R0 = 16, R1 = 8, R2 = 16, R3 = 1; static constexpr int NST = 3; };
template <> struct FftPlan<2048, 2> { static constexpr int R0 = 8, R1 = 16, R2 = 16, R3 = 1; static constexpr int NST = 3; };
// radix plans per grid edge (plan indices 0 .. gx_fft_plans(N) - 1)
constexpr int GX_FFT_PLANS = 4;
__host__ __device__ constexpr int gx_fft_plans(long long N) { return N == 1024 ? 4 : N == 2048 ? 3 : 1; }

template <int N, int PL = 0> struct FftFirst { static constexpr int R = FftPlan<N, PL>::R0; };
template <int N, int PL = 0> struct FftLast {
  static constexpr int R = FftPlan<N, PL>::NST == 4 ? FftPlan<N, PL>::R3 : FftPlan<N, PL>::R2;
};

// Runs the whole transform.  On entry v holds the first-stage inputs
// (v[t R0 + r] = x[lane + 64 t + r N/R0]); on exit v holds the outputs in
// last-stage order (v[t RL + r] = X[lane + 64 t + r N/RL]).  `lds` is this
// wave's private buffer of N complex values; it is free again on return.
template <int N, typename C, int PLAN = 0>
__device__ __forceinline__ void fft_wave(C (&v)[N / GX_WAVE], C* lds, const C* __restrict__ tw, int lane) {
  constexpr int P = N / GX_WAVE;
  using PL = FftPlan<N, PLAN>;
  constexpr int R0 = PL::R0, R1 = PL::R1, R2 = PL::R2, R3 = PL::R3;
  fft_butterflies<N, R0, 1, P, C>(v, tw, lane);
  fft_store<N, R0, 1, P, C>(v, lds, lane);
  wave_sync();
  fft_load<N, R1, P, C>(v, lds, lane);
  wave_sync();
  fft_butterflies<N, R1, R0, P, C>(v, tw, lane);
  fft_store<N, R1, R0, P, C>(v, lds, lane);
  wave_sync();
  fft_load<N, R2, P, C>(v, lds, lane);
  wave_sync();
  fft_butterflies<N, R2, R0 * R1, P, C>(v, tw, lane);
  if constexpr (PL::NST == 4) {
    fft_store<N, R2, R0 * R1, P, C>(v, lds, lane);
    wave_sync();
    fft_load<N, R3, P, C>(v, lds, lane);
    wave_sync();
    fft_butterflies<N, R3, R0 * R1 * R2, P, C>(v, tw, lane);
  }
}

// ------------------------------------------------------------------------------------
// Banded variant of fft_wave for the three-stage plans (N = 512, 1024, 2048), used by
// the forward row and column passes, which keep only the outputs whose signed
// frequency lies in a band [klo, khi] around 0.
//
// * LDS layout.  fft_wave's first exchange stores butterfly j's R0 outputs at
//   j R0 + r, so the lanes of one store are R0 complex apart and meet in the same
//   LDS banks.  Here element i of the first exchange sits at i + i / R0 and element i
//   of the second at i + R0 (i / (R0 R1)): for every store and load of both exchanges
//   the address is a lane-dependent base plus a compile-time offset (immediate
//   offsets, as before) and consecutive lanes fall in different banks.  The buffer
//   holds N + N / min(R0, R1) values (fft_band_buf).  Values and operation order are
//   fft_wave's, so results are bitwise the same.
// * Last stage.  Butterfly j produces X[j + r N/RL], r < RL; with the band narrower
//   than N / RL only r = 0 (k = j) and r = RL - 1 (k = j - N/RL, signed) can be kept.
//   For each group of 64 butterflies (wave-uniform) the needed outputs are computed:
//   all of them through Dft<RL> when a middle r is needed, else only r = 0 and / or
//   RL - 1, with Dft<RL>'s own additions in its own order (bitwise the same values).
//   Outputs that are not computed are left undefined; callers use only k in the band.
// ------------------------------------------------------------------------------------
template <int N, int PLAN = 0> struct FftBand {
  using PL = FftPlan<N, PLAN>;
  static_assert(PL::NST == 3, "banded transforms use the three-stage plans");
  static constexpr int R0 = PL::R0, R1 = PL::R1, RL = PL::R2;
  static constexpr int PAD = (N / R0 > N / R1) ? N / R0 : N / R1;
  static constexpr int BUF = N + PAD;  // complex values per wave
};
// Buffer of every plan of N (kernels size their LDS once for all plans)
template <int N> constexpr int fft_band_buf() {
  return N + N / 8;  // >= N + max(N / R0, N / R1) for radices of 8 and 16
}

// r = 0 and r = R - 1 outputs of Dft<R> (R = 4, 8) in Dft<R>'s operation order
template <int R, typename C> struct DftEnds;
template <typename C> struct DftEnds<4, C> {
  static __device__ __forceinline__ void run(C *v, bool want0, bool wantL) {
    if (want0) {
      const C t0 = cadd(v[0], v[2]), t2 = cadd(v[1], v[3]);
      const C x0 = cadd(t0, t2);
      if (wantL) {
        const C t1 = csub(v[0], v[2]), t3 = cmul_mi(csub(v[1], v[3]));
        v[3] = csub(t1, t3);
      }
      v[0] = x0;
    } else if (wantL) {
      const C t1 = csub(v[0], v[2]), t3 = cmul_mi(csub(v[1], v[3]));
      v[3] = csub(t1, t3);
    }
  }
};
template <typename C> struct DftEnds<8, C> {
  static __device__ __forceinline__ void run(C *v, bool want0, bool wantL) {
    using T = typename Real<C>::T;
    const T c = (T)0.70710678118654752440;
    // Dft<8>: e = Dft<4>(v0, v2, v4, v6), o = Dft<4>(v1, v3, v5, v7);
    // out0 = e0 + o0, out7 = e3 - w3 with w3 = (c (o3.y - o3.x), -c (o3.x + o3.y))
    C e0, o0, e3, o3;
    if (want0) {
      e0 = cadd(cadd(v[0], v[4]), cadd(v[2], v[6]));
      o0 = cadd(cadd(v[1], v[5]), cadd(v[3], v[7]));
    }
    if (wantL) {
      e3 = csub(csub(v[0], v[4]), cmul_mi(csub(v[2], v[6])));
      o3 = csub(csub(v[1], v[5]), cmul_mi(csub(v[3], v[7])));
    }
    if (want0) v[0] = cadd(e0, o0);
    if (wantL) {
      const C w3 = C{c * (o3.y - o3.x), -c * (o3.x + o3.y)};
      v[7] = csub(e3, w3);
    }
  }
};

template <int N, typename C, int PLAN = 0>
__device__ __forceinline__ void fft_wave_band(C (&v)[N / GX_WAVE], C *lds, const C *__restrict__ tw, int lane,
                                              int klo, int khi) {
  constexpr int P = N / GX_WAVE;
  using FB = FftBand<N, PLAN>;
  static_assert(FB::BUF <= fft_band_buf<N>(), "plan buffer exceeds the per-wave LDS");
  constexpr int R0 = FB::R0, R1 = FB::R1, RL = FB::RL;
  constexpr int NB0 = P / R0, NB1 = P / R1, NBL = P / RL;
  // stage 1 (radix R0, no twiddles); element i -> i + i / R0
  fft_butterflies<N, R0, 1, P, C>(v, tw, lane);
#pragma unroll
  for (int t = 0; t < NB0; ++t) {
    const int j = lane + GX_WAVE * t;
    C *p = lds + j * (R0 + 1);
#pragma unroll
    for (int r = 0; r < R0; ++r) p[r] = v[t * R0 + r];
  }
  wave_sync();
  {
    const C *p = lds + lane + lane / R0;
#pragma unroll
    for (int t = 0; t < NB1; ++t)
#pragma unroll
      for (int r = 0; r < R1; ++r)
        v[t * R1 + r] = p[(GX_WAVE * t) + (GX_WAVE * t) / R0 + r * (N / R1 + N / (R0 * R1))];
  }
  wave_sync();
  // stage 2 (radix R1, Ns = R0); element i -> i + R0 (i / (R0 R1))
  fft_butterflies<N, R1, R0, P, C>(v, tw, lane);
#pragma unroll
  for (int t = 0; t < NB1; ++t) {
    const int j = lane + GX_WAVE * t;
    C *p = lds + (j / R0) * (R0 * R1 + R0) + (j & (R0 - 1));
#pragma unroll
    for (int r = 0; r < R1; ++r) p[r * R0] = v[t * R1 + r];
  }
  wave_sync();
  {
    const C *p = lds + lane;
#pragma unroll
    for (int t = 0; t < NBL; ++t)
#pragma unroll
      for (int r = 0; r < RL; ++r) v[t * RL + r] = p[GX_WAVE * t + r * (N / RL + R0)];
  }
  wave_sync();
  // stage 3 (radix RL, Ns = R0 R1 = N / RL), outputs pruned to the band
  constexpr int NS = R0 * R1;
#pragma unroll
  for (int t = 0; t < NBL; ++t) {
    const int j0 = GX_WAVE * t;  // first butterfly of this group (j = j0 .. j0 + 63)
    // signed k of output r: r = 0: j; r = RL - 1: j - NS; 1 <= r <= RL/2: j + r NS >= j0 + NS;
    // RL/2 < r <= RL - 2: j + r NS - N <= j0 + 63 - 2 NS
    const bool want0 = j0 <= khi && j0 + GX_WAVE - 1 >= klo;
    const bool wantL = j0 - NS <= khi && j0 + GX_WAVE - 1 - NS >= klo;
    const bool wantMid = khi >= j0 + NS || klo <= j0 + GX_WAVE - 1 - 2 * NS;
    if (!(want0 || wantL || wantMid)) continue;
    const int j = lane + j0;
    const int k = j & (NS - 1);
#pragma unroll
    for (int r = 1; r < RL; ++r) v[t * RL + r] = cmul(v[t * RL + r], tw[(r * k) * (N / (NS * RL))]);
    if constexpr (RL == 4 || RL == 8) {
      if (wantMid)
        Dft<RL, C>::run(&v[t * RL]);
      else
        DftEnds<RL, C>::run(&v[t * RL], want0, wantL);
    } else {
      Dft<RL, C>::run(&v[t * RL]);
    }
  }
}

// Per-wave LDS (complex values) of fft_wave_kept and the transform itself: the banded,
// padded transform for the three-stage plans, fft_wave (all outputs) for N = 256.
template <int N> constexpr int fft_kept_buf() {
  if constexpr (FftPlan<N>::NST == 3) return fft_band_buf<N>();
  return N;
}
template <int N, typename C, int PLAN = 0>
__device__ __forceinline__ void fft_wave_kept(C (&v)[N / GX_WAVE], C *lds, const C *__restrict__ tw, int lane,
                                              int klo, int khi) {
  if constexpr (FftPlan<N, PLAN>::NST == 3)
    fft_wave_band<N, C, PLAN>(v, lds, tw, lane, klo, khi);
  else
    fft_wave<N, C, PLAN>(v, lds, tw, lane);
}

// ------------------------------------------------------------------------------------
// Block-level Stockham transforms of length M over GX_BLK_FFT = 256 threads
// (P = M / 256 points per thread), used by the chirp-z (Bluestein) path for
// grid edges that are not powers of two.  The plans have equal first and last
// radix, so a transform's output registers are already laid out as the next
// transform's inputs (index j + r M/R with j = tid + 256 t).  Block barriers
// separate the stages; the caller must reach this with the whole block.
// ------------------------------------------------------------------------------------
#define GX_BLK_FFT 256
template <int M> struct BlkPlan;
template <> struct BlkPlan<1024> { static constexpr int R = 4; static constexpr int NST = 5; };   // 4^5
template <> struct BlkPlan<2048> { static constexpr int R = 8; static constexpr int NST = 4; };   // 8 4 8 8
template <> struct BlkPlan<4096> { static constexpr int R = 16; static constexpr int NST = 3; };  // 16^3

template <int M, typename C>
__device__ __forceinline__ void fft_block(C (&v)[M / GX_BLK_FFT], C* lds, const C* __restrict__ tw, int tid) {
  constexpr int P = M / GX_BLK_FFT;
  constexpr int L = GX_BLK_FFT;
  if constexpr (M == 1024) {
    fft_butterflies<M, 4, 1, P, C, L>(v, tw, tid);
    fft_store<M, 4, 1, P, C, L>(v, lds, tid); __syncthreads();
    fft_load<M, 4, P, C, L>(v, lds, tid); __syncthreads();
    fft_butterflies<M, 4, 4, P, C, L>(v, tw, tid);
    fft_store<M, 4, 4, P, C, L>(v, lds, tid); __syncthreads();
    fft_load<M, 4, P, C, L>(v, lds, tid); __syncthreads();
    fft_butterflies<M, 4, 16, P, C, L>(v, tw, tid);
    fft_store<M, 4, 16, P, C, L>(v, lds, tid); __syncthreads();
    fft_load<M, 4, P, C, L>(v, lds, tid); __syncthreads();
    fft_butterflies<M, 4, 64, P, C, L>(v, tw, tid);
    fft_store<M, 4, 64, P, C, L>(v, lds, tid); __syncthreads();
    fft_load<M, 4, P, C, L>(v, lds, tid); __syncthreads();
    fft_butterflies<M, 4, 256, P, C, L>(v, tw, tid);
  } else if constexpr (M == 2048) {
    fft_butterflies<M, 8, 1, P, C, L>(v, tw, tid);
    fft_store<M, 8, 1, P, C, L>(v, lds, tid); __syncthreads();
    fft_load<M, 4, P, C, L>(v, lds, tid); __syncthreads();
    fft_butterflies<M, 4, 8, P, C, L>(v, tw, tid);
    fft_store<M, 4, 8, P, C, L>(v, lds, tid); __syncthreads();
    fft_load<M, 8, P, C, L>(v, lds, tid); __syncthreads();
    fft_butterflies<M, 8, 32, P, C, L>(v, tw, tid);
    fft_store<M, 8, 32, P, C, L>(v, lds, tid); __syncthreads();
    fft_load<M, 8, P, C, L>(v, lds, tid); __syncthreads();
    fft_butterflies<M, 8, 256, P, C, L>(v, tw, tid);
  } else {
    static_assert(M == 4096, "block transform lengths: 1024, 2048, 4096");
    fft_butterflies<M, 16, 1, P, C, L>(v, tw, tid);
    fft_store<M, 16, 1, P, C, L>(v, lds, tid); __syncthreads();
    fft_load<M, 16, P, C, L>(v, lds, tid); __syncthreads();
    fft_butterflies<M, 16, 16, P, C, L>(v, tw, tid);
    fft_store<M, 16, 16, P, C, L>(v, lds, tid); __syncthreads();
    fft_load<M, 16, P, C, L>(v, lds, tid); __syncthreads();
    fft_butterflies<M, 16, 256, P, C, L>(v, tw, tid);
  }
}

// Zoomed chirp-z band transform on a whole block (see the chirp-z passes): the circular
// convolution IFFT_M(FFT_M(v) K) of the block's length-M input v with the kernel spectrum
// K (already divided by M), as conj(FFT(conj(FFT(v) K))); the output is in v's layout.
template <int M, typename C>
__device__ __forceinline__ void chirpz_block(C (&v)[M / GX_BLK_FFT], C *lds, const C *__restrict__ K,
                                             const C *__restrict__ tw, int tid) {
  constexpr int P = M / GX_BLK_FFT;
  constexpr int R = BlkPlan<M>::R;
  fft_block<M, C>(v, lds, tw, tid);
#pragma unroll
  for (int t = 0; t < P / R; ++t)
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = tid + GX_BLK_FFT * t + r * (M / R);
      const C x = cmul(v[t * R + r], K[k]);
      v[t * R + r] = C{x.x, -x.y};
    }
  __syncthreads();
  fft_block<M, C>(v, lds, tw, tid);
#pragma unroll
  for (int q = 0; q < P; ++q) v[q] = C{v[q].x, -v[q].y};
}

// ------------------------------------------------------------------------------------
// XCD-aware tile order.  The workgroups of a launch are dispatched round-robin over
// the 8 XCDs (block b runs on XCD b % 8), each with its own L2.  xcd_tile maps the
// linear block id b of an n-block launch to a work item so that each XCD receives one
// contiguous range of work items; with work items numbered row-block major and frame
// minor, the blocks resident on one XCD at a time are the frames of one or two row
// blocks, which then share those rows' atoms in that XCD's L2.
// ------------------------------------------------------------------------------------
constexpr int GX_XCDS = 8;
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t n) {
  const int64_t q = n / GX_XCDS, r = n % GX_XCDS;
  const int64_t x = b % GX_XCDS, slot = b / GX_XCDS;
  return x < r ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

// (row block, frame) of block b of a launch over nrb row blocks x nf frames (a 1-D grid
// of nrb nf blocks).  Each XCD takes one contiguous range of work items (xcd_tile),
// numbered in tiles of GX_ROW_TILE neighbouring row blocks x all frames with the row
// block fastest: the blocks resident on one XCD at a time are a few frames of those row
// blocks, which share the rows' atoms and the frame-buffer lines of neighbouring rows
// (a 128-byte line holds 16 rows of a z-contiguous column) in that XCD's L2.
#ifndef GX_ROW_TILE
#define GX_ROW_TILE 8
#endif
struct RowFrame {
  int rb, fb;
};
__device__ __forceinline__ RowFrame row_frame_tile(int64_t b, int64_t nrb, int64_t nf) {
  const int64_t item = xcd_tile(b, nrb * nf);
  const int64_t span = (int64_t)GX_ROW_TILE * nf;
  const int64_t t = item / span, rem = item - t * span;
  const int64_t tw = min((int64_t)GX_ROW_TILE, nrb - t * GX_ROW_TILE);  // row blocks in this tile
  const int64_t f = rem / tw;
  return {(int)(t * GX_ROW_TILE + (rem - f * tw)), (int)f};
}

// ------------------------------------------------------------------------------------
// rectangular_collapse_lengths (tools/voxelgrids.py:238-309) at one x value,
// from the per-frame record prepared by the host planner (same branch order
// and the same FP64 operations as the reference's per-x loop).
// ------------------------------------------------------------------------------------
__device__ __forceinline__ double collapse_length(const double *fp, double x) {
  const int mode = (int)fp[GX_FP_LEN_MODE];
  if (mode == 1) return (x < fp[GX_FP_HOR]) ? fp[GX_FP_VER] : 0.0;
  if (mode == 2) return (x < fp[GX_FP_VER]) ? fp[GX_FP_HOR] : 0.0;
  if (x == 0.0) return 0.0;
  if (x < fp[GX_FP_STOP1]) {
    double tot = fp[GX_FP_A] + x * fp[GX_FP_TP];
    double miss = (fp[GX_FP_VC] - x) * fp[GX_FP_TT];
    return tot - miss;
  }
  if (x <= fp[GX_FP_STOP2]) return fp[GX_FP_MID];
  if (x < fp[GX_FP_S12]) {
    double rem = fp[GX_FP_HOR] - ((x - fp[GX_FP_VC]) / fp[GX_FP_CP]);
    return rem / fp[GX_FP_CT];
  }
  return 0.0;
}

// ------------------------------------------------------------------------------------
// Pieces of the row and column passes shared by the forward and the
// differentiable variant (same arithmetic in both, by construction).
//
// stage_row_band: the kept fftshifted column band [jlo, jlo + wf) of one row
// transform v (last-stage register layout) into stage[w * SP + rb] (SP, the tile's
// row stride, is odd so that the lanes' stores spread over the LDS banks).  A row
// outside the filled z range is the constant A = avg_voxel_f * max_voxels (only
// k = 0 is non-zero); inside, with fill or smoothing, the frame's background
// spectrum U is added: o = mz (v + U) (+ A N at k = 0 when smoothing), in the
// transform's precision.
// ------------------------------------------------------------------------------------
template <int N, typename C, int PLAN = 0>
__device__ __forceinline__ void stage_row_band(const C (&v)[N / GX_WAVE], int lane, int jlo, int wf, bool inside,
                                               const gx_frames &d, double mz, const C *tspec, C *stage, int SP,
                                               int rb) {
  using T = typename Real<C>::T;
  constexpr int P = N / GX_WAVE, RL = FftLast<N, PLAN>::R;
  const bool dense = d.fill_bkg || d.smooth;
  const T mzt = (T)mz;
  const C a0 = C{(T)(d.avg_re * d.max_voxels * (double)N), (T)(d.avg_im * d.max_voxels * (double)N)};
#pragma unroll
  for (int t = 0; t < P / RL; ++t) {
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      const int k = lane + GX_WAVE * t + q * (N / RL);
      const int w = ((k + N / 2) & (N - 1)) - jlo;
      if (w < 0 || w >= wf) continue;
      C o = v[t * RL + q];
      if (!inside) {
        o = (k == 0) ? a0 : C{(T)0, (T)0};
      } else if (dense) {
        const C u = tspec[w];
        T ore = mzt * (o.x + u.x);
        T oim = mzt * (o.y + u.y);
        if (d.smooth && k == 0) {
          ore = ore + a0.x;
          oim = oim + a0.y;
        }
        o = C{ore, oim};
      }
      stage[w * SP + rb] = o;
    }
  }
}

// col_input: element z of a kept column (fftshifted column j = jlo + wcol) for the
// transform along z.  Rows z < Z come from the row pass; rows >= Z hold no atoms:
// with fill they are the constant A (only the k = 0 column, k0), otherwise
// mz[z] U (+ A N at k = 0) when smoothing, else zero.
template <typename C>
__device__ __forceinline__ C col_input(const gx_frames &d, const C *col, int z, bool k0, C u, const double *sz_row) {
  using T = typename Real<C>::T;
  if (z < d.Z) return col[z];
  const double avv_re = d.avg_re * d.max_voxels, avv_im = d.avg_im * d.max_voxels;
  if (d.fill_bkg) return k0 ? C{(T)(avv_re * (double)d.N), (T)(avv_im * (double)d.N)} : C{(T)0, (T)0};
  if (d.smooth) {
    const double mz = sz_row[z];
    double xre = mz * (double)u.x, xim = mz * (double)u.y;
    if (k0) {
      xre = xre + avv_re * (double)d.N;
      xim = xim + avv_im * (double)d.N;
    }
    return C{(T)xre, (T)xim};
  }
  return C{(T)0, (T)0};
}

// avg_voxel_f of the deposited f values, when the caller supplies it on the device
// (gx_frames.avg_q), replaces the descriptor's host value; a kernel that reads the
// average (directly or through stage_row_band / col_input) resolves it once on entry.
#define GX_RESOLVE_AVG(d)          \
  do {                             \
    if ((d).avg_q) {               \
      (d).avg_re = (d).avg_q[0];   \
      (d).avg_im = (d).avg_q[1];   \
    }                              \
  } while (0)

// Transform of the fill constant A N (A = avg_voxel_f max_voxels) on the rows [Lc, N) of
// the k = 0 column at natural frequency kk, which a chirp-z column band shorter than
// N + K - 1 leaves out: A N (N - Lc) at kk = 0, else A N (w^Lc - 1) / (1 - w) with
// w = exp(-2 pi i kk / N), FP64, the angles reduced in integers.  d.avg_* resolved.
__device__ __forceinline__ double2 chirpz_fill_tail(const gx_frames &d, int kk, int Lc) {
  const int N = (int)d.N;
  const double tre = d.avg_re * d.max_voxels * (double)N, tim = d.avg_im * d.max_voxels * (double)N;
  if (kk == 0) return make_double2(tre * (double)(N - Lc), tim * (double)(N - Lc));
  double s1, c1, sL, cL;
  sincos(-2.0 * M_PI * (double)kk / (double)N, &s1, &c1);
  sincos(-2.0 * M_PI * (double)(((int64_t)Lc * kk) % N) / (double)N, &sL, &cL);
  const double nre = cL - 1.0, nim = sL, dre = 1.0 - c1, dim = -s1;
  const double den = dre * dre + dim * dim;
  const double qre = (nre * dre + nim * dim) / den, qim = (nim * dre - nre * dim) / den;
  return make_double2(tre * qre - tim * qim, tre * qim + tim * qre);
}

// numpy linspace element i of n: i * step + start, the last one exactly stop.
__device__ __forceinline__ double lin_axis(int64_t i, int64_t n, double start, double step, double stop) {
  return (i == n - 1) ? stop : (double)i * step + start;
}

// (f0(|q|) / z_elem)^2 from the nine Cromer-Mann coefficients a1 b1 .. a4 b4 c
// (add_f0_q_3d, tools/voxelgrids.py:828-857; f0 of tools/utilities.py:320-340).
__device__ __forceinline__ double f0_weight(const double *aff, double q2, double z_elem, double c16pi2) {
  double f0 = aff[0] * exp(-aff[1] * q2 / c16pi2);
  f0 = f0 + aff[2] * exp(-aff[3] * q2 / c16pi2);
  f0 = f0 + aff[4] * exp(-aff[5] * q2 / c16pi2);
  f0 = f0 + aff[6] * exp(-aff[7] * q2 / c16pi2);
  f0 = f0 + aff[8];
  const double fn = f0 / z_elem;
  return fn * fn;
}

}  // namespace gx
