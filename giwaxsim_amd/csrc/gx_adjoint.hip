// gfx950 kernels of the differentiable phi-frame path and its hand-written adjoint.
//
// The reference (tools/voxelgrids.py rotate_project_fft_coords 311-416 and
// process_file2 464-506) has no gradients.  Its nearest-grid-point deposit is
// piecewise constant in the atom positions, so this variant deposits with
// cloud-in-cell (bilinear) weights; everything after the deposit is the forward
// path's (k_colpass binning, k_bins_finalize, the detector gather).
//
// Forward per batch of frames:
//   gx_frames_stats   min y' per frame (origin of u)
//   k_adj_rowpass<N>  per block of z rows: atoms of rows v-1 (weight tv) and v (weight 1-tv),
//                     bilinear in u, exact fixed-point LDS sums, FFT along y, kept column
//                     band staged and stored transposed (d.inter)
//   k_colpass<N>      (forward path) FFT along z, |F|^2, bins + counts
// Backward per batch (frames recomputed, never stored; the forward's per-frame
// min y' in d.ymin_key is reused):
//   k_adj_rowpass<N>  as above
//   k_adj_colpass<N>  per kept column: FFT along z, cotangent 2 g_bins[bin] F on the
//                     binned pixels, adjoint DFT (conj FFT conj) back to z, in place;
//                     also resets each frame's argmin slot for k_adj_rowgrad
//   k_adj_rowgrad<N>  per block of rows (+1 halo row): adjoint DFT along y of the kept
//                     band, then every atom of the rows gathers its four nodes:
//                     dL/dy', dL/dz, dL/df per (frame, atom) into g_part
//                     and each block's share of sum_a dL/dy'_a and the argmin of y'
//   k_adj_frame_origin origin term: the argmin atom of min y' takes -sum_a dL/dy'_a
//   k_adj_reduce      g_atom += sum_f (sin, cos) dL/dy', dL/dz, dL/df
// The regrouping is stable (original order within a row), sums over atoms run in
// that order with fixed partitions, and the row deposit is exact, so every result
// is bitwise reproducible.
#include "gx_device.h"
#include "gx_host.h"

#include <limits.h>
#include <math.h>

using namespace gx;

namespace {

constexpr int BLK = 256;
constexpr int NWAVE = BLK / GX_WAVE;
constexpr int RED = GX_ADJ_RED_BLOCKS;

// fractional grid coordinate along y of an atom in a frame: u = (y' - ymin) / r,
// u0 = exact floor (the same index arithmetic as the forward path), tu in [0, 1]
// returns the cell u0 when both nodes u0, u0 + 1 lie in a row of N points, else -1
// (then tu is not set); the floor division is the reference's `//`, and the int32
// conversion of the integral quotient is exact in that range
__device__ __forceinline__ int cic_u(double x, double y, double s, double c, double ymin, double r, double rinv,
                                     int N, double &tu) {
  const double yy = (x * s + y * c) - ymin;
  const double q = floor_div(yy, r, rinv);
  if (!(q >= 0.0 && q < (double)(N - 1))) return -1;
  const double t = yy * rinv - q;
  tu = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  return (int)q;
}

template <typename C>
__device__ __forceinline__ C cconj(C a) {
  return C{a.x, -a.y};
}

// block-wide reductions in a fixed order (wave shuffles, then the waves in order)
__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, GX_WAVE);
  const int lane = threadIdx.x & (GX_WAVE - 1), w = threadIdx.x / GX_WAVE;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  for (int i = 0; i < NWAVE; ++i) r += sh[i];
  return r;
}

__device__ __forceinline__ int block_min_int(int v, int *sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, GX_WAVE));
  const int lane = threadIdx.x & (GX_WAVE - 1), w = threadIdx.x / GX_WAVE;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  int r = INT_MAX;
  for (int i = 0; i < NWAVE; ++i) r = min(r, sh[i]);
  return r;
}

// NV block-wide sums at once, in block_sum's order; sh holds NWAVE * NV doubles
template <int NV>
__device__ __forceinline__ void block_sum_multi(double (&v)[NV], double *sh) {
#pragma unroll
  for (int i = 0; i < NV; ++i)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off, GX_WAVE);
  const int lane = threadIdx.x & (GX_WAVE - 1), w = threadIdx.x / GX_WAVE;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) sh[w * NV + i] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double r = 0.0;
    for (int k = 0; k < NWAVE; ++k) r += sh[k * NV + i];
    v[i] = r;
  }
}

// ------------------------------------------------------------------------------------
// preparation
// ------------------------------------------------------------------------------------
// Deposit exponent of row v: the row receives f w from atoms of rows v-1 and v with
// w <= 1, so 2^S sum |f| + (rounding) stays below 2^62 (int64 accumulators, FP64
// transforms) or 2^31 - 1 (int32, FP32 transforms).
__device__ __forceinline__ int adj_row_exp(const gx_frames &d, const gx_adjoint &g, int v) {
  double bound = 0.0;
  double cnt = 0.0;
  for (int e = 0; e < g.n_elem; ++e) {
    const double fm = fmax(fabs(g.f_table[2 * e]), fabs(g.f_table[2 * e + 1]));
    double c = (double)g.ecnt[(int64_t)v * g.n_elem + e];
    if (v > 0) c += (double)g.ecnt[(int64_t)(v - 1) * g.n_elem + e];
    bound += c * fm;
    cnt += c;
  }
  int S = 0;
  if (bound > 0.0) {
    const double room = d.fft_f64 ? 4611686018427387904.0 : 2147483647.0;  // int64 / int32 accumulators
    S = ilogb((room - 2.0 * cnt - 2.0) / bound);
    if (S > 200) S = 200;
    if (S < -200) S = -200;
  }
  return S;
}

// Per-call preparation in one launch: the rows' deposit exponents, the inverse of the
// regrouping permutation, zeroed per-atom gradient accumulators and avg_voxel_f slots.
__global__ __launch_bounds__(BLK) void k_adj_prepare(gx_frames d, gx_adjoint g) {
  const int64_t t0 = blockIdx.x * (int64_t)BLK + threadIdx.x, stride = (int64_t)gridDim.x * BLK;
  const int64_t n = d.n_atoms;
  if (t0 < 2 && g.g_avg) g.g_avg[t0] = 0.0;
  if (g.e_sum && t0 < 2 * g.n_elem) g.e_sum[t0] = 0.0;
  for (int64_t v = t0; v < d.N; v += stride) g.row_exp[v] = adj_row_exp(d, g, (int)v);
  for (int64_t a = t0; a < n; a += stride) {
    g.ainv[g.aperm[a]] = (int32_t)a;
#pragma unroll
    for (int k = 0; k < 5; ++k) g.g_atom[k * n + a] = 0.0;
  }
}

__global__ __launch_bounds__(BLK) void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) p[i] = v;
}

// ------------------------------------------------------------------------------------
// forward row pass with the cloud-in-cell deposit
// ------------------------------------------------------------------------------------
// Accumulator of the fixed-point deposit: int32 pairs under FP32 transforms (the
// quantum, 2^-31 of the row bound, is below FP32 rounding), int64 under FP64.  A
// row's accumulators take exactly N * sizeof(C) bytes, the FFT buffer of the row.
template <typename C> struct DepositAcc;
template <> struct DepositAcc<float2> { using T = int32_t; using U = unsigned int; static constexpr int BITS = 31; };
template <> struct DepositAcc<double2> { using T = long long; using U = unsigned long long; static constexpr int BITS = 62; };

// a deposit term rounded to the nearest integer (ties to even, as llrint) in the
// accumulator's type: the row exponent keeps every term inside it, so the int32 case
// converts directly (two instructions instead of llrint's int64 conversion)
template <typename AT> __device__ __forceinline__ AT fx_round(double x);
template <> __device__ __forceinline__ int32_t fx_round<int32_t>(double x) { return (int32_t)rint(x); }
template <> __device__ __forceinline__ long long fx_round<long long>(double x) { return llrint(x); }

#ifndef GX_ADJ_COL_PLAN
#define GX_ADJ_COL_PLAN 0
#endif

// Atoms per thread in flight in the atom loops of the row-wise kernels: the loops are
// bound by the latency of the atoms' loads otherwise.
constexpr int AU_UNROLL = 4;

// Rows per block of the differentiable row pass: RB rows of accumulators in 32 KB of
// LDS (at most 8), so that three blocks share a CU; at N = 1024 FP32 two rows (16 KB:
// more blocks per CU outweigh the atoms of row z0 - 1 read by twice as many blocks: the
// forward of the 110k-atom slab's 581 frames 2.08 -> 1.97 ms; eight rows: 2.42 ms), at
// N = 512 FP32 four (the 5.6k-atom graphite forward 0.217 -> 0.210 ms; two: 0.225 ms)
template <int N, typename C>
struct RowPassShape {
  static constexpr int FIT = 32768 / (N * (int)sizeof(C));
  static constexpr int RB = (N == 1024 && sizeof(C) == 8) ? 2 : (N == 512 && sizeof(C) == 8) ? 4 : (FIT < 1 ? 1 : (FIT < 8 ? FIT : 8));
};

// One block = RB consecutive z rows of one frame.  All 256 threads deposit the atoms
// of rows z0-1 .. z0+RB-1 once (an atom of row v0 reaches rows v0 and v0+1), then
// each wave transforms its rows.
template <int N, typename C>
__global__ __launch_bounds__(BLK) void k_adj_rowpass(gx_frames d, gx_adjoint g, int64_t f0, int64_t nf) {
  using T = typename Real<C>::T;
  using AT = typename DepositAcc<C>::T;
  using AU = typename DepositAcc<C>::U;
  constexpr int P = N / GX_WAVE;
  constexpr int RB = RowPassShape<N, C>::RB;
  constexpr int R0 = FftFirst<N>::R;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  AU *accs = reinterpret_cast<AU *>(smem_raw);                 // [RB][2][N]
  constexpr int SP = RB + 1;                                    // staging row stride (odd)
  C *stage = reinterpret_cast<C *>(accs + (size_t)RB * 2 * N);  // [W][SP]
  C *tspec = stage + d.W * SP;                                  // [W] background spectrum
  double *ftab = reinterpret_cast<double *>(tspec + d.W);       // [n_elem][2]
  __shared__ double rsc[RB];
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int wave = threadIdx.x / GX_WAVE;

  const RowFrame rf = row_frame_tile(blockIdx.x, (d.Z + RB - 1) / RB, nf);
  const int fb = rf.fb;
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wf = jhi - jlo + 1;
  if (wf <= 0) return;  // block-uniform
  if (wf > d.W) {
    if (threadIdx.x == 0 && rf.rb == 0) atomicAdd(&d.status[1], 1);
    return;
  }
  const int z0 = rf.rb * RB;
  for (int i = threadIdx.x; i < 2 * g.n_elem; i += BLK) ftab[i] = g.f_table[i];
  if (d.fill_bkg || d.smooth) {
    const C *spec = reinterpret_cast<const C *>(d.col_spec) + (int64_t)fb * d.W;
    for (int w = threadIdx.x; w < wf; w += BLK) tspec[w] = spec[w];
  }
  if (threadIdx.x < RB) rsc[threadIdx.x] = ldexp(1.0, g.row_exp[min(z0 + (int)threadIdx.x, (int)d.N - 1)]);
  for (int i = threadIdx.x; i < RB * 2 * N; i += BLK) accs[i] = 0;
  __syncthreads();

  const double s = fp[GX_FP_SIN], c = fp[GX_FP_COS];
  const double ymin = dkey_inv(d.ymin_key[3 * fb]);
  const int32_t *bnd = d.idx_bounds + 4 * fb;
  const int ylo = bnd[0], yhi = bnd[1], zlo = bnd[2], zhi = bnd[3];
  // nodes in columns the fill overwrites keep no deposit: ya <= y < yb survive
  const int ya = d.fill_bkg ? (ylo == INT_MAX ? N : ylo + 1) : 0;
  const int yb = d.fill_bkg ? (yhi == INT_MIN ? 0 : yhi) : N;
  const double r = d.r, rinv = 1.0 / d.r;

  // 1) deposit: atoms of rows z0-1 .. z0+RB-1 into the block's rows, AU_UNROLL atoms
  //    per thread in flight (every load of a group issued before the first is used)
  const int a_lo = d.row_off[z0 > 0 ? z0 - 1 : 0], a_hi = d.row_off[min(z0 + RB, (int)d.Z)];
  for (int a = a_lo + threadIdx.x; a < a_hi; a += BLK * AU_UNROLL) {
    double xs[AU_UNROLL], ys[AU_UNROLL], tvs[AU_UNROLL];
    int zr[AU_UNROLL], es[AU_UNROLL];
#pragma unroll
    for (int k = 0; k < AU_UNROLL; ++k) {
      const int ak = a + k * BLK;
      const bool ok = ak < a_hi;
      xs[k] = ok ? d.ax[ak] : 0.0;
      ys[k] = ok ? d.ay[ak] : 0.0;
      tvs[k] = ok ? g.afz[ak] : 0.0;
      zr[k] = ok ? d.azrow[ak] : 0;
      es[k] = ok ? g.aelem[ak] : 0;
    }
#pragma unroll
    for (int k = 0; k < AU_UNROLL; ++k) {
      if (a + k * BLK >= a_hi) break;
      const int lv = zr[k] - z0;  // -1 .. RB-1
      double tu = 0.0;
      const int u0 = cic_u(xs[k], ys[k], s, c, ymin, r, rinv, N, tu);
      if (u0 < 0) {
        atomicAdd(&d.status[2], 1);
        continue;
      }
      const double tv = tvs[k];
      const double fr = ftab[2 * es[k]], fi = ftab[2 * es[k] + 1];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int row = lv + h;
        if (row < 0 || row >= RB) continue;
        const double wv = h ? tv : 1.0 - tv;
        // the power-of-two row scale sc goes into the weights: f (w sc) rounds to
        // exactly (f w) sc
        const double sc = rsc[row];
        const double w0 = (wv * (1.0 - tu)) * sc, w1 = (wv * tu) * sc;
        AU *ar = accs + (size_t)row * 2 * N;
        atomicAdd(&ar[u0], (AU)fx_round<AT>(fr * w0));
        atomicAdd(&ar[N + u0], (AU)fx_round<AT>(fi * w0));
        atomicAdd(&ar[u0 + 1], (AU)fx_round<AT>(fr * w1));
        atomicAdd(&ar[N + u0 + 1], (AU)fx_round<AT>(fi * w1));
      }
    }
  }
  __syncthreads();

  // 2) per row: node mask (smoothed edge x filled region), FFT along y, kept band
  T my[P];
#pragma unroll
  for (int t = 0; t < P / R0; ++t)
#pragma unroll
    for (int q = 0; q < R0; ++q) {
      const int y = lane + GX_WAVE * t + q * (N / R0);
      const T m = d.smooth ? (T)d.col_sy[(int64_t)fb * N + y] : (T)1;
      my[t * R0 + q] = (y >= ya && y < yb) ? m : (T)0;
    }
  const C *tw = reinterpret_cast<const C *>(d.twiddle);
  for (int rb = wave; rb < RB; rb += NWAVE) {
    const int v = z0 + rb;
    const bool inside = !d.fill_bkg || (v >= zlo && v < zhi);
    const bool atoms = inside && d.row_off[v + 1] > d.row_off[v > 0 ? v - 1 : 0];
    AU *ar = accs + (size_t)rb * 2 * N;
    C x[P];
    if (atoms) {
      const double inv = 1.0 / rsc[rb];
#pragma unroll
      for (int t = 0; t < P / R0; ++t) {
#pragma unroll
        for (int q = 0; q < R0; ++q) {
          const int y = lane + GX_WAVE * t + q * (N / R0);
          const T m = my[t * R0 + q];
          x[t * R0 + q] = C{(T)((double)(AT)ar[y] * inv) * m, (T)((double)(AT)ar[N + y] * inv) * m};
        }
      }
      wave_sync();
      fft_wave<N, C>(x, reinterpret_cast<C *>(ar), tw, lane);
    } else {
#pragma unroll
      for (int q = 0; q < P; ++q) x[q] = C{(T)0, (T)0};
    }
    const double mz = d.smooth ? d.col_sz[(int64_t)fb * N + v] : 1.0;
    stage_row_band<N, C>(x, lane, jlo, wf, inside, d, mz, tspec, stage, SP, rb);
  }
  __syncthreads();
  C *out = reinterpret_cast<C *>(d.inter) + (int64_t)fb * d.W * d.Z;
  for (int idx = threadIdx.x; idx < wf * RB; idx += BLK) {
    const int w = idx / RB, q = idx - w * RB;
    if (z0 + q < d.Z) out[(int64_t)w * d.Z + z0 + q] = stage[w * SP + q];
  }
}

// ------------------------------------------------------------------------------------
// adjoint of the column pass (in place on d.inter)
// ------------------------------------------------------------------------------------
template <int N, typename C>
__global__ __launch_bounds__(BLK) void k_adj_colpass(gx_frames d, gx_adjoint g, int64_t f0, bool stored) {
  using T = typename Real<C>::T;
  constexpr int P = N / GX_WAVE;
  constexpr int R0 = FftFirst<N>::R, RL = FftLast<N>::R;
  __shared__ C lds[NWAVE][N];
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int wave = threadIdx.x / GX_WAVE;
  const int fb = blockIdx.y;
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wcol = blockIdx.x * NWAVE + wave;
  if (blockIdx.x == 0 && threadIdx.x == 0) g.amin[fb] = INT_MAX;  // argmin slot of k_adj_rowgrad
  if (wcol > jhi - jlo || wcol >= d.W) return;  // wave-uniform; no block barrier below
  const bool dense = d.fill_bkg || d.smooth;
  double2 *apart = reinterpret_cast<double2 *>(g.avg_part);
  C *col = reinterpret_cast<C *>(d.inter) + ((int64_t)fb * d.W + wcol) * d.Z;
  const int j = jlo + wcol;
  // the forward's column -> bin arithmetic (k_colpass)
  const double hx = (j == N - 1) ? fp[GX_FP_HX_STOP] : (double)j * fp[GX_FP_HX_STEP] + fp[GX_FP_HX_START];
  const double hy = (j == N - 1) ? fp[GX_FP_HY_STOP] : (double)j * fp[GX_FP_HY_STEP] + fp[GX_FP_HY_START];
  const double dqinv = 1.0 / d.dq;
  const long long bx = to_int(floor_div(hx - d.qmin, d.dq, dqinv)) - d.c_lo;
  const long long by = to_int(floor_div(hy - d.qmin, d.dq, dqinv)) - d.c_lo;
  if (bx < 0 || bx >= d.c_n || by < 0 || by >= d.c_n) {  // column never binned: zero cotangent
    for (int z = lane; z < d.Z; z += GX_WAVE) col[z] = C{(T)0, (T)0};
    if (dense && lane == 0) apart[(int64_t)fb * d.W + wcol] = make_double2(0.0, 0.0);
    return;
  }
  const bool k0 = (j == N / 2);
  C u = C{(T)0, (T)0};
  if (!d.fill_bkg && d.smooth) u = reinterpret_cast<const C *>(d.col_spec)[(int64_t)fb * d.W + wcol];
  const double *sz_row = d.col_sz + (int64_t)fb * N;
  C v[P];
  const C *tw = reinterpret_cast<const C *>(d.twiddle);
  C *buf = lds[wave];
  // F at its natural index k into LDS: the kept rows i_lo..i_hi (the only ones read
  // below) from the forward's stored band, or the column's transform recomputed (plan 0,
  // as the forward's column pass with fft_plan_col = 0: bitwise the same values)
  if (stored) {
    const C *fsrc = reinterpret_cast<const C *>(d.f_band) + ((int64_t)fb * d.W + wcol) * (d.i_hi - d.i_lo + 1);
    for (int i = (int)d.i_lo + lane; i <= (int)d.i_hi; i += GX_WAVE) buf[(i + N / 2) & (N - 1)] = fsrc[i - d.i_lo];
  } else {
#pragma unroll
    for (int t = 0; t < P / R0; ++t)
#pragma unroll
      for (int q = 0; q < R0; ++q)
        v[t * R0 + q] = col_input(d, (const C *)col, lane + GX_WAVE * t + q * (N / R0), k0, u, sz_row);
    fft_wave<N, C>(v, buf, tw, lane);
    wave_sync();
#pragma unroll
    for (int t = 0; t < P / RL; ++t)
#pragma unroll
      for (int q = 0; q < RL; ++q) buf[lane + GX_WAVE * t + q * (N / RL)] = v[t * RL + q];
  }
  wave_sync();
  // per k (rolled loop, each lane its own k): conj of the cotangent dL/dF = 2 (dL/dI) F on
  // binned pixels, and with fill or smoothing this column's share
  // sum_k dL/dF_k conj(Pz^[k]) conj(Py^[j]) of dL/d avg (+ max_voxels N^2 dL/dF_0 on the
  // k = 0 column, the pattern's constant)
  const int64_t cbase = (by * d.c_n + bx) * d.c_n;
  const double val = d.q_val;
  const double2 *zhat = reinterpret_cast<const double2 *>(g.avg_spec) + (int64_t)fb * (N + d.W);
  double are = 0.0, aim = 0.0, dc_re = 0.0, dc_im = 0.0;
  // (the dL/d avg sums take only the binned k: every other term is an exact zero)
#pragma unroll 2
  for (int k = lane; k < N; k += GX_WAVE) {
    const int i = (k + N / 2) & (N - 1);
    double gr = 0.0, gi = 0.0;
    if (i >= d.i_lo && i <= d.i_hi) {
      const int kint = (k < N / 2) ? k : k - N;
      const double qz = (((double)kint * val) * 2.0) * M_PI;
      const long long bz = to_int(floor_div(qz - d.qmin, d.dq, dqinv)) - d.c_lo;
      if (bz >= 0 && bz < d.c_n) {
        const double gI = 2.0 * g.g_bins[cbase + bz];
        const C F = buf[k];
        gr = gI * (double)F.x;
        gi = gI * (double)F.y;
        if (dense) {
          const double2 zh = zhat[k];
          are += gr * zh.x + gi * zh.y;
          aim += gi * zh.x - gr * zh.y;
        }
      }
    }
    if (dense && k == 0) {
      dc_re = gr;
      dc_im = gi;
    }
    buf[k] = C{(T)gr, (T)(-gi)};
  }
  if (dense) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      are += __shfl_xor(are, off, GX_WAVE);
      aim += __shfl_xor(aim, off, GX_WAVE);
    }
    if (lane == 0) {
      const double2 yh = zhat[N + wcol];
      double pre = are * yh.x + aim * yh.y, pim = aim * yh.x - are * yh.y;
      if (k0) {
        const double c0 = d.max_voxels * (double)N * (double)N;
        pre = pre + c0 * dc_re;
        pim = pim + c0 * dc_im;
      }
      apart[(int64_t)fb * d.W + wcol] = make_double2(pre, pim);
    }
  }
  wave_sync();
  // the adjoint transform, radix plan ADJ_COL_PLAN
  constexpr int PA = (GX_ADJ_COL_PLAN < gx_fft_plans(N)) ? GX_ADJ_COL_PLAN : 0;
  constexpr int A0 = FftFirst<N, PA>::R, AL = FftLast<N, PA>::R;
#pragma unroll
  for (int t = 0; t < P / A0; ++t)
#pragma unroll
    for (int q = 0; q < A0; ++q) v[t * A0 + q] = buf[lane + GX_WAVE * t + q * (N / A0)];
  wave_sync();
  fft_wave<N, C, PA>(v, buf, tw, lane);
  // dL/dx[z] = conj(FFT(conj(dL/dF)))[z]
#pragma unroll
  for (int t = 0; t < P / AL; ++t) {
#pragma unroll
    for (int q = 0; q < AL; ++q) {
      const int z = lane + GX_WAVE * t + q * (N / AL);
      if (z < d.Z) col[z] = cconj(v[t * AL + q]);
    }
  }
}

// Per frame: Pz^ = FFT of Pz (all k) and Py^ = FFT of Py on the kept columns, FP64,
// one wave each (grid (nf, 2)):  fill: Pz = inz (sz), Py = iny (nm - max_voxels) (sy);
// smooth only: Pz = sz, Py = -max_voxels sy  (sz, sy the smoothed masks when smoothing).
template <int N>
__global__ __launch_bounds__(GX_WAVE) void k_adj_avg_spectra(gx_frames d, gx_adjoint g, int64_t f0) {
  constexpr int P = N / GX_WAVE;
  constexpr int R0 = FftFirst<N>::R, RL = FftLast<N>::R;
  __shared__ double2 lds[N];
  const int lane = threadIdx.x;
  const int fb = blockIdx.x;
  const bool along_y = blockIdx.y == 1;
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wf = max(0, min(jhi - jlo + 1, (int)d.W));
  const int32_t *bnd = d.idx_bounds + 4 * fb;
  const int ylo = bnd[0], yhi = bnd[1], zlo = bnd[2], zhi = bnd[3];
  const int ya = ylo == INT_MAX ? N : ylo + 1, yb = yhi == INT_MIN ? 0 : yhi;
  double2 v[P];
#pragma unroll
  for (int t = 0; t < P / R0; ++t) {
#pragma unroll
    for (int q = 0; q < R0; ++q) {
      const int i = lane + GX_WAVE * t + q * (N / R0);
      double p;
      if (!along_y) {
        p = (!d.fill_bkg || (i >= zlo && i < zhi)) ? 1.0 : 0.0;
        if (d.smooth) p = p * d.col_sz[(int64_t)fb * N + i];
      } else {
        if (d.fill_bkg)
          p = (i >= ya && i < yb) ? (double)d.col_nm[(int64_t)fb * N + i] - d.max_voxels : 0.0;
        else
          p = -d.max_voxels;
        if (d.smooth) p = p * d.col_sy[(int64_t)fb * N + i];
      }
      v[t * R0 + q] = make_double2(p, 0.0);
    }
  }
  fft_wave<N, double2>(v, lds, reinterpret_cast<const double2 *>(d.twiddle64), lane);
  double2 *out = reinterpret_cast<double2 *>(g.avg_spec) + (int64_t)fb * (N + d.W);
#pragma unroll
  for (int t = 0; t < P / RL; ++t) {
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      const int k = lane + GX_WAVE * t + q * (N / RL);
      if (!along_y) {
        out[k] = v[t * RL + q];
      } else {
        const int w = ((k + N / 2) & (N - 1)) - jlo;
        if (w >= 0 && w < wf) out[N + w] = v[t * RL + q];
      }
    }
  }
}

// Block per frame: the frame's kept columns of avg_part summed in a fixed tree order,
// the total left in the frame's column 0.
__global__ __launch_bounds__(BLK) void k_adj_avg_frames(gx_frames d, gx_adjoint g, int64_t f0) {
  __shared__ double sre[NWAVE], sim[NWAVE];
  double2 *apart = reinterpret_cast<double2 *>(g.avg_part) + (int64_t)blockIdx.x * d.W;
  const double *fp = d.fparams + (f0 + blockIdx.x) * GX_FP_STRIDE;
  const int wf = max(0, min((int)fp[GX_FP_JHI] - (int)fp[GX_FP_JLO] + 1, (int)d.W));
  double re = 0.0, im = 0.0;
  for (int w = threadIdx.x; w < wf; w += BLK) {
    re += apart[w].x;
    im += apart[w].y;
  }
  re = block_sum(re, sre);
  im = block_sum(im, sim);
  if (threadIdx.x == 0) apart[0] = make_double2(wf > 0 ? re : 0.0, wf > 0 ? im : 0.0);
}

// One block: g_avg += the frame totals in frame order.
__global__ __launch_bounds__(BLK) void k_adj_avg_reduce(gx_frames d, gx_adjoint g, int64_t nf) {
  __shared__ double2 tot[BLK];
  const double2 *apart = reinterpret_cast<const double2 *>(g.avg_part);
  double re = 0.0, im = 0.0;
  for (int64_t base = 0; base < nf; base += BLK) {
    if (base + threadIdx.x < nf) tot[threadIdx.x] = apart[(base + threadIdx.x) * d.W];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int64_t k = 0; k < BLK && base + k < nf; ++k) {
        re += tot[k].x;
        im += tot[k].y;
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    g.g_avg[0] += re;
    g.g_avg[1] += im;
  }
}

// ------------------------------------------------------------------------------------
// adjoint of the row pass: rows of the grid's cotangent, gathered at the atoms
// ------------------------------------------------------------------------------------
// The gather of the gradient kernels, once `rows` holds the masked cotangent of grid rows
// z0 .. z0 + RB (row stride N): every atom of rows z0 .. z0 + RB - 1 gathers its four
// nodes and writes its dL/dy', dL/dz (and, with more than GX_ADJ_ELEM_REG elements,
// dL/df re, im) for this frame to g_part; the block's share of sum_a dL/dy'_a (fixed
// partition of the regrouped order) goes to red, the argmin of y' to amin and, with
// at most GX_ADJ_ELEM_REG elements, the block's per-element dL/df to e_part.
// With `mzr` given, the rows hold the unmasked cotangent and the gather applies the node
// mask M = (mz inz)[row] (my iny)[y] itself: mzr[h] for the block's rows, my from col_sy.
__device__ __forceinline__ double node_my(const gx_frames &d, const double *sy_row, int ya, int yb, long long y) {
  return (y >= ya && y < yb) ? (d.smooth ? sy_row[y] : 1.0) : 0.0;
}

template <typename C>
__device__ __forceinline__ void rowgrad_gather(const gx_frames &d, const gx_adjoint &g, const double *fp, int fb,
                                               int rb, int nrb, int z0, int RB, int N, const C *rows, const double *ftab,
                                               double *shd, int *shi, const double *mzr = nullptr,
                                               int ya = 0, int yb = 0) {
  using T = typename Real<C>::T;
  const int64_t n = d.n_atoms;
  const double *sy_row = d.col_sy + (int64_t)fb * N;
  const double s = fp[GX_FP_SIN], c = fp[GX_FP_COS];
  const double ymin = dkey_inv(d.ymin_key[3 * fb]);
  const double r = d.r, rinv = 1.0 / d.r;
  const int a_lo = d.row_off[z0], a_hi = d.row_off[min(z0 + RB, (int)d.Z)];
  T *gp = reinterpret_cast<T *>(g.g_part) + (int64_t)fb * 4 * n;
  double su = 0.0;   // this block's share of sum_a dL/dy'_a
  int am = INT_MAX;  // smallest regrouped index attaining min y'
  // dL/df per element in registers (each thread's atoms in increasing order), reduced
  // over the block at the end, when there are at most GX_ADJ_ELEM_REG elements
  const bool esum = g.n_elem <= GX_ADJ_ELEM_REG;
  double ef[2 * GX_ADJ_ELEM_REG];
#pragma unroll
  for (int q = 0; q < 2 * GX_ADJ_ELEM_REG; ++q) ef[q] = 0.0;
  // AU_UNROLL atoms per thread in flight; each thread still takes its atoms in
  // increasing order (the same partial sums)
  for (int a0 = a_lo + threadIdx.x; a0 < a_hi; a0 += BLK * AU_UNROLL) {
    double xs[AU_UNROLL], ys[AU_UNROLL], tvs[AU_UNROLL];
    int zr[AU_UNROLL], es[AU_UNROLL];
#pragma unroll
    for (int k = 0; k < AU_UNROLL; ++k) {
      const int ak = a0 + k * BLK;
      const bool ok = ak < a_hi;
      xs[k] = ok ? d.ax[ak] : 0.0;
      ys[k] = ok ? d.ay[ak] : 0.0;
      tvs[k] = ok ? g.afz[ak] : 0.0;
      zr[k] = ok ? d.azrow[ak] : 0;
      es[k] = ok ? g.aelem[ak] : 0;
    }
#pragma unroll
    for (int k = 0; k < AU_UNROLL; ++k) {
      const int a = a0 + k * BLK;
      if (a >= a_hi) break;
      if ((xs[k] * s + ys[k] * c) == ymin) am = min(am, a);
      const int vr = zr[k] - z0;  // 0 .. RB-1
      double tu = 0.0;
      const int u0 = cic_u(xs[k], ys[k], s, c, ymin, r, rinv, N, tu);
      const double tv = tvs[k];
      const double fr = ftab[2 * es[k]], fi = ftab[2 * es[k] + 1];
      double gu = 0.0, gv = 0.0, gfr = 0.0, gfi = 0.0;
      if (u0 >= 0 && vr >= 0 && vr < RB) {
        const C *r0 = rows + vr * N + u0;
        const C *r1 = r0 + N;
        double g00r = (double)r0[0].x, g00i = (double)r0[0].y;
        double g01r = (double)r0[1].x, g01i = (double)r0[1].y;
        double g10r = (double)r1[0].x, g10i = (double)r1[0].y;
        double g11r = (double)r1[1].x, g11i = (double)r1[1].y;
        if (mzr) {
          const double my0 = node_my(d, sy_row, ya, yb, u0), my1 = node_my(d, sy_row, ya, yb, u0 + 1);
          const double m00 = mzr[vr] * my0, m01 = mzr[vr] * my1;
          const double m10 = mzr[vr + 1] * my0, m11 = mzr[vr + 1] * my1;
          g00r *= m00; g00i *= m00; g01r *= m01; g01i *= m01;
          g10r *= m10; g10i *= m10; g11r *= m11; g11i *= m11;
        }
        // h = Re(conj(dL/dG) f) at the four nodes
        const double h00 = g00r * fr + g00i * fi, h01 = g01r * fr + g01i * fi;
        const double h10 = g10r * fr + g10i * fi, h11 = g11r * fr + g11i * fi;
        const double w00 = (1.0 - tv) * (1.0 - tu), w01 = (1.0 - tv) * tu;
        const double w10 = tv * (1.0 - tu), w11 = tv * tu;
        gu = ((1.0 - tv) * (h01 - h00) + tv * (h11 - h10)) * rinv;
        gv = ((1.0 - tu) * (h10 - h00) + tu * (h11 - h01)) * rinv;
        gfr = ((w00 * g00r + w01 * g01r) + w10 * g10r) + w11 * g11r;
        gfi = ((w00 * g00i + w01 * g01i) + w10 * g10i) + w11 * g11i;
      } else {
        atomicAdd(&d.status[2], 1);
      }
      gp[a] = (T)gu;
      gp[n + a] = (T)gv;
      if (esum) {
        // ef[e] += term for the atom's element, as fma(1, term, ef) there and
        // fma(0, term, ef) elsewhere: the same sums, without a branch per element
#pragma unroll
        for (int q = 0; q < GX_ADJ_ELEM_REG; ++q) {
          const double sel = es[k] == q ? 1.0 : 0.0;
          ef[2 * q] = fma(sel, gfr, ef[2 * q]);
          ef[2 * q + 1] = fma(sel, gfi, ef[2 * q + 1]);
        }
      } else {
        gp[2 * n + a] = (T)gfr;
        gp[3 * n + a] = (T)gfi;
      }
      su += gu;
    }
  }
  su = block_sum(su, shd);
  am = block_min_int(am, shi);
  if (threadIdx.x == 0) {
    g.red[(int64_t)fb * nrb + rb] = su;
    if (am != INT_MAX) atomicMin(&g.amin[fb], am);
  }
  if (esum) {  // block-uniform
    __shared__ double she[NWAVE * 2 * GX_ADJ_ELEM_REG];
    block_sum_multi<2 * GX_ADJ_ELEM_REG>(ef, she);
    if (threadIdx.x == 0) {
      double *ep = g.e_part + ((int64_t)fb * nrb + rb) * 2 * g.n_elem;
      for (int q = 0; q < 2 * g.n_elem; ++q) ep[q] = ef[q];
    }
  }
}

template <int N, typename C>
struct RowGradShape {
  // H rows per block in 32 KB of LDS, RB = H - 1 rows own atoms, the last one is the
  // halo row v0 + 1.  At N = 2048 FP32 three rows (48 KB): each row transformed 1.5
  // times instead of twice (measured 1.2 ms faster per 1307-frame slab; four rows,
  // 64 KB, lost more to occupancy than they saved)
  static constexpr int FIT = 32768 / (N * (int)sizeof(C));
  static constexpr int H = (N == 2048 && sizeof(C) == 8) ? 3 : (FIT < 2 ? 2 : (FIT < 8 ? FIT : 8));
};

// One block = RB rows (+ the halo row) of one frame: the kept band of each row,
// conjugated, is transformed (the adjoint DFT along y) and masked; every atom of the
// rows gathers its four nodes and writes its dL/dy', dL/dz, dL/df (re, im) for this
// frame to g_part (precision of the transforms).  The block's share of
// sum_a dL/dy'_a (regrouped order, fixed partition) goes to red, the argmin of y' to
// amin.
template <int N, typename C>
__global__ __launch_bounds__(BLK) void k_adj_rowgrad(gx_frames d, gx_adjoint g, int64_t f0, int64_t nf) {
  using T = typename Real<C>::T;
  constexpr int P = N / GX_WAVE;
  constexpr int H = RowGradShape<N, C>::H, RB = H - 1;
  constexpr int HW = (H + NWAVE - 1) / NWAVE;  // rows per wave
  constexpr int R0 = FftFirst<N>::R, RL = FftLast<N>::R;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *rows = reinterpret_cast<C *>(smem_raw);  // [H][N] cotangent of the grid rows
  C *stage = rows;                            // [W][H], consumed before rows is written
  double *ftab = reinterpret_cast<double *>(rows + H * N);
  __shared__ double mzr[H];                   // z part of the node mask of the block's rows
  __shared__ double shd[NWAVE];
  __shared__ int shi[NWAVE];
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int wave = threadIdx.x / GX_WAVE;
  const int nrb = (d.Z + RB - 1) / RB;
  const RowFrame rf = row_frame_tile(blockIdx.x, nrb, nf);
  const int fb = rf.fb;
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wf = max(0, min(jhi - jlo + 1, (int)d.W));
  const int z0 = rf.rb * RB;
  const int64_t n = d.n_atoms;

  for (int i = threadIdx.x; i < 2 * g.n_elem; i += BLK) ftab[i] = g.f_table[i];
  const int32_t *bnd = d.idx_bounds + 4 * fb;
  const int ylo = bnd[0], yhi = bnd[1], zlo = bnd[2], zhi = bnd[3];
  const int ya = d.fill_bkg ? (ylo == INT_MAX ? N : ylo + 1) : 0;
  const int yb = d.fill_bkg ? (yhi == INT_MIN ? 0 : yhi) : N;
  const bool masked = d.fill_bkg || d.smooth;
  if (threadIdx.x < H) {
    const int v = z0 + threadIdx.x;
    double mzv = 1.0;
    if (d.fill_bkg && !(v >= zlo && v < zhi)) mzv = 0.0;
    if (d.smooth && v < N) mzv = mzv * d.col_sz[(int64_t)fb * N + v];
    mzr[threadIdx.x] = mzv;
  }
  const C *src = reinterpret_cast<const C *>(d.inter) + (int64_t)fb * d.W * d.Z;
  for (int idx = threadIdx.x; idx < wf * H; idx += BLK) {
    const int w = idx / H, q = idx - w * H;
    const int v = z0 + q;
    stage[idx] = (v < d.Z) ? src[(int64_t)w * d.Z + v] : C{(T)0, (T)0};
  }
  __syncthreads();
  // this wave's rows q = wave + NWAVE h: the kept band, conjugated, into registers
  C x[HW][P];
#pragma unroll
  for (int h = 0; h < HW; ++h) {
    const int q = wave + NWAVE * h;
#pragma unroll
    for (int t = 0; t < P / R0; ++t) {
#pragma unroll
      for (int qq = 0; qq < R0; ++qq) {
        const int k = lane + GX_WAVE * t + qq * (N / R0);
        const int w = ((k + N / 2) & (N - 1)) - jlo;
        x[h][t * R0 + qq] = (q < H && w >= 0 && w < wf) ? cconj(stage[w * H + q]) : C{(T)0, (T)0};
      }
    }
  }
  __syncthreads();
  const C *tw = reinterpret_cast<const C *>(d.twiddle);
  // the transformed rows leave the registers conjugated and unmasked: the gather applies
  // the node mask M = (mz inz) x (my iny) at the four nodes it reads (no N-double mask
  // table in LDS, one more block per CU at N = 2048)
#pragma unroll
  for (int h = 0; h < HW; ++h) {
    const int q = wave + NWAVE * h;
    if (q >= H) break;  // wave-uniform
    C *buf = rows + q * N;
    fft_wave<N, C>(x[h], buf, tw, lane);
    wave_sync();
#pragma unroll
    for (int t = 0; t < P / RL; ++t)
#pragma unroll
      for (int qq = 0; qq < RL; ++qq) buf[lane + GX_WAVE * t + qq * (N / RL)] = cconj(x[h][t * RL + qq]);
  }
  __syncthreads();

  rowgrad_gather<C>(d, g, fp, fb, rf.rb, nrb, z0, RB, N, rows, ftab, shd, shi, masked ? mzr : nullptr, ya, yb);
}

// e_sum += the batch's per-frame, per-row-block dL/df sums, in two steps with fixed
// partitions: grid (RED, n_elem) sums chunk blockIdx.x of element blockIdx.y's nfb
// entries (in order per thread, then a fixed tree) into red; one block adds the RED
// chunk sums of each element in order.
__global__ __launch_bounds__(BLK) void k_adj_esum(gx_adjoint g, int64_t nfb) {
  __shared__ double shd[NWAVE];
  const int b = blockIdx.x, e = blockIdx.y;
  const int64_t chunk = (nfb + RED - 1) / RED;
  const int64_t i0 = b * chunk, i1 = min(nfb, i0 + chunk);
  double re = 0.0, im = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += BLK) {
    re += g.e_part[(i * g.n_elem + e) * 2];
    im += g.e_part[(i * g.n_elem + e) * 2 + 1];
  }
  re = block_sum(re, shd);
  im = block_sum(im, shd);
  if (threadIdx.x == 0) {
    g.red[((int64_t)e * RED + b) * 2] = re;
    g.red[((int64_t)e * RED + b) * 2 + 1] = im;
  }
}

__global__ __launch_bounds__(BLK) void k_adj_esum_fin(gx_adjoint g) {
  const int e = threadIdx.x;
  if (e >= g.n_elem) return;
  double re = 0.0, im = 0.0;
  for (int b = 0; b < RED; ++b) {
    re += g.red[((int64_t)e * RED + b) * 2];
    im += g.red[((int64_t)e * RED + b) * 2 + 1];
  }
  g.e_sum[2 * e] += re;
  g.e_sum[2 * e + 1] += im;
}

// ------------------------------------------------------------------------------------
// grid edges that are not powers of two: the differentiable passes on the forward
// path's zoomed chirp-z band transforms (block-wide, length M)
//
// The band transform X = S F_N x has the adjoint x_bar = conj(F_N^T conj(S^T X_bar)): the
// band DFT run from the K band values (k = k_lo + j) to the N grid points o.  With
// o k = (o^2 + k^2 - (k - o)^2) / 2 it is c[o] sum_j (u_j c[|k_j|]) conj(c[k_lo - (o - j)]),
// a convolution in the lag o - j, evaluated like the forward band with the same length M
// and twiddles and the kernel spectrum gx_adjoint.bs_kernel_adj (_c for columns).
// ------------------------------------------------------------------------------------
constexpr int BS_RB = 8;       // rows per block of the chirp-z row pass
constexpr int BS_GRAD_H = 4;   // rows per block of the chirp-z gradient gather (+ halo)

// forward: cloud-in-cell deposit of rows v - 1 and v, node mask, band transform along y,
// staged band stored transposed (as k_adj_rowpass)
template <int M, typename C>
__global__ __launch_bounds__(GX_BLK_FFT) void k_adj_rowpass_bs(gx_frames d, gx_adjoint g, int64_t f0) {
  using T = typename Real<C>::T;
  using AT = typename DepositAcc<C>::T;
  using AU = typename DepositAcc<C>::U;
  constexpr int P = M / GX_BLK_FFT;
  constexpr int R = BlkPlan<M>::R;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *fbuf = reinterpret_cast<C *>(smem_raw);  // [M] transform exchange; aliased by the accumulators [2][N]
  AU *acc = reinterpret_cast<AU *>(fbuf);
  const int N = (int)d.N;
  static_assert(2 * sizeof(AU) == sizeof(C), "the accumulators of a row take N elements of C");
  C *stage = fbuf + (M > N ? M : N);                      // [W][BS_RB]
  C *tspec = stage + d.W * BS_RB;                         // [W]
  double *ftab = reinterpret_cast<double *>(tspec + d.W);  // [n_elem][2]
  const int tid = threadIdx.x;
  const int fb = blockIdx.y;
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wf = jhi - jlo + 1;
  if (wf <= 0) return;  // block-uniform
  // the band's convolution is sized to the rows the deposit reaches (nodes u0 and
  // u0 + 1 <= y_max + 1 must stay below bs_M - bs_rk + 1)
  const int32_t *bnd = d.idx_bounds + 4 * fb;
  if (wf > d.W || jlo < d.bs_rlo + N / 2 || jhi > d.bs_rlo + d.bs_rk - 1 + N / 2 ||
      (bnd[1] != INT_MIN && bnd[1] + 1 > d.bs_M - d.bs_rk)) {
    if (tid == 0 && blockIdx.x == 0) atomicAdd(&d.status[1], 1);
    return;
  }
  const bool dense = d.fill_bkg || d.smooth;
  if (dense) {
    const C *spec = reinterpret_cast<const C *>(d.col_spec) + (int64_t)fb * d.W;
    for (int w = tid; w < wf; w += GX_BLK_FFT) tspec[w] = spec[w];
  }
  for (int i = tid; i < 2 * g.n_elem; i += GX_BLK_FFT) ftab[i] = g.f_table[i];
  const double s = fp[GX_FP_SIN], c = fp[GX_FP_COS];
  const double ymin = dkey_inv(d.ymin_key[3 * fb]);
  const int ylo = bnd[0], yhi = bnd[1], zlo = bnd[2], zhi = bnd[3];
  const int ya = d.fill_bkg ? (ylo == INT_MAX ? N : ylo + 1) : 0;
  const int yb = d.fill_bkg ? (yhi == INT_MIN ? 0 : yhi) : N;
  const double r = d.r, rinv = 1.0 / d.r;
  const double avv_re = d.avg_re * d.max_voxels, avv_im = d.avg_im * d.max_voxels;
  const C *chirp = reinterpret_cast<const C *>(d.bs_chirp);
  const C *kern = reinterpret_cast<const C *>(d.bs_kernel);
  const C *tw = reinterpret_cast<const C *>(d.bs_twiddle);
  const int z0 = blockIdx.x * BS_RB;
  __syncthreads();
  for (int rb = 0; rb < BS_RB; ++rb) {
    const int v = z0 + rb;
    const bool inside = !d.fill_bkg || (v >= zlo && v < zhi);
    const int a_lo = d.row_off[v > 0 ? v - 1 : 0], a_hi = d.row_off[v + 1];
    const bool atoms = inside && a_hi > a_lo;  // block-uniform
    C x[P];
    if (atoms) {
      for (int y = tid; y < 2 * N; y += GX_BLK_FFT) acc[y] = 0;
      __syncthreads();
      const double sc = ldexp(1.0, g.row_exp[v]);
      for (int a = a_lo + tid; a < a_hi; a += GX_BLK_FFT) {
        const bool own = d.azrow[a] == v;  // row v: weight 1 - tv; row v - 1: tv
        double tu = 0.0;
        const int u0 = cic_u(d.ax[a], d.ay[a], s, c, ymin, r, rinv, N, tu);
        if (u0 < 0) {
          atomicAdd(&d.status[2], 1);
          continue;
        }
        const double tv = g.afz[a];
        const int e = g.aelem[a];
        const double fr = ftab[2 * e], fi = ftab[2 * e + 1];
        const double wv = own ? 1.0 - tv : tv;
        const double w0 = (wv * (1.0 - tu)) * sc, w1 = (wv * tu) * sc;  // f (w sc) = (f w) sc exactly
        atomicAdd(&acc[u0], (AU)fx_round<AT>(fr * w0));
        atomicAdd(&acc[N + u0], (AU)fx_round<AT>(fi * w0));
        atomicAdd(&acc[u0 + 1], (AU)fx_round<AT>(fr * w1));
        atomicAdd(&acc[N + u0 + 1], (AU)fx_round<AT>(fi * w1));
      }
      __syncthreads();
      const double inv = 1.0 / sc;
#pragma unroll
      for (int t = 0; t < P / R; ++t)
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int n = tid + GX_BLK_FFT * t + q * (M / R);
          C xv = C{(T)0, (T)0};
          if (n < N) {
            T m = (T)0;
            if (n >= ya && n < yb) m = d.smooth ? (T)d.col_sy[(int64_t)fb * N + n] : (T)1;
            xv = cmul(C{(T)((double)(AT)acc[n] * inv) * m, (T)((double)(AT)acc[N + n] * inv) * m}, chirp[n]);
          }
          x[t * R + q] = xv;
        }
      __syncthreads();
      chirpz_block<M, C>(x, fbuf, kern, tw, tid);
    } else {
#pragma unroll
      for (int q = 0; q < P; ++q) x[q] = C{(T)0, (T)0};
    }
    const double mz = d.smooth ? d.col_sz[(int64_t)fb * N + v] : 1.0;
#pragma unroll
    for (int t = 0; t < P / R; ++t)
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int j = tid + GX_BLK_FFT * t + q * (M / R);
        if (j >= (int)d.bs_rk) continue;
        const int k = (int)d.bs_rlo + j;  // signed frequency
        const int w = k + N / 2 - jlo;    // fftshifted column within the frame's band
        if (w < 0 || w >= wf) continue;
        C o = C{(T)0, (T)0};
        if (atoms) o = cmul(x[t * R + q], chirp[k < 0 ? -k : k]);
        if (!inside) {
          o = (k == 0) ? C{(T)(avv_re * (double)N), (T)(avv_im * (double)N)} : C{(T)0, (T)0};
        } else if (dense) {
          const C u = tspec[w];
          double ore = mz * ((double)o.x + (double)u.x);
          double oim = mz * ((double)o.y + (double)u.y);
          if (d.smooth && k == 0) {
            ore = ore + avv_re * (double)N;
            oim = oim + avv_im * (double)N;
          }
          o = C{(T)ore, (T)oim};
        }
        stage[w * BS_RB + rb] = o;
      }
    __syncthreads();
  }
  C *out = reinterpret_cast<C *>(d.inter) + (int64_t)fb * d.W * d.Z;
  for (int idx = tid; idx < wf * BS_RB; idx += GX_BLK_FFT) {
    const int w = idx / BS_RB, q = idx - w * BS_RB;
    if (z0 + q < d.Z) out[(int64_t)w * d.Z + z0 + q] = stage[idx];
  }
}

// adjoint of the chirp-z column pass (in place on d.inter), one block per kept column:
// band transform along z, cotangent 2 (dL/d sum)[bin] F on binned pixels (and the
// column's share of dL/d avg), adjoint band transform back to z < Z
template <int M, typename C>
__global__ __launch_bounds__(GX_BLK_FFT) void k_adj_colpass_bs(gx_frames d, gx_adjoint g, int64_t f0) {
  using T = typename Real<C>::T;
  constexpr int P = M / GX_BLK_FFT;
  constexpr int R = BlkPlan<M>::R;
  __shared__ C fbuf[M];
  __shared__ double shd[GX_BLK_FFT / GX_WAVE];
  const int N = (int)d.N;
  const int tid = threadIdx.x;
  const int fb = blockIdx.y;
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  if (blockIdx.x == 0 && tid == 0) g.amin[fb] = INT_MAX;  // argmin slot of the gradient gather
  const int wcol = blockIdx.x;
  if (wcol > jhi - jlo || wcol >= d.W) return;  // block-uniform
  const bool dense = d.fill_bkg || d.smooth;
  double2 *apart = reinterpret_cast<double2 *>(g.avg_part);
  C *col = reinterpret_cast<C *>(d.inter) + ((int64_t)fb * d.W + wcol) * d.Z;
  const int j = jlo + wcol;
  const double hx = (j == N - 1) ? fp[GX_FP_HX_STOP] : (double)j * fp[GX_FP_HX_STEP] + fp[GX_FP_HX_START];
  const double hy = (j == N - 1) ? fp[GX_FP_HY_STOP] : (double)j * fp[GX_FP_HY_STEP] + fp[GX_FP_HY_START];
  const double dqinv = 1.0 / d.dq;
  const long long bx = to_int(floor_div(hx - d.qmin, d.dq, dqinv)) - d.c_lo;
  const long long by = to_int(floor_div(hy - d.qmin, d.dq, dqinv)) - d.c_lo;
  if (bx < 0 || bx >= d.c_n || by < 0 || by >= d.c_n) {  // never binned: zero cotangent
    for (int z = tid; z < d.Z; z += GX_BLK_FFT) col[z] = C{(T)0, (T)0};
    if (dense && tid == 0) apart[(int64_t)fb * d.W + wcol] = make_double2(0.0, 0.0);
    return;
  }
  const bool k0 = (j == N / 2);
  C u = C{(T)0, (T)0};
  if (!d.fill_bkg && d.smooth) u = reinterpret_cast<const C *>(d.col_spec)[(int64_t)fb * d.W + wcol];
  const double *sz_row = d.col_sz + (int64_t)fb * N;
  const C *chirp = reinterpret_cast<const C *>(d.bs_chirp);
  const C *tw = reinterpret_cast<const C *>(d.bs_twiddle_c);
  // a band shorter than N + bs_ck - 1 takes the rows z < Lc (>= Z) and the k = 0 column's
  // fill constant beyond them in closed form, as the forward column pass (k_colpass_bs)
  const int Lc = (int)min((int64_t)N, d.bs_Mc - d.bs_ck + 1);
  const bool tail = k0 && d.fill_bkg && Lc < N;
  C v[P];
#pragma unroll
  for (int t = 0; t < P / R; ++t)
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int z = tid + GX_BLK_FFT * t + q * (M / R);
      C e = C{(T)0, (T)0};
      if (z < Lc) e = cmul(col_input(d, (const C *)col, z, k0, u, sz_row), chirp[z]);
      v[t * R + q] = e;
    }
  chirpz_block<M, C>(v, fbuf, reinterpret_cast<const C *>(d.bs_kernel_c), tw, tid);
  const int64_t cbase = (by * d.c_n + bx) * d.c_n;
  const double val = d.q_val;
  const int nn = (N - 1) / 2 + 1;  // fftfreq: k < nn are the non-negative frequencies
  const double2 *zhat = reinterpret_cast<const double2 *>(g.avg_spec) + (int64_t)fb * (N + d.W);
  double are = 0.0, aim = 0.0, dc_re = 0.0, dc_im = 0.0;
#pragma unroll
  for (int t = 0; t < P / R; ++t)
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int jj = tid + GX_BLK_FFT * t + q * (M / R);
      C a = C{(T)0, (T)0};
      if (jj < (int)d.bs_ck) {
        const int k = (int)d.bs_clo + jj;  // signed frequency
        const int i = k + N / 2;           // fftshifted row
        const int kk = k < 0 ? k + N : k;  // natural index
        const C ck = chirp[k < 0 ? -k : k];
        double gr = 0.0, gi = 0.0;
        if (i >= d.i_lo && i <= d.i_hi) {
          const int kint = (kk < nn) ? kk : kk - N;
          const double qz = (((double)kint * val) * 2.0) * M_PI;
          const long long bz = to_int(floor_div(qz - d.qmin, d.dq, dqinv)) - d.c_lo;
          if (bz >= 0 && bz < d.c_n) {
            const double gI = 2.0 * g.g_bins[cbase + bz];
            C F = cmul(v[t * R + q], ck);
            if (tail) {
              const double2 gt = chirpz_fill_tail(d, kk, Lc);
              F = C{(T)((double)F.x + gt.x), (T)((double)F.y + gt.y)};
            }
            gr = gI * (double)F.x;
            gi = gI * (double)F.y;
          }
        }
        if (dense) {
          const double2 zh = zhat[kk];
          are += gr * zh.x + gi * zh.y;
          aim += gi * zh.x - gr * zh.y;
          if (kk == 0) {
            dc_re = gr;
            dc_im = gi;
          }
        }
        a = cmul(C{(T)gr, (T)(-gi)}, ck);  // conj(dL/dF) c[|k|]
      }
      v[t * R + q] = a;
    }
  if (dense) {  // block-uniform
    are = block_sum(are, shd);
    aim = block_sum(aim, shd);
    dc_re = block_sum(dc_re, shd);
    dc_im = block_sum(dc_im, shd);
    if (tid == 0) {
      const double2 yh = zhat[N + wcol];
      double pre = are * yh.x + aim * yh.y, pim = aim * yh.x - are * yh.y;
      if (k0) {
        const double c0 = d.max_voxels * (double)N * (double)N;
        pre = pre + c0 * dc_re;
        pim = pim + c0 * dc_im;
      }
      apart[(int64_t)fb * d.W + wcol] = make_double2(pre, pim);
    }
  }
  __syncthreads();  // fbuf is read by the last stage of the forward band transform
  chirpz_block<M, C>(v, fbuf, reinterpret_cast<const C *>(g.bs_kernel_adj_c), tw, tid);
  // dL/dx[z] = conj(c[z] y[z])
#pragma unroll
  for (int t = 0; t < P / R; ++t)
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int z = tid + GX_BLK_FFT * t + q * (M / R);
      if (z < d.Z) col[z] = cconj(cmul(v[t * R + q], chirp[z]));
    }
}

// FFT(Pz) at the column band and FFT(Py) on the kept columns by direct FP64 DFTs (as
// k_adj_avg_spectra for power-of-two edges), grid (nf, 2)
__global__ __launch_bounds__(BLK) void k_adj_avg_spectra_dft(gx_frames d, gx_adjoint g, int64_t f0) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double *p = reinterpret_cast<double *>(smem_raw);  // [N]
  const int N = (int)d.N;
  const int fb = blockIdx.x;
  const bool along_y = blockIdx.y == 1;
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wf = max(0, min(jhi - jlo + 1, (int)d.W));
  const int32_t *bnd = d.idx_bounds + 4 * fb;
  const int ylo = bnd[0], yhi = bnd[1], zlo = bnd[2], zhi = bnd[3];
  const int ya = ylo == INT_MAX ? N : ylo + 1, yb = yhi == INT_MIN ? 0 : yhi;
  for (int i = threadIdx.x; i < N; i += BLK) {
    double v;
    if (!along_y) {
      v = (!d.fill_bkg || (i >= zlo && i < zhi)) ? 1.0 : 0.0;
      if (d.smooth) v = v * d.col_sz[(int64_t)fb * N + i];
    } else {
      if (d.fill_bkg)
        v = (i >= ya && i < yb) ? (double)d.col_nm[(int64_t)fb * N + i] - d.max_voxels : 0.0;
      else
        v = -d.max_voxels;
      if (d.smooth) v = v * d.col_sy[(int64_t)fb * N + i];
    }
    p[i] = v;
  }
  __syncthreads();
  const double2 *tw = reinterpret_cast<const double2 *>(d.twiddle64);
  double2 *out = reinterpret_cast<double2 *>(g.avg_spec) + (int64_t)fb * (N + d.W);
  const int n_out = along_y ? wf : (int)d.bs_ck;
  for (int o = threadIdx.x; o < n_out; o += BLK) {
    int k;  // natural frequency index
    if (along_y) {
      k = (jlo + o + N - N / 2) % N;
    } else {
      const int ks = (int)d.bs_clo + o;
      k = ks < 0 ? ks + N : ks;
    }
    double re = 0.0, im = 0.0;
    int m = 0;
    for (int i = 0; i < N; ++i) {
      const double2 t = tw[m];
      re = fma(p[i], t.x, re);
      im = fma(p[i], t.y, im);
      m += k;
      if (m >= N) m -= N;
    }
    if (along_y)
      out[N + o] = make_double2(re, im);
    else
      out[k] = make_double2(re, im);
  }
}

// adjoint of the chirp-z row pass: per block BS_GRAD_H rows (the last the halo row), the
// kept band of each row, conjugated, through the adjoint band transform to the N grid
// points, masked; then the gather of k_adj_rowgrad
template <int M, typename C>
__global__ __launch_bounds__(GX_BLK_FFT) void k_adj_rowgrad_bs(gx_frames d, gx_adjoint g, int64_t f0) {
  using T = typename Real<C>::T;
  constexpr int P = M / GX_BLK_FFT;
  constexpr int R = BlkPlan<M>::R;
  constexpr int H = BS_GRAD_H, RB = H - 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int N = (int)d.N;
  C *rows = reinterpret_cast<C *>(smem_raw);                // [H][N] cotangent of the grid rows
  C *fbuf = rows + H * N;                                   // [M]
  C *stage = fbuf + M;                                      // [W][H]
  double *ftab = reinterpret_cast<double *>(stage + d.W * H);  // [n_elem][2]
  double *msk = ftab + 2 * g.n_elem;                        // [N]
  __shared__ double shd[GX_BLK_FFT / GX_WAVE];
  __shared__ int shi[GX_BLK_FFT / GX_WAVE];
  const int tid = threadIdx.x;
  const int fb = blockIdx.y;
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wf = max(0, min(jhi - jlo + 1, (int)d.W));
  const int z0 = blockIdx.x * RB;
  const int32_t *bnd = d.idx_bounds + 4 * fb;
  // outputs o < bs_M - bs_rk + 1 of the adjoint band transform are exact; the gather
  // reads nodes u0 and u0 + 1 <= y_max + 1 (block-uniform, before any barrier)
  if (bnd[1] != INT_MIN && bnd[1] + 1 > d.bs_M - d.bs_rk) {
    if (tid == 0 && blockIdx.x == 0) atomicAdd(&d.status[1], 1);
    return;
  }
  for (int i = tid; i < 2 * g.n_elem; i += GX_BLK_FFT) ftab[i] = g.f_table[i];
  const int ylo = bnd[0], yhi = bnd[1], zlo = bnd[2], zhi = bnd[3];
  const int ya = d.fill_bkg ? (ylo == INT_MAX ? N : ylo + 1) : 0;
  const int yb = d.fill_bkg ? (yhi == INT_MIN ? 0 : yhi) : N;
  const bool masked = d.fill_bkg || d.smooth;
  for (int y = tid; y < N; y += GX_BLK_FFT)
    msk[y] = (y >= ya && y < yb) ? (d.smooth ? d.col_sy[(int64_t)fb * N + y] : 1.0) : 0.0;
  const C *src = reinterpret_cast<const C *>(d.inter) + (int64_t)fb * d.W * d.Z;
  for (int idx = tid; idx < wf * H; idx += GX_BLK_FFT) {
    const int w = idx / H, q = idx - w * H;
    const int v = z0 + q;
    C e = C{(T)0, (T)0};
    if (v < d.Z) e = src[(int64_t)w * d.Z + v];
    stage[idx] = e;
  }
  __syncthreads();
  const C *chirp = reinterpret_cast<const C *>(d.bs_chirp);
  const C *kadj = reinterpret_cast<const C *>(g.bs_kernel_adj);
  const C *tw = reinterpret_cast<const C *>(d.bs_twiddle);
  for (int h = 0; h < H; ++h) {
    const int v = z0 + h;
    double mzv = 1.0;
    if (d.fill_bkg && !(v >= zlo && v < zhi)) mzv = 0.0;
    if (d.smooth && v < N) mzv = mzv * d.col_sz[(int64_t)fb * N + v];
    C x[P];
#pragma unroll
    for (int t = 0; t < P / R; ++t)
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int jj = tid + GX_BLK_FFT * t + q * (M / R);
        C a = C{(T)0, (T)0};
        if (jj < (int)d.bs_rk) {
          const int k = (int)d.bs_rlo + jj;
          const int w = k + N / 2 - jlo;
          if (w >= 0 && w < wf) a = cmul(cconj(stage[w * H + h]), chirp[k < 0 ? -k : k]);
        }
        x[t * R + q] = a;
      }
    __syncthreads();  // the previous row's transform has read fbuf
    chirpz_block<M, C>(x, fbuf, kadj, tw, tid);
#pragma unroll
    for (int t = 0; t < P / R; ++t)
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int o = tid + GX_BLK_FFT * t + q * (M / R);
        if (o >= N) continue;
        const C y = cconj(cmul(x[t * R + q], chirp[o]));
        const double m = masked ? mzv * msk[o] : 1.0;
        rows[h * N + o] = masked ? C{(T)((double)y.x * m), (T)((double)y.y * m)} : y;
      }
  }
  __syncthreads();
  rowgrad_gather<C>(d, g, fp, fb, blockIdx.x, gridDim.x, z0, RB, N, rows, ftab, shd, shi);
}

// ------------------------------------------------------------------------------------
// reductions (fixed partitions of the regrouped atom order: bitwise reproducible)
// ------------------------------------------------------------------------------------
// Origin term, one wave per frame: S_f = the sum of the frame's nb row-block partials
// of dL/dy' (lane l adds blocks l, l + 64, ... in order, then a fixed butterfly), and
// since d(min y')/dy'_a = [a = argmin_f], the argmin atom's dL/dy' of frame f becomes
// dL/dy' - S_f.  Frames own distinct rows of g_part.
template <typename T>
__global__ __launch_bounds__(BLK) void k_adj_frame_origin(gx_frames d, gx_adjoint g, int64_t nf, int nb, double *S) {
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int64_t fb = (blockIdx.x * (int64_t)BLK + threadIdx.x) / GX_WAVE;
  if (fb >= nf) return;  // wave-uniform
  double acc = 0.0;
  for (int b = lane; b < nb; b += GX_WAVE) acc += g.red[fb * nb + b];
#pragma unroll
  for (int o = GX_WAVE / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, GX_WAVE);
  if (lane != 0) return;
  S[fb] = acc;
  const int am = g.amin[fb];
  if (am == INT_MAX) return;
  T *u = reinterpret_cast<T *>(g.g_part) + fb * 4 * d.n_atoms + am;
  *u = (T)((double)*u - acc);
}

// g_atom += sum over the batch's frames of (sin, cos) dL/dy', dL/dz, dL/df.
// Block = RA atoms x RF frame lanes; lane j sums frames j, j + RF, ... in order,
// then the RF lane sums are added in a fixed tree order.  A wave reads 64 consecutive
// atoms' partials of one frame (whole cache lines): 1.5 ms faster per 1307-frame,
// 1.3M-atom slab than 16 atoms x 16 frame lanes.
constexpr int RA = 64, RF = BLK / RA;
template <typename T>
__global__ __launch_bounds__(BLK) void k_adj_reduce(gx_frames d, gx_adjoint g, int64_t f0, int64_t nf) {
  __shared__ double sh[5][RF][RA + 1];
  const int64_t n = d.n_atoms, n_rowed = d.row_off[d.N];
  const int al = threadIdx.x % RA, j = threadIdx.x / RA;
  const int64_t a = blockIdx.x * (int64_t)RA + al;
  const T *gpart = reinterpret_cast<const T *>(g.g_part);
  const bool esum = g.n_elem <= GX_ADJ_ELEM_REG;  // dL/df is in e_part instead
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (a < n_rowed) {
#pragma unroll 4
    for (int64_t fb = j; fb < nf; fb += RF) {
      const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
      const T *gp = gpart + fb * 4 * n;
      const double u = (double)gp[a];
      acc[0] += fp[GX_FP_SIN] * u;
      acc[1] += fp[GX_FP_COS] * u;
      acc[2] += (double)gp[n + a];
      if (!esum) {
        acc[3] += (double)gp[2 * n + a];
        acc[4] += (double)gp[3 * n + a];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) sh[k][j][al] = acc[k];
  __syncthreads();
  for (int w = RF / 2; w > 0; w >>= 1) {
    if (j < w)
#pragma unroll
      for (int k = 0; k < 5; ++k) sh[k][j][al] += sh[k][j + w][al];
    __syncthreads();
  }
  if (j == 0 && a < n_rowed)
#pragma unroll
    for (int k = 0; k < 5; ++k) g.g_atom[k * n + a] += sh[k][0][al];
}

// grid (RED, 1 + n_elem): y = 0 sums dL/dz over atoms (regrouped order) and finds the
// smallest original index attaining each bounding-box extreme (slots: x min, x max,
// y min, y max, z min, z max; only z min without fill / smoothing); y = 1 + e sums
// dL/df of element e
__global__ __launch_bounds__(BLK) void k_adj_fin_sums(gx_frames d, gx_adjoint g, int64_t slot) {
  __shared__ double shd[NWAVE];
  __shared__ int shi[NWAVE];
  const int b = blockIdx.x, y = blockIdx.y;
  const int64_t n = d.n_atoms;
  const int64_t chunk = (n + RED - 1) / RED;
  const int64_t i0 = b * chunk, i1 = min(n, i0 + chunk);
  const bool box = d.fill_bkg || d.smooth;
  const bool esum = g.n_elem <= GX_ADJ_ELEM_REG;
  double s0 = 0.0, s1 = 0.0;
  int am[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MAX, INT_MAX, INT_MAX};
  for (int64_t i = i0 + threadIdx.x; i < i1; i += BLK) {
    // sums over regrouped atoms i; extremes over original atoms i
    if (y == 0) {
      s0 += g.g_atom[2 * n + i];
      if (g.z[i] == g.zmin) am[4] = min(am[4], (int)i);
      if (box) {
        if (g.x[i] == g.ext[0]) am[0] = min(am[0], (int)i);
        if (g.x[i] == g.ext[1]) am[1] = min(am[1], (int)i);
        if (g.y[i] == g.ext[2]) am[2] = min(am[2], (int)i);
        if (g.y[i] == g.ext[3]) am[3] = min(am[3], (int)i);
        if (g.z[i] == g.ext[5]) am[5] = min(am[5], (int)i);
      }
    } else if (!esum && g.aelem[i] == y - 1) {
      s0 += g.g_atom[3 * n + i];
      s1 += g.g_atom[4 * n + i];
    }
  }
  if (esum && y > 0 && b == 0 && threadIdx.x == 0) {  // dL/df reduced per element already
    s0 = g.e_sum[2 * (y - 1)];
    s1 = g.e_sum[2 * (y - 1) + 1];
  }
  s0 = block_sum(s0, shd);
  s1 = block_sum(s1, shd);
  if (y == 0)
    for (int k = 0; k < 6; ++k) am[k] = block_min_int(am[k], shi);
  if (threadIdx.x == 0) {
    double *o = g.red + ((int64_t)y * RED + b) * 2;
    o[0] = s0;
    o[1] = s1;
    if (y == 0)
      for (int k = 0; k < 6; ++k)
        if (am[k] != INT_MAX) atomicMin(&g.amin[slot + k], am[k]);
  }
}

// y = 0: the origin term of min z; with fill / smoothing also the bounding-box
// terms of avg_voxel_f = sum f r^3 / (x_b y_b z_b): dL/dx_b = -Re(conj(dL/d avg) avg) / x_b
// (likewise y, z) on the extreme atoms.  y = 1 + e: dL/df_e, plus count_e r^3 / V dL/d avg.
__global__ __launch_bounds__(BLK) void k_adj_fin_apply(gx_frames d, gx_adjoint g, int64_t slot, double *g_f) {
  const int y = threadIdx.x;
  if (y > g.n_elem) return;
  const bool box = d.fill_bkg || d.smooth;
  double s0 = 0.0, s1 = 0.0;
  for (int b = 0; b < RED; ++b) {
    s0 += g.red[((int64_t)y * RED + b) * 2];
    s1 += g.red[((int64_t)y * RED + b) * 2 + 1];
  }
  const int64_t n = d.n_atoms;
  if (y == 0) {
    const int am = g.amin[slot + 4];
    if (am != INT_MAX) g.g_atom[2 * n + g.ainv[am]] -= s0;  // origin term of min z
    if (box) {
      const double t = g.g_avg[0] * d.avg_re + g.g_avg[1] * d.avg_im;
      for (int k = 0; k < 3; ++k) {
        const double ext = g.ext[2 * k + 1] - g.ext[2 * k];
        if (!(ext > 0.0)) continue;
        const double dl = -t / ext;
        const int lo = g.amin[slot + 2 * k], hi = g.amin[slot + 2 * k + 1];
        if (hi != INT_MAX) g.g_atom[k * n + g.ainv[hi]] += dl;
        if (lo != INT_MAX) g.g_atom[k * n + g.ainv[lo]] -= dl;
      }
    }
  } else {
    if (box) {
      const double ce = (double)g.ecount[y - 1] * g.vol_r3;
      s0 = s0 + ce * g.g_avg[0];
      s1 = s1 + ce * g.g_avg[1];
    }
    g_f[2 * (y - 1)] = s0;
    g_f[2 * (y - 1) + 1] = s1;
  }
}

__global__ __launch_bounds__(BLK) void k_adj_fin_out(gx_frames d, gx_adjoint g, double *g_coords) {
  const int64_t n = d.n_atoms, n_rowed = d.row_off[d.N];
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    const int a = g.ainv[i];
    const bool in = a < n_rowed;
    g_coords[3 * i] = in ? g.g_atom[a] : 0.0;
    g_coords[3 * i + 1] = in ? g.g_atom[n + a] : 0.0;
    g_coords[3 * i + 2] = in ? g.g_atom[2 * n + a] : 0.0;
  }
}

struct AffArgs {
  double a[9];
};

__global__ __launch_bounds__(BLK) void k_bins_adjoint(const unsigned long long *__restrict__ g_fx, int64_t n_rep,
                                                      const double *fx_scale, const int32_t *__restrict__ bcnt,
                                                      int64_t c_n, int64_t c_lo, int64_t q_num, double qa_start,
                                                      double qa_step, double qa_stop, AffArgs aff, double z_elem,
                                                      double c16pi2, double *g_bins) {
  const int64_t n = c_n * c_n * c_n;
  const double inv = 1.0 / fx_scale[0];
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    const int32_t cnt = bcnt[i];
    double gb = 0.0;
    if (cnt != 0) {
      const int64_t iz = i % c_n, ix = (i / c_n) % c_n, iy = i / (c_n * c_n);
      const double qx = lin_axis(c_lo + ix, q_num, qa_start, qa_step, qa_stop);
      const double qy = lin_axis(c_lo + iy, q_num, qa_start, qa_step, qa_stop);
      const double qz = lin_axis(c_lo + iz, q_num, qa_start, qa_step, qa_stop);
      const double q2 = (qx * qx + qy * qy) + qz * qz;
      unsigned long long acc = 0;  // exact (wrapping) integer sum of the copies
      for (int64_t r = 0; r < n_rep; ++r) acc += g_fx[r * n + i];
      const double giq = (double)(long long)acc * inv;
      gb = (giq * f0_weight(aff.a, q2, z_elem, c16pi2)) / (double)cnt;
    }
    g_bins[i] = gb;
  }
}

// ------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------
template <int N, typename C>
int launch_adj_rowpass_t(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, hipStream_t st) {
  constexpr int RB = RowPassShape<N, C>::RB;
  const size_t lds = (size_t)RB * N * sizeof(C) + (size_t)d->W * (RB + 2) * sizeof(C) + (size_t)g->n_elem * 16;
  if (lds > 150 * 1024)
    return set_error(GX_ERR_UNSUPPORTED, "differentiable row pass needs %zu bytes of LDS (N=%d, W=%lld)", lds, N,
                     (long long)d->W);
  static bool attr_set = false;
  if (!attr_set) {
    GX_HIP(hipFuncSetAttribute((const void *)k_adj_rowpass<N, C>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               150 * 1024));
    attr_set = true;
  }
  const int64_t nb = (d->Z + RB - 1) / RB * nf;
  GX_REQUIRE(nb < (1ll << 31), "%lld row blocks x frames in one launch", (long long)nb);
  k_adj_rowpass<N, C><<<(unsigned)nb, BLK, lds, st>>>(*d, *g, f0, nf);
  return check_launch("k_adj_rowpass");
}

template <int N, typename C>
int launch_adj_colpass_t(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, hipStream_t st,
                         bool stored) {
  dim3 grid(grid1d(d->W, NWAVE), (unsigned)nf);
  k_adj_colpass<N, C><<<grid, BLK, 0, st>>>(*d, *g, f0, stored);
  return check_launch("k_adj_colpass");
}

template <int N, typename C>
int launch_adj_rowgrad_t(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, hipStream_t st) {
  constexpr int H = RowGradShape<N, C>::H, RB = H - 1;
  // rows (the staging tile aliases them, W <= N) + f table
  const size_t lds = (size_t)H * N * sizeof(C) + (size_t)g->n_elem * 16;
  if (lds > 150 * 1024)
    return set_error(GX_ERR_UNSUPPORTED, "adjoint row pass needs %zu bytes of LDS (N=%d, W=%lld)", lds, N,
                     (long long)d->W);
  static bool attr_set = false;
  if (!attr_set) {
    GX_HIP(hipFuncSetAttribute((const void *)k_adj_rowgrad<N, C>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               150 * 1024));
    attr_set = true;
  }
  const int64_t nb = (d->Z + RB - 1) / RB * nf;
  GX_REQUIRE(nb < (1ll << 31), "%lld row blocks x frames in one launch", (long long)nb);
  k_adj_rowgrad<N, C><<<(unsigned)nb, BLK, lds, st>>>(*d, *g, f0, nf);
  return check_launch("k_adj_rowgrad");
}

// row blocks per frame of k_adj_rowgrad (its partial sums per frame)
int rowgrad_blocks(const gx_frames *d) {
  if (d->bs_M) return (int)((d->Z + BS_GRAD_H - 2) / (BS_GRAD_H - 1));
  int rb = 1;
  switch (d->N) {
    case 256: rb = d->fft_f64 ? RowGradShape<256, double2>::H - 1 : RowGradShape<256, float2>::H - 1; break;
    case 512: rb = d->fft_f64 ? RowGradShape<512, double2>::H - 1 : RowGradShape<512, float2>::H - 1; break;
    case 1024: rb = d->fft_f64 ? RowGradShape<1024, double2>::H - 1 : RowGradShape<1024, float2>::H - 1; break;
    case 2048: rb = d->fft_f64 ? RowGradShape<2048, double2>::H - 1 : RowGradShape<2048, float2>::H - 1; break;
  }
  return (int)((d->Z + rb - 1) / rb);
}

// OP_COLPASS_STORED: the column adjoint on the kept band the forward's column pass
// stored (gx_frames.f_band) instead of recomputing each column's transform
enum Op { OP_ROWPASS, OP_COLPASS, OP_COLPASS_STORED, OP_ROWGRAD, OP_AVGSPEC };

template <int N>
int dispatch_n(Op op, const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, hipStream_t st) {
  if (op == OP_AVGSPEC) {
    k_adj_avg_spectra<N><<<dim3((unsigned)nf, 2), GX_WAVE, 0, st>>>(*d, *g, f0);
    return check_launch("k_adj_avg_spectra");
  }
  if (d->fft_f64) {
    if (op == OP_ROWPASS) return launch_adj_rowpass_t<N, double2>(d, g, f0, nf, st);
    if (op == OP_COLPASS || op == OP_COLPASS_STORED)
      return launch_adj_colpass_t<N, double2>(d, g, f0, nf, st, op == OP_COLPASS_STORED);
    return launch_adj_rowgrad_t<N, double2>(d, g, f0, nf, st);
  }
  if (op == OP_ROWPASS) return launch_adj_rowpass_t<N, float2>(d, g, f0, nf, st);
  if (op == OP_COLPASS || op == OP_COLPASS_STORED)
    return launch_adj_colpass_t<N, float2>(d, g, f0, nf, st, op == OP_COLPASS_STORED);
  return launch_adj_rowgrad_t<N, float2>(d, g, f0, nf, st);
}

template <int M, typename C>
int launch_adj_bs_t(Op op, const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, hipStream_t st) {
  if (op == OP_ROWPASS) {
    const size_t lds = (size_t)(M > d->N ? M : d->N) * sizeof(C) + (size_t)d->W * (BS_RB + 1) * sizeof(C) +
                       (size_t)g->n_elem * 16;
    if (lds > 150 * 1024)
      return set_error(GX_ERR_UNSUPPORTED, "chirp-z differentiable row pass needs %zu bytes of LDS", lds);
    static bool attr_set = false;
    if (!attr_set) {
      GX_HIP(hipFuncSetAttribute((const void *)k_adj_rowpass_bs<M, C>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 150 * 1024));
      attr_set = true;
    }
    dim3 grid((unsigned)((d->Z + BS_RB - 1) / BS_RB), (unsigned)nf);
    k_adj_rowpass_bs<M, C><<<grid, GX_BLK_FFT, lds, st>>>(*d, *g, f0);
    return check_launch("k_adj_rowpass_bs");
  }
  if (op == OP_COLPASS) {
    dim3 grid((unsigned)d->W, (unsigned)nf);
    k_adj_colpass_bs<M, C><<<grid, GX_BLK_FFT, 0, st>>>(*d, *g, f0);
    return check_launch("k_adj_colpass_bs");
  }
  // OP_ROWGRAD (150 KB of dynamic LDS at most: the gather's reductions use static LDS)
  const size_t lds = ((size_t)BS_GRAD_H * d->N + M + (size_t)d->W * BS_GRAD_H) * sizeof(C) +
                     (size_t)g->n_elem * 16 + (size_t)d->N * 8;
  if (lds > 150 * 1024)
    return set_error(GX_ERR_UNSUPPORTED, "chirp-z adjoint row pass needs %zu bytes of LDS (N=%lld, W=%lld)", lds,
                     (long long)d->N, (long long)d->W);
  static bool attr_set = false;
  if (!attr_set) {
    GX_HIP(hipFuncSetAttribute((const void *)k_adj_rowgrad_bs<M, C>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               150 * 1024));
    attr_set = true;
  }
  dim3 grid((unsigned)rowgrad_blocks(d), (unsigned)nf);
  k_adj_rowgrad_bs<M, C><<<grid, GX_BLK_FFT, lds, st>>>(*d, *g, f0);
  return check_launch("k_adj_rowgrad_bs");
}

template <typename C>
int dispatch_bs_c(Op op, const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, hipStream_t st) {
  const int64_t M = (op == OP_COLPASS) ? d->bs_Mc : d->bs_M;
  switch (M) {
    case 1024: return launch_adj_bs_t<1024, C>(op, d, g, f0, nf, st);
    case 2048: return launch_adj_bs_t<2048, C>(op, d, g, f0, nf, st);
    case 4096: return launch_adj_bs_t<4096, C>(op, d, g, f0, nf, st);
    default: return set_error(GX_ERR_UNSUPPORTED, "chirp-z transform length %lld", (long long)M);
  }
}

int dispatch_bs(Op op, const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, hipStream_t st) {
  if (op == OP_AVGSPEC) {
    k_adj_avg_spectra_dft<<<dim3((unsigned)nf, 2), BLK, (size_t)d->N * sizeof(double), st>>>(*d, *g, f0);
    return check_launch("k_adj_avg_spectra_dft");
  }
  if (d->fft_f64) return dispatch_bs_c<double2>(op, d, g, f0, nf, st);
  return dispatch_bs_c<float2>(op, d, g, f0, nf, st);
}

int dispatch(Op op, const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, hipStream_t st) {
  if (d->bs_M) return dispatch_bs(op, d, g, f0, nf, st);
  switch (d->N) {
    case 256: return dispatch_n<256>(op, d, g, f0, nf, st);
    case 512: return dispatch_n<512>(op, d, g, f0, nf, st);
    case 1024: return dispatch_n<1024>(op, d, g, f0, nf, st);
    case 2048: return dispatch_n<2048>(op, d, g, f0, nf, st);
    default: return set_error(GX_ERR_UNSUPPORTED, "differentiable path: grid edge N=%lld is not 256/512/1024/2048",
                              (long long)d->N);
  }
}

int check_adj(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, bool backward) {
  GX_REQUIRE(d != nullptr && g != nullptr, "NULL descriptor");
  if (d->bs_M) {
    GX_REQUIRE(d->N >= 16 && d->N <= 2048, "differentiable path: grid edge N=%lld outside [16, 2048]",
               (long long)d->N);
    GX_REQUIRE(d->bs_chirp && d->bs_kernel && d->bs_twiddle && d->bs_kernel_c && d->bs_twiddle_c && d->twiddle64,
               "chirp-z grid edge N=%lld needs the band tables of gx_frames", (long long)d->N);
    GX_REQUIRE((d->bs_M == 1024 || d->bs_M == 2048 || d->bs_M == 4096) &&
                   (d->bs_Mc == 1024 || d->bs_Mc == 2048 || d->bs_Mc == 4096) && d->bs_M >= d->bs_rk &&
                   (d->bs_Mc >= d->N + d->bs_ck - 1 ||
                    (d->bs_Mc - d->bs_ck + 1 >= d->Z && (d->fill_bkg || !d->smooth))),
               "chirp-z transform lengths (%lld, %lld) do not fit N=%lld", (long long)d->bs_M, (long long)d->bs_Mc,
               (long long)d->N);
  } else {
    GX_REQUIRE(d->N == 256 || d->N == 512 || d->N == 1024 || d->N == 2048,
               "differentiable path: grid edge N=%lld is not 256/512/1024/2048 (other edges need chirp-z tables)",
               (long long)d->N);
  }
  const bool dense = d->fill_bkg || d->smooth;
  GX_REQUIRE(!d->fill_bkg || d->col_nm, "fill_bkg needs col_nm");
  GX_REQUIRE(!d->smooth || (d->col_sy && d->col_sz && d->gauss_cw), "smooth needs col_sy, col_sz, gauss_cw");
  GX_REQUIRE(!dense || (d->col_spec && d->twiddle64), "fill_bkg / smooth need col_spec and twiddle64");
  GX_REQUIRE(d->bs_M == 0 || (d->N != 256 && d->N != 512 && d->N != 1024 && d->N != 2048),
             "differentiable path: chirp-z tables given for a power-of-two grid");
  GX_REQUIRE(!d->inter_rm, "differentiable path: the intermediate is [batch][W][Z] (inter_rm = 0)");
  GX_REQUIRE(!d->f_band || (d->bs_M == 0 && d->fft_plan_col == 0),
             "differentiable path: f_band needs a power-of-two grid and fft_plan_col = 0 (the column "
             "adjoint recomputes the other batches' transforms with that plan)");
  GX_REQUIRE(d->r > 0.0, "r_voxel_size must be positive");
  GX_REQUIRE(nf >= 0 && nf <= d->batch, "batch of %lld frames exceeds workspace batch %lld", (long long)nf,
             (long long)d->batch);
  GX_REQUIRE(f0 >= 0 && f0 + nf <= d->n_frames, "frames [%lld, %lld) outside [0, %lld)", (long long)f0,
             (long long)(f0 + nf), (long long)d->n_frames);
  GX_REQUIRE(d->W >= 1 && d->W <= d->N, "intermediate width W=%lld outside [1, N]", (long long)d->W);
  GX_REQUIRE(d->Z >= 8 && d->Z <= d->N && d->Z % 8 == 0, "row range Z=%lld must be a multiple of 8 in [8, N]",
             (long long)d->Z);
  GX_REQUIRE(d->c_n >= 1, "empty cropped grid");
  GX_REQUIRE(d->fparams && d->ymin_key && d->idx_bounds && d->inter && d->twiddle && d->status && d->row_off &&
                 d->ax && d->ay && d->azrow,
             "a required device buffer of gx_frames is NULL");
  GX_REQUIRE(g->n_elem >= 1 && g->n_elem <= 128, "n_elem=%lld outside [1, 128]", (long long)g->n_elem);
  GX_REQUIRE(d->batch <= 8192, "batch of %lld frames exceeds 8192", (long long)d->batch);
  GX_REQUIRE(g->f_table && g->afz && g->aperm && g->aelem && g->ecnt && g->row_exp && g->ainv,
             "a required buffer of gx_adjoint is NULL");
  if (backward) {
    GX_REQUIRE(g->g_bins && g->g_part && g->red && g->amin && g->g_atom, "a backward buffer of gx_adjoint is NULL");
    GX_REQUIRE(g->n_elem > GX_ADJ_ELEM_REG || (g->e_part && g->e_sum),
               "n_elem <= %d needs gx_adjoint.e_part and e_sum", GX_ADJ_ELEM_REG);
    GX_REQUIRE(!d->bs_M || (g->bs_kernel_adj && g->bs_kernel_adj_c),
               "chirp-z grid edge needs gx_adjoint.bs_kernel_adj and bs_kernel_adj_c");
    GX_REQUIRE(!dense || (g->x && g->y && g->ext && g->ecount && g->avg_spec && g->avg_part && g->g_avg),
               "fill_bkg / smooth need gx_adjoint.x, y, ext, ecount, avg_spec, avg_part, g_avg");
  }
  return GX_OK;
}

}  // namespace

extern "C" {

int gx_adj_prepare(const gx_frames *d, const gx_adjoint *g, gx_stream_t stream) {
  clear_error();
  int rc = check_adj(d, g, 0, 0, false);
  if (rc) return rc;
  GX_REQUIRE(g->g_atom, "g_atom is NULL");
  hipStream_t st = as_stream(stream);
  unsigned gr = grid1d(d->n_atoms > d->N ? d->n_atoms : d->N, BLK);
  k_adj_prepare<<<gr > 4096 ? 4096 : gr, BLK, 0, st>>>(*d, *g);
  return check_launch("gx_adj_prepare");
}

int gx_adj_rowpass(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream) {
  clear_error();
  int rc = check_adj(d, g, f0, nf, false);
  if (rc || nf == 0) return rc;
  return dispatch(OP_ROWPASS, d, g, f0, nf, as_stream(stream));
}

int gx_adj_forward(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream) {
  clear_error();
  int rc = check_adj(d, g, f0, nf, false);
  if (rc || nf == 0) return rc;
  if ((rc = gx_frames_stats(d, f0, nf, stream))) return rc;
  if ((d->fill_bkg || d->smooth) && (rc = gx_frames_columns(d, f0, nf, stream))) return rc;
  if ((rc = dispatch(OP_ROWPASS, d, g, f0, nf, as_stream(stream)))) return rc;
  return gx_frames_colpass(d, f0, nf, stream);
}

int gx_adj_colpass(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream) {
  clear_error();
  int rc = check_adj(d, g, f0, nf, true);
  if (rc || nf == 0) return rc;
  return dispatch(OP_COLPASS, d, g, f0, nf, as_stream(stream));
}

int gx_adj_rowgrad(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream) {
  clear_error();
  int rc = check_adj(d, g, f0, nf, true);
  if (rc || nf == 0) return rc;
  k_fill_i32<<<grid1d(nf, BLK), BLK, 0, as_stream(stream)>>>(g->amin, nf, INT_MAX);
  return dispatch(OP_ROWGRAD, d, g, f0, nf, as_stream(stream));
}

int gx_adj_reduce(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream) {
  clear_error();
  int rc = check_adj(d, g, f0, nf, true);
  if (rc || nf == 0) return rc;
  hipStream_t st = as_stream(stream);
  double *S = g->red + d->batch * d->N;  // per-frame sums of dL/dy'
  const unsigned gw = grid1d(nf * GX_WAVE, BLK);  // a wave per frame
  if (d->fft_f64)
    k_adj_frame_origin<double><<<gw, BLK, 0, st>>>(*d, *g, nf, rowgrad_blocks(d), S);
  else
    k_adj_frame_origin<float><<<gw, BLK, 0, st>>>(*d, *g, nf, rowgrad_blocks(d), S);
  if (d->n_atoms > 0) {
    if (d->fft_f64)
      k_adj_reduce<double><<<grid1d(d->n_atoms, RA), BLK, 0, st>>>(*d, *g, f0, nf);
    else
      k_adj_reduce<float><<<grid1d(d->n_atoms, RA), BLK, 0, st>>>(*d, *g, f0, nf);
  }
  if (g->n_elem <= GX_ADJ_ELEM_REG) {  // red is free again (k_adj_frame_origin has read it)
    k_adj_esum<<<dim3(RED, (unsigned)g->n_elem), BLK, 0, st>>>(*g, nf * rowgrad_blocks(d));
    k_adj_esum_fin<<<1, BLK, 0, st>>>(*g);
  }
  return check_launch("gx_adj_reduce");
}

static int adj_backward(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream,
                        bool recompute) {
  clear_error();
  int rc = check_adj(d, g, f0, nf, true);
  if (rc || nf == 0) return rc;
  hipStream_t st = as_stream(stream);
  const bool dense = d->fill_bkg || d->smooth;
  if (dense) {
    if ((rc = gx_frames_columns(d, f0, nf, stream))) return rc;
    if ((rc = dispatch(OP_AVGSPEC, d, g, f0, nf, st))) return rc;
  }
  if (recompute && (rc = dispatch(OP_ROWPASS, d, g, f0, nf, st))) return rc;
  // the batch the forward left also has its columns' kept band stored (when f_band is given)
  const Op col = (!recompute && d->f_band) ? OP_COLPASS_STORED : OP_COLPASS;
  if ((rc = dispatch(col, d, g, f0, nf, st))) return rc;
  if (dense) {
    k_adj_avg_frames<<<(unsigned)nf, BLK, 0, st>>>(*d, *g, f0);
    k_adj_avg_reduce<<<1, BLK, 0, st>>>(*d, *g, nf);
    if ((rc = check_launch("k_adj_avg_reduce"))) return rc;
  }
  // k_adj_colpass has reset the frames' argmin slots
  if ((rc = dispatch(OP_ROWGRAD, d, g, f0, nf, st))) return rc;
  return gx_adj_reduce(d, g, f0, nf, stream);
}

int gx_adj_backward(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream) {
  return adj_backward(d, g, f0, nf, stream, true);
}

int gx_adj_backward_stored(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream) {
  return adj_backward(d, g, f0, nf, stream, false);
}

int gx_adj_finish(const gx_frames *d, const gx_adjoint *g, double *g_coords, double *g_f, gx_stream_t stream) {
  clear_error();
  int rc = check_adj(d, g, 0, 0, true);
  if (rc) return rc;
  GX_REQUIRE(g_coords && g_f, "NULL output");
  GX_REQUIRE(g->z, "gx_adjoint.z (original z) is NULL");
  GX_REQUIRE(g->n_elem + 1 <= BLK, "n_elem");
  hipStream_t st = as_stream(stream);
  const int64_t slot = d->batch;
  k_fill_i32<<<1, BLK, 0, st>>>(g->amin + slot, 6, INT_MAX);
  k_adj_fin_sums<<<dim3(RED, (unsigned)(1 + g->n_elem)), BLK, 0, st>>>(*d, *g, slot);
  k_adj_fin_apply<<<1, BLK, 0, st>>>(*d, *g, slot, g_f);
  if (d->n_atoms > 0) {
    unsigned gr = grid1d(d->n_atoms, BLK);
    k_adj_fin_out<<<gr > 4096 ? 4096 : gr, BLK, 0, st>>>(*d, *g, g_coords);
  }
  return check_launch("gx_adj_finish");
}

int gx_bins_adjoint(const unsigned long long *g_fx, int64_t n_rep, const double *fx_scale, const int32_t *bin_cnt,
                    int64_t c_n,
                    int64_t c_lo, int64_t q_num, double qa_start, double qa_step, double qa_stop,
                    const double *host_aff, double z_elem, double *g_bins, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(g_fx && fx_scale && bin_cnt && host_aff && g_bins, "NULL argument");
  GX_REQUIRE(n_rep >= 1 && n_rep <= 64, "n_rep=%lld outside [1, 64]", (long long)n_rep);
  GX_REQUIRE(c_n >= 1 && c_lo >= 0 && c_lo + c_n <= q_num, "crop [%lld, %lld) outside the q axis of %lld",
             (long long)c_lo, (long long)(c_lo + c_n), (long long)q_num);
  AffArgs aff;
  for (int i = 0; i < 9; ++i) aff.a[i] = host_aff[i];
  const double c16pi2 = 16 * (M_PI * M_PI);
  const int64_t n = c_n * c_n * c_n;
  unsigned gr = grid1d(n, BLK);
  k_bins_adjoint<<<gr > 8192 ? 8192 : gr, BLK, 0, as_stream(stream)>>>(g_fx, n_rep, fx_scale, bin_cnt, c_n, c_lo, q_num,
                                                                       qa_start, qa_step, qa_stop, aff, z_elem,
                                                                       c16pi2, g_bins);
  return check_launch("gx_bins_adjoint");
}

}  // extern "C"
