// gfx950 kernels of the phi-frame pipeline (reference: tools/voxelgrids.py
// rotate_project_fft_coords 311-416 and process_file2 464-506, driven by
// tools/comparison.py voxelgridmaker_fitting 673-788).
//
// Data flow per batch of frames (all device resident):
//   k_frames_ymin     min_y'(phi) over all atoms              (voxelgrids.py:319-323)
//   k_frames_bounds   y/z index bounds over valid atoms       (voxelgrids.py:328-349)
//   k_frames_columns  num_missing_voxels and smoothed masks   (voxelgrids.py:350-374)
//   k_rowpass<N>      per z row: LDS fixed-point scatter of the row's atoms,
//                     fill + blend, FFT along y, keep the fftshifted column band
//                     that reaches the cropped bins, LDS-staged transposed store
//   k_colpass<N>      per kept column: FFT along z, |F|^2, bin indices in FP64,
//                     deterministic fixed-point bin sums + int32 counts
#include "gx_device.h"
#include "gx_host.h"

#include <limits.h>
#include <math.h>

using namespace gx;

namespace {

constexpr int BLK = 256;

// ------------------------------------------------------------------------------------
// slab preparation
// ------------------------------------------------------------------------------------
__global__ void k_slab_bounds_init(unsigned long long *keys, unsigned long long *counts, int64_t n_elem) {
  int t = threadIdx.x;
  if (t < 6) keys[t] = (t % 2 == 0) ? ~0ull : 0ull;
  for (int e = t; e < n_elem; e += blockDim.x) counts[e] = 0;
}

constexpr int MAX_ELEM = 128;

// Each thread takes PT atoms per grid stride, all loads issued before the first use;
// the block reduces its extremes in LDS and adds one atomic per key.
template <int PT>
__global__ __launch_bounds__(BLK) void k_slab_bounds(const double *__restrict__ x, const double *__restrict__ y,
                                                     const double *__restrict__ z, const int32_t *__restrict__ el,
                                                     int64_t n, int64_t n_elem, unsigned long long *keys,
                                                     unsigned long long *counts) {
  constexpr int NW = BLK / GX_WAVE;
  __shared__ unsigned int hist[MAX_ELEM];
  __shared__ double ext[NW][6];
  for (int e = threadIdx.x; e < MAX_ELEM; e += BLK) hist[e] = 0;
  __syncthreads();
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t a0 = blockIdx.x * (int64_t)(BLK * PT) + threadIdx.x; a0 < n; a0 += (int64_t)gridDim.x * (BLK * PT)) {
    double v[PT][3];
    int e[PT];
#pragma unroll
    for (int k = 0; k < PT; ++k) {
      const int64_t a = a0 + (int64_t)k * BLK;
      v[k][0] = v[k][1] = v[k][2] = 0.0;
      e[k] = -1;
      if (a < n) {
        v[k][0] = x[a];
        v[k][1] = y[a];
        v[k][2] = z[a];
        e[k] = el[a];
      }
    }
#pragma unroll
    for (int k = 0; k < PT; ++k) {
      if (a0 + (int64_t)k * BLK >= n) break;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        mn[c] = fmin(mn[c], v[k][c]);
        mx[c] = fmax(mx[c], v[k][c]);
      }
      if (e[k] >= 0 && e[k] < n_elem) atomicAdd(&hist[e[k]], 1u);
    }
  }
  const int lane = threadIdx.x & (GX_WAVE - 1), w = threadIdx.x / GX_WAVE;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double a = mn[c], b = mx[c];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      a = fmin(a, __shfl_xor(a, o, GX_WAVE));
      b = fmax(b, __shfl_xor(b, o, GX_WAVE));
    }
    if (lane == 0) {
      ext[w][2 * c] = a;
      ext[w][2 * c + 1] = b;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n_elem; e += BLK)
    if (hist[e]) atomicAdd(&counts[e], (unsigned long long)hist[e]);
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    double v = ext[0][k];
    for (int i = 1; i < NW; ++i) v = (k & 1) ? fmax(v, ext[i][k]) : fmin(v, ext[i][k]);
    if (k & 1)
      atomicMax(&keys[k], dkey(v));
    else
      atomicMin(&keys[k], dkey(v));
  }
}

__global__ void k_slab_bounds_fin(unsigned long long *keys) {
  int t = threadIdx.x;
  if (t < 6) {
    double v = dkey_inv(keys[t]);
    reinterpret_cast<double *>(keys)[t] = v;
  }
}

// work layout: [N * n_elem] per-row element counts | [N + 1] row counts | [N + 1] cursors
//
// Projected z row of atom a (z_indices, tools/voxelgrids.py:329): N for rows outside
// [0, N) and for unknown element ids, which never reach a row.
__device__ __forceinline__ int slab_row_of(const gx_slab &s, int64_t a, double rinv) {
  const int N = (int)s.N;
  const double zz = s.z[a] - s.zmin;
  const long long zi = to_int(floor_div(zz, s.r, rinv));
  int row = (zi >= 0 && zi < N) ? (int)zi : N;
  const int e = s.elem[a];
  if (e < 0 || e >= s.n_elem) row = N;
  return row;
}

// Row counting and placement work on chunks of PT * BLK atoms per block: counts
// go to LDS histograms first and reach the global counters once per (block, row),
// so a row shared by many atoms costs one global atomic per block, not per atom.
template <int PT> constexpr int sort_chunk() { return PT * BLK; }
constexpr int SORT_LDS_INTS = 16384;  // 64 KB of LDS histograms at most

// LDS: [N + 1] row counts | [N * n_elem] per-row element counts when elem_lds
// (otherwise the element counts go straight to the global counters).
template <int PT>
__global__ __launch_bounds__(BLK) void k_slab_rowcount(gx_slab s, int elem_lds) {
  constexpr int SORT_CHUNK = sort_chunk<PT>();
  extern __shared__ int hist[];
  const int N = (int)s.N, M = N + 1;
  const int ne = (int)s.n_elem;
  int *hrow = hist, *hel = hist + M;
  const int nh = M + (elem_lds ? N * ne : 0);
  for (int i = threadIdx.x; i < nh; i += BLK) hist[i] = 0;
  __syncthreads();
  const double rinv = 1.0 / s.r;
  int32_t *ecnt = s.work;
  int32_t *rcnt = s.work + (int64_t)N * ne;
  const int64_t a0 = blockIdx.x * (int64_t)SORT_CHUNK, a1 = min(s.n_atoms, a0 + SORT_CHUNK);
  for (int64_t a = a0 + threadIdx.x; a < a1; a += BLK) {
    const int row = slab_row_of(s, a, rinv);
    atomicAdd(&hrow[row], 1);
    if (row < N) {
      const int e = s.elem[a];
      if (elem_lds) atomicAdd(&hel[row * ne + e], 1);
      else atomicAdd(&ecnt[(int64_t)row * ne + e], 1);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < M; i += BLK)
    if (hrow[i]) atomicAdd(&rcnt[i], hrow[i]);
  if (elem_lds)
    for (int i = threadIdx.x; i < N * ne; i += BLK)
      if (hel[i]) atomicAdd(&ecnt[i], hel[i]);
}

// Single block: exclusive scan of the N+1 row counts, per-row fixed-point exponent.
__global__ __launch_bounds__(1024) void k_slab_scan(gx_slab s) {
  const int N = (int)s.N;
  const int M = N + 1;
  int32_t *ecnt = s.work;
  int32_t *rcnt = s.work + (int64_t)N * s.n_elem;
  int32_t *cursor = rcnt + M;
  __shared__ long long part[1024];
  const int T = blockDim.x, t = threadIdx.x;
  const int chunk = (M + T - 1) / T;
  const int b0 = t * chunk, b1 = min(M, b0 + chunk);
  long long sum = 0;
  for (int i = b0; i < b1; ++i) sum += rcnt[i];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < T; off <<= 1) {
    long long v = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  long long run = part[t] - sum;  // exclusive prefix of this chunk
  for (int i = b0; i < b1; ++i) {
    s.row_off[i] = (int32_t)run;
    cursor[i] = (int32_t)run;
    run += rcnt[i];
  }
  if (t == T - 1) s.row_off[M] = (int32_t)part[T - 1];
  // Fixed-point exponent, one for the whole slab: the largest S with
  // 2^S * sum|f| + cnt/2 <= 2^31 - 1 for re and im in every row.  With one scale an
  // element's f is rounded to the same value in every row, so the deposit is exactly
  // that of f' = f + eps_e: for one element a pure scale of the image, and no
  // row-to-row pattern that interference between rows could amplify (per-row scales
  // left ~1e-5 relative errors at the weakest voxels of layered slabs).
  __shared__ int smin;
  if (t == 0) smin = 60;
  __syncthreads();
  int my_min = 60;
  for (int row = t; row < N; row += T) {
    double bound = 0.0;
    int cnt = 0;
    for (int e = 0; e < s.n_elem; ++e) {
      int c = ecnt[(int64_t)row * s.n_elem + e];
      double fr = fabs(s.f_table[2 * e]), fi = fabs(s.f_table[2 * e + 1]);
      bound += (double)c * fmax(fr, fi);
      cnt += c;
    }
    if (bound > 0.0) {
      double room = 2147483647.0 - 0.5 * (double)cnt - 1.0;
      const int S = ilogb(room / bound);  // floor(log2(room / bound))
      my_min = min(my_min, S);
    }
  }
  atomicMin(&smin, my_min);
  __syncthreads();
  for (int row = t; row < N; row += T) s.row_exp[row] = smin;
}

// avg_voxel_f of the deposited values f'_e = rint(f_e 2^S) 2^-S (one thread; n_elem is
// small): (sum_e n_e f'_e / volume) * r^3, the reference's operation order
// (tools/comparison.py:741-744) on the quantised table.
__global__ void k_slab_avg(const double *__restrict__ f_table, const int32_t *__restrict__ row_exp,
                           const int64_t *__restrict__ cnt, int64_t n_elem, double volume, double r3,
                           double *out) {
  if (threadIdx.x != 0) return;
  const double sc = ldexp(1.0, row_exp[0]), isc = ldexp(1.0, -row_exp[0]);
  double re = 0.0, im = 0.0;
  for (int64_t e = 0; e < n_elem; ++e) {
    const double n = (double)cnt[e];
    re = re + n * (rint(f_table[2 * e] * sc) * isc);
    im = im + n * (rint(f_table[2 * e + 1] * sc) * isc);
  }
  out[0] = (re / volume) * r3;
  out[1] = (im / volume) * r3;
}

// Placement: every atom of the block takes a rank within its row from the LDS
// counts; the block then reserves one contiguous range per row with a single
// global atomic on the row's cursor and writes the atoms there.
// LDS: [N + 1] block counts | [N + 1] block bases.
template <int PT>
__global__ __launch_bounds__(BLK) void k_slab_scatter(gx_slab s) {
  constexpr int SORT_CHUNK = sort_chunk<PT>(), SORT_PER_THREAD = PT;
  extern __shared__ int hist[];
  const int N = (int)s.N, M = N + 1;
  int *cnt = hist, *base = hist + M;
  for (int i = threadIdx.x; i < M; i += BLK) cnt[i] = 0;
  __syncthreads();
  const double rinv = 1.0 / s.r;
  int32_t *rcnt = s.work + (int64_t)N * s.n_elem;
  int32_t *cursor = rcnt + M;
  const int64_t a0 = blockIdx.x * (int64_t)SORT_CHUNK;
  int row[SORT_PER_THREAD], rank[SORT_PER_THREAD];
#pragma unroll
  for (int k = 0; k < SORT_PER_THREAD; ++k) {
    const int64_t a = a0 + threadIdx.x + k * BLK;
    row[k] = a < s.n_atoms ? slab_row_of(s, a, rinv) : -1;
    rank[k] = row[k] >= 0 ? atomicAdd(&cnt[row[k]], 1) : 0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < M; i += BLK)
    if (cnt[i]) base[i] = atomicAdd(&cursor[i], cnt[i]);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SORT_PER_THREAD; ++k) {
    if (row[k] < 0) continue;
    const int64_t a = a0 + threadIdx.x + k * BLK;
    const int rw = row[k];
    const int pos = base[rw] + rank[k];
    const int e = s.elem[a];
    s.sx[pos] = s.x[a];
    s.sy[pos] = s.y[a];
    s.szrow[pos] = rw;
    int qr = 0, qi = 0;
    if (rw < N) {
      const double sc = ldexp(1.0, s.row_exp[rw]);
      qr = (int)rint(s.f_table[2 * e] * sc);
      qi = (int)rint(s.f_table[2 * e + 1] * sc);
    }
    s.sq[2 * (int64_t)pos] = qr;
    s.sq[2 * (int64_t)pos + 1] = qi;
    if (s.szf) {
      double t = 0.0;
      if (rw < N) {
        const double zz = s.z[a] - s.zmin;
        t = zz * rinv - (double)to_int(floor_div(zz, s.r, rinv));
        t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
      }
      s.szf[pos] = t;
    }
    if (s.sperm) s.sperm[pos] = (int32_t)a;
    if (s.selem) s.selem[pos] = e;
  }
}

// ------------------------------------------------------------------------------------
// stable regrouping: atoms keep their original order within a row.  Chunks of
// GX_STABLE_CHUNK atoms (original order) count their rows; a per-row scan over the
// chunks gives each chunk's first slot in every row; each chunk then places its
// atoms in order, 256 at a time, ranking equal rows within a wave by ballots and
// across waves in wave order.
// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLK) void k_slab_stable_count(gx_slab s) {
  extern __shared__ int hist[];  // [N + 1]
  const int M = (int)s.N + 1;
  const double rinv = 1.0 / s.r;
  for (int i = threadIdx.x; i < M; i += BLK) hist[i] = 0;
  __syncthreads();
  const int64_t a0 = (int64_t)blockIdx.x * GX_STABLE_CHUNK;
  const int64_t a1 = min(s.n_atoms, a0 + GX_STABLE_CHUNK);
  for (int64_t a = a0 + threadIdx.x; a < a1; a += BLK) atomicAdd(&hist[slab_row_of(s, a, rinv)], 1);
  __syncthreads();
  int32_t *tab = s.stable_work + (int64_t)blockIdx.x * M;
  for (int i = threadIdx.x; i < M; i += BLK) tab[i] = hist[i];
}

// per row: exclusive scan over the chunks, offset by the row's first slot
__global__ __launch_bounds__(BLK) void k_slab_stable_scan(gx_slab s, int64_t n_chunks) {
  const int M = (int)s.N + 1;
  const int row = blockIdx.x * BLK + threadIdx.x;
  if (row >= M) return;
  int32_t run = s.row_off[row];
  for (int64_t c = 0; c < n_chunks; ++c) {
    int32_t *t = s.stable_work + c * M + row;
    const int32_t v = *t;
    *t = run;
    run += v;
  }
}

__global__ __launch_bounds__(BLK) void k_slab_stable_scatter(gx_slab s) {
  extern __shared__ int cur[];  // [N + 1] next free slot of each row for this chunk
  const int N = (int)s.N, M = N + 1;
  const double rinv = 1.0 / s.r;
  const int lane = threadIdx.x & (GX_WAVE - 1), wave = threadIdx.x / GX_WAVE;
  const int32_t *tab = s.stable_work + (int64_t)blockIdx.x * M;
  for (int i = threadIdx.x; i < M; i += BLK) cur[i] = tab[i];
  __syncthreads();
  const int64_t a0 = (int64_t)blockIdx.x * GX_STABLE_CHUNK;
  const int64_t a1 = min(s.n_atoms, a0 + GX_STABLE_CHUNK);
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int64_t base = a0; base < a1; base += BLK) {
    const int64_t a = base + threadIdx.x;
    const bool valid = a < a1;
    const int row = valid ? slab_row_of(s, a, rinv) : -1;
    // rank among the lanes of this wave with the same row (lower lanes first)
    int rank = 0, cnt = 0;
    bool leader = false;
    unsigned long long rem = __ballot(valid);
    while (rem) {
      const int l = __ffsll((long long)rem) - 1;
      const int lr = __shfl(row, l, GX_WAVE);
      const unsigned long long m = __ballot(valid && row == lr);
      if (valid && row == lr) {
        rank = __popcll(m & lt);
        cnt = __popcll(m);
        leader = (lane == l);
      }
      rem &= ~m;
    }
    for (int w = 0; w < BLK / GX_WAVE; ++w) {
      if (wave == w && valid) {
        const int pos = cur[row] + rank;
        wave_sync();
        if (leader) cur[row] += cnt;
        s.sx[pos] = s.x[a];
        s.sy[pos] = s.y[a];
        s.szrow[pos] = row;
        int qr = 0, qi = 0;
        double t = 0.0;
        const int e = s.elem[a];
        if (row < N) {
          const double sc = ldexp(1.0, s.row_exp[row]);
          qr = (int)rint(s.f_table[2 * e] * sc);
          qi = (int)rint(s.f_table[2 * e + 1] * sc);
          const double zz = s.z[a] - s.zmin;
          t = zz * rinv - (double)to_int(floor_div(zz, s.r, rinv));
          t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
        }
        s.sq[2 * (int64_t)pos] = qr;
        s.sq[2 * (int64_t)pos + 1] = qi;
        if (s.szf) s.szf[pos] = t;
        s.sperm[pos] = (int32_t)a;
        if (s.selem) s.selem[pos] = e;
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------------------------
// per-frame statistics
// ------------------------------------------------------------------------------------
__global__ void k_frames_init(unsigned long long *ymin_key, int32_t *bounds, int64_t nf) {
  int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (f < nf) {
    ymin_key[3 * f] = ~0ull;
    ymin_key[3 * f + 1] = ~0ull;
    ymin_key[3 * f + 2] = 0ull;
    bounds[4 * f + 0] = INT_MAX;
    bounds[4 * f + 1] = INT_MIN;
    bounds[4 * f + 2] = INT_MAX;
    bounds[4 * f + 3] = INT_MIN;
  }
}

constexpr int STAT_CHUNK = 16384;  // atoms per block in the statistics kernels

__device__ __forceinline__ void block_minmax4(int (&v)[4], int *sh) {
  v[0] = wave_min_i(v[0]);
  v[1] = wave_max_i(v[1]);
  v[2] = wave_min_i(v[2]);
  v[3] = wave_max_i(v[3]);
  const int lane = threadIdx.x & (GX_WAVE - 1), w = threadIdx.x / GX_WAVE;
  if (lane == 0)
    for (int k = 0; k < 4; ++k) sh[4 * w + k] = v[k];
  __syncthreads();
  for (int i = 1; i < BLK / GX_WAVE; ++i) {
    v[0] = min(v[0], sh[4 * i + 0]);
    v[1] = max(v[1], sh[4 * i + 1]);
    v[2] = min(v[2], sh[4 * i + 2]);
    v[3] = max(v[3], sh[4 * i + 3]);
  }
}

// min y' = x sin(phi) + y cos(phi) over all atoms (canonical rotate_coords_z,
// voxelgrids.py:319-323) and min / max y' over the atoms whose z row is inside the
// grid.  Grid (atom chunks, groups of YMIN_FRAMES frames): each atom is read once per
// group and rotated for every frame of it; one atomic per block, frame and key.
// min/max of doubles are exact, so the result does not depend on the order.
constexpr int YMIN_FRAMES = 8;
constexpr int YMIN_PER_THREAD = 8;                   // atom loads in flight per thread
constexpr int YMIN_ROUNDS = 4;                       // rounds per block (amortise the reduction)
constexpr int YMIN_CHUNK = YMIN_ROUNDS * YMIN_PER_THREAD * BLK;

__global__ __launch_bounds__(BLK) void k_frames_ymin(gx_frames d, int64_t f0, int64_t nf) {
  constexpr int NW = BLK / GX_WAVE, K = 3 * YMIN_FRAMES;
  __shared__ double sh[NW][K];
  const int64_t fg = (int64_t)blockIdx.y * YMIN_FRAMES;
  const int nfg = (int)min((int64_t)YMIN_FRAMES, nf - fg);
  double sn[YMIN_FRAMES], cs[YMIN_FRAMES];
#pragma unroll
  for (int j = 0; j < YMIN_FRAMES; ++j) {
    const double *fp = d.fparams + (f0 + fg + min(j, nfg - 1)) * GX_FP_STRIDE;
    sn[j] = fp[GX_FP_SIN];
    cs[j] = fp[GX_FP_COS];
  }
  // atoms with a z row inside the grid come first (regrouped order): a round of a
  // thread is all rowed, all not, or (once) straddles; rowed atoms feed mz / Mz,
  // the others mo, and min y' over all atoms is min(mz, mo)
  const int64_t n_rowed = d.row_off[d.N];
  double mo[YMIN_FRAMES], mz[YMIN_FRAMES], Mz[YMIN_FRAMES];
#pragma unroll
  for (int j = 0; j < YMIN_FRAMES; ++j) {
    mo[j] = INFINITY;
    mz[j] = INFINITY;
    Mz[j] = -INFINITY;
  }
  for (int rd = 0; rd < YMIN_ROUNDS; ++rd) {
    const int64_t a0 = (int64_t)blockIdx.x * YMIN_CHUNK + (int64_t)rd * YMIN_PER_THREAD * BLK + threadIdx.x;
    if (a0 >= d.n_atoms) break;
    // every load of the round is issued before the first use
    double x[YMIN_PER_THREAD], y[YMIN_PER_THREAD];
#pragma unroll
    for (int k = 0; k < YMIN_PER_THREAD; ++k) {
      const int64_t a = a0 + k * BLK;
      x[k] = a < d.n_atoms ? d.ax[a] : 0.0;
      y[k] = a < d.n_atoms ? d.ay[a] : 0.0;
    }
    const int64_t a_last = min(a0 + (YMIN_PER_THREAD - 1) * BLK, (int64_t)d.n_atoms - 1);
    if (a_last < n_rowed) {  // all rowed
#pragma unroll
      for (int k = 0; k < YMIN_PER_THREAD; ++k) {
        if (a0 + k * BLK >= d.n_atoms) break;
#pragma unroll
        for (int j = 0; j < YMIN_FRAMES; ++j) {
          const double v = x[k] * sn[j] + y[k] * cs[j];
          mz[j] = fmin(mz[j], v);
          Mz[j] = fmax(Mz[j], v);
        }
      }
    } else if (a0 >= n_rowed) {  // none rowed
#pragma unroll
      for (int k = 0; k < YMIN_PER_THREAD; ++k) {
        if (a0 + k * BLK >= d.n_atoms) break;
#pragma unroll
        for (int j = 0; j < YMIN_FRAMES; ++j) mo[j] = fmin(mo[j], x[k] * sn[j] + y[k] * cs[j]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < YMIN_PER_THREAD; ++k) {
        const int64_t a = a0 + k * BLK;
        if (a >= d.n_atoms) break;
#pragma unroll
        for (int j = 0; j < YMIN_FRAMES; ++j) {
          const double v = x[k] * sn[j] + y[k] * cs[j];
          if (a < n_rowed) {
            mz[j] = fmin(mz[j], v);
            Mz[j] = fmax(Mz[j], v);
          } else {
            mo[j] = fmin(mo[j], v);
          }
        }
      }
    }
  }
  double m[YMIN_FRAMES];
#pragma unroll
  for (int j = 0; j < YMIN_FRAMES; ++j) m[j] = fmin(mz[j], mo[j]);
  const int lane = threadIdx.x & (GX_WAVE - 1), w = threadIdx.x / GX_WAVE;
#pragma unroll
  for (int j = 0; j < YMIN_FRAMES; ++j) {
    const double a = wave_min(m[j]), b = wave_min(mz[j]), c = -wave_min(-Mz[j]);
    if (lane == 0) {
      sh[w][3 * j] = a;
      sh[w][3 * j + 1] = b;
      sh[w][3 * j + 2] = c;
    }
  }
  __syncthreads();
  if (threadIdx.x < 3 * nfg) {
    const int j = threadIdx.x / 3, kind = threadIdx.x - 3 * j;
    double v = sh[0][threadIdx.x];
    for (int i = 1; i < NW; ++i) v = kind == 2 ? fmax(v, sh[i][threadIdx.x]) : fmin(v, sh[i][threadIdx.x]);
    unsigned long long *key = d.ymin_key + 3 * (fg + j) + kind;
    if (kind == 2) {
      if (v > -INFINITY) atomicMax(key, dkey(v));
    } else if (v < INFINITY) {
      atomicMin(key, dkey(v));
    }
  }
}

// Index bounds from the extremes (floor division is monotone): valid when every
// atom with z row < N also has y row < N and every atom's z row is < N; false for
// frames where that fails, which take the per-atom pass below.
__device__ __forceinline__ bool frame_bounds_fast(const gx_frames &d, int fb, int32_t (&b)[4]) {
  const unsigned long long *k = d.ymin_key + 3 * fb;
  if (d.zrow_max < 0) return false;
  const double ymin = dkey_inv(k[0]);
  const double rinv = 1.0 / d.r;
  const long long lo = to_int(floor_div(dkey_inv(k[1]) - ymin, d.r, rinv));
  const long long hi = to_int(floor_div(dkey_inv(k[2]) - ymin, d.r, rinv));
  if (hi >= d.N) return false;
  b[0] = (int32_t)lo;
  b[1] = (int32_t)hi;
  b[2] = 0;
  b[3] = (int32_t)d.zrow_max;
  return true;
}

// y/z index bounds over valid atoms (voxelgrids.py:328-335, 346-349): from the
// extremes when the fast path above settles the frame (block 0 writes them), else
// per atom.  Every block of a frame takes the same branch.
__global__ __launch_bounds__(BLK) void k_frames_bounds(gx_frames d, int64_t f0, int64_t nf) {
  __shared__ int sh[4 * (BLK / GX_WAVE)];
  const int N = (int)d.N;
  const double rinv = 1.0 / d.r;
  const int fb = blockIdx.y;
  if (d.ymin_key[3 * fb + 1] == ~0ull) return;  // no atom with a z row inside the grid
  int32_t fast[4];
  if (frame_bounds_fast(d, fb, fast)) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
      for (int k = 0; k < 4; ++k) d.idx_bounds[4 * fb + k] = fast[k];
    return;
  }
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const double s = fp[GX_FP_SIN], c = fp[GX_FP_COS];
  const double ymin = dkey_inv(d.ymin_key[3 * fb]);
  const int64_t n_rowed = d.row_off[N];
  const int64_t a0 = (int64_t)blockIdx.x * STAT_CHUNK;
  const int64_t a1 = min(n_rowed, a0 + STAT_CHUNK);
  int v[4] = {INT_MAX, INT_MIN, INT_MAX, INT_MIN};
  for (int64_t a = a0 + threadIdx.x; a < a1; a += BLK) {
    const double yy = (d.ax[a] * s + d.ay[a] * c) - ymin;
    const long long yi = to_int(floor_div(yy, d.r, rinv));
    if (yi >= 0 && yi < N) {
      const int zr = d.azrow[a];
      v[0] = min(v[0], (int)yi);
      v[1] = max(v[1], (int)yi);
      v[2] = min(v[2], zr);
      v[3] = max(v[3], zr);
    }
  }
  block_minmax4(v, sh);
  if (threadIdx.x == 0 && v[0] != INT_MAX) {
    int32_t *b = d.idx_bounds + 4 * fb;
    atomicMin(&b[0], v[0]);
    atomicMax(&b[1], v[1]);
    atomicMin(&b[2], v[2]);
    atomicMax(&b[3], v[3]);
  }
}

}  // namespace

namespace {

// python slice [a:b] of a length-n axis, normalised to [lo, hi)
__device__ __forceinline__ void py_slice(long long a, long long b, long long n, long long &lo, long long &hi) {
  if (a < 0) a += n;
  if (a < 0) a = 0;
  if (a > n) a = n;
  if (b < 0) b += n;
  if (b < 0) b = 0;
  if (b > n) b = n;
  lo = a;
  hi = b > a ? b : a;
}

// gaussian_filter1d(mask, sigma, mode='wrap') of the indicator of [lo, hi):
// s[i] = sum_{d=-R..R} w[d] [lo <= (i + d) mod n < hi], evaluated from prefix
// sums of w over the (at most a few) runs of d whose wrapped index is inside.
__device__ double smooth_mask(const double *cw, int R, long long lo, long long hi, int i, int n) {
  if (hi <= lo) return 0.0;
  double acc = 0.0;
  const long long x0 = (long long)i - R, x1 = (long long)i + R;  // window [x0, x1]
  // copies [lo + m n, hi + m n) of the interval that meet the window
  long long m0 = (x0 - hi) / n - 1, m1 = (x1 - lo) / n + 1;
  for (long long m = m0; m <= m1; ++m) {
    long long a = lo + m * n, b = hi + m * n - 1;  // inclusive
    if (a < x0) a = x0;
    if (b > x1) b = x1;
    if (a > b) continue;
    acc += cw[b - i + R + 1] - cw[a - i + R];
  }
  return acc;
}

__global__ __launch_bounds__(BLK) void k_frames_columns(gx_frames d, int64_t f0, int64_t nf) {
  const int N = (int)d.N;
  const int fb = blockIdx.y;
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (fb >= nf || i >= N) return;
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const int32_t *b = d.idx_bounds + 4 * fb;
  if (i == 0 && b[0] == INT_MAX) atomicAdd(&d.status[0], 1);
  if (d.fill_bkg) {
    const double rinv = 1.0 / d.r;
    double x = (double)i * d.r;
    double len = collapse_length(fp, x);
    double nm = d.max_voxels - floor_div(len, d.r, rinv);
    d.col_nm[(int64_t)fb * N + i] = (int32_t)to_int(nm);
  }
  if (d.smooth) {
    long long lo, hi;
    py_slice((long long)b[0] + d.smooth, (long long)b[1] - d.smooth, N, lo, hi);
    d.col_sy[(int64_t)fb * N + i] = smooth_mask(d.gauss_cw, (int)d.smooth_radius, lo, hi, i, N);
    py_slice((long long)b[2] + d.smooth, (long long)b[3] - d.smooth, N, lo, hi);
    d.col_sz[(int64_t)fb * N + i] = smooth_mask(d.gauss_cw, (int)d.smooth_radius, lo, hi, i, N);
  }
}

// ------------------------------------------------------------------------------------
// per-frame background spectrum
//
// With my[y] = smooth ? s_y[y] : 1, mz = smooth ? s_z[z] : 1, A = avg_voxel_f *
// max_voxels and b[y] = (y_min < y < y_max ? num_missing[y] * avg_voxel_f : A),
// the grid of a row inside the filled region is
//   G[z, y] = (S[z, y] + b[y]) my[y] mz + A (1 - mz my[y])        (smooth)
//   G[z, y] =  S[z, y] + b[y]                                     (no smooth)
// (without fill b = 0, with S = atoms), so its transform along y is
//   F[z, k] = mz (FFT(S my)[k] + U[k]) + A N delta_k0 [smooth],
//   U = FFT((fill ? b : 0) my - (smooth ? A my : 0)),
// and rows outside the filled z range are the constant A.  U is one FP64 FFT
// per frame; only rows that hold atoms need a transform of their own.
// ------------------------------------------------------------------------------------
template <int N, typename C>
__global__ __launch_bounds__(GX_WAVE) void k_frames_spectra(gx_frames d, int64_t f0) {
  GX_RESOLVE_AVG(d);
  using T = typename Real<C>::T;
  constexpr int P = N / GX_WAVE;
  constexpr int R0 = FftFirst<N>::R, RL = FftLast<N>::R;
  __shared__ double2 lds[N];
  const int lane = threadIdx.x;
  const int fb = blockIdx.x;
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wf = jhi - jlo + 1;
  if (wf <= 0 || wf > d.W) return;
  const int32_t *bnd = d.idx_bounds + 4 * fb;
  const int ylo = bnd[0], yhi = bnd[1];
  const double avv_re = d.avg_re * d.max_voxels, avv_im = d.avg_im * d.max_voxels;
  const int32_t *nm_row = d.col_nm + (int64_t)fb * N;
  const double *sy_row = d.col_sy + (int64_t)fb * N;
  double2 v[P];
#pragma unroll
  for (int t = 0; t < P / R0; ++t) {
#pragma unroll
    for (int q = 0; q < R0; ++q) {
      const int y = lane + GX_WAVE * t + q * (N / R0);
      const double my = d.smooth ? sy_row[y] : 1.0;
      double ure = 0.0, uim = 0.0;
      if (d.fill_bkg) {
        double bre = avv_re, bim = avv_im;
        if (y > ylo && y < yhi) {
          const double nm = (double)nm_row[y];
          bre = nm * d.avg_re;
          bim = nm * d.avg_im;
        }
        ure = bre * my;
        uim = bim * my;
      }
      if (d.smooth) {
        ure = ure - avv_re * my;
        uim = uim - avv_im * my;
      }
      v[t * R0 + q] = make_double2(ure, uim);
    }
  }
  fft_wave<N, double2>(v, lds, reinterpret_cast<const double2 *>(d.twiddle64), lane);
  C *spec = reinterpret_cast<C *>(d.col_spec) + (int64_t)fb * d.W;
#pragma unroll
  for (int t = 0; t < P / RL; ++t) {
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      const int k = lane + GX_WAVE * t + q * (N / RL);
      const int w = ((k + N / 2) & (N - 1)) - jlo;
      if (w >= 0 && w < wf) spec[w] = C{(T)v[t * RL + q].x, (T)v[t * RL + q].y};
    }
  }
}

// ------------------------------------------------------------------------------------
// row pass
// ------------------------------------------------------------------------------------
constexpr int SCATTER_UNROLL = 4;

template <int N, int RPW, typename C, int PL = 0>
__global__ __launch_bounds__(BLK) void k_rowpass(gx_frames d, int64_t f0, int64_t nf) {
  GX_RESOLVE_AVG(d);
  using T = typename Real<C>::T;
  constexpr int P = N / GX_WAVE;
  constexpr int RB = (BLK / GX_WAVE) * RPW;  // rows per block
  constexpr int R0 = FftFirst<N, PL>::R, RL = FftLast<N, PL>::R;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *smem = reinterpret_cast<C *>(smem_raw);
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int wave = threadIdx.x / GX_WAVE;
  constexpr int WB = fft_kept_buf<N>();   // per-wave transform buffer (complex values)
  C *wbuf = smem + wave * WB;
  constexpr int SP = RB + 1;               // row stride of the staging tile (odd)
  C *stage = smem + (BLK / GX_WAVE) * WB;  // [W][SP]
  C *tspec = stage + d.W * SP;            // [W] background spectrum of this frame
  int32_t *acc_re = reinterpret_cast<int32_t *>(wbuf);
  int32_t *acc_im = acc_re + N;

  // tile order: frame major (the rows of one frame together), or XCD-aware with the
  // frames of a row block together (row_variant bit 2)
  const int64_t nzb = gridDim.x / nf;
  int fb, zb;
  if (d.row_variant & 4) {
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    fb = (int)(tile % nf);
    zb = (int)(tile / nf);
  } else {
    fb = (int)(blockIdx.x / nzb);
    zb = (int)(blockIdx.x % nzb);
  }
  const int64_t f = f0 + fb;
  const double *fp = d.fparams + f * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wf = jhi - jlo + 1;
  if (wf <= 0) return;  // block-uniform
  if (wf > d.W) {       // planner guarantees wf <= W; never overrun the staging tile
    if (threadIdx.x == 0 && zb == 0) atomicAdd(&d.status[1], 1);
    return;
  }
  const bool dense = d.fill_bkg || d.smooth;
  if (dense) {
    const C *spec = reinterpret_cast<const C *>(d.col_spec) + (int64_t)fb * d.W;
    for (int w = threadIdx.x; w < wf; w += BLK) tspec[w] = spec[w];
  }
  __syncthreads();

  const double s = fp[GX_FP_SIN], c = fp[GX_FP_COS];
  const double ymin = dkey_inv(d.ymin_key[3 * fb]);
  const int32_t *bnd = d.idx_bounds + 4 * fb;
  const int ylo = bnd[0], yhi = bnd[1], zlo = bnd[2], zhi = bnd[3];
  // atoms that land in y columns overwritten by the fill never contribute
  const int ya = d.fill_bkg ? (ylo == INT_MAX ? N : ylo + 1) : 0;   // keep ya <= y < yb
  const int yb = d.fill_bkg ? (yhi == INT_MIN ? 0 : yhi) : N;
  const double r = d.r, rinv = 1.0 / d.r;
  const C *tw = reinterpret_cast<const C *>(d.twiddle);
  const int z0 = zb * RB;

  // y-mask of this lane's first-stage inputs (same for every row of the frame)
  T my[P];
#pragma unroll
  for (int t = 0; t < P / R0; ++t)
#pragma unroll
    for (int q = 0; q < R0; ++q)
      my[t * R0 + q] = d.smooth ? (T)d.col_sy[(int64_t)fb * N + lane + GX_WAVE * t + q * (N / R0)] : (T)1;

  for (int rr = 0; rr < RPW; ++rr) {
    const int rb = wave * RPW + rr;
    const int z = z0 + rb;
    const bool inside = !d.fill_bkg || (z >= zlo && z < zhi);
    const int a0 = d.row_off[z], a1 = d.row_off[z + 1];
    const bool atoms = inside && a1 > a0;
    C v[P];
    if (atoms) {
      // 1) zero the fixed-point accumulators
      for (int y = lane; y < N; y += GX_WAVE) {
        acc_re[y] = 0;
        acc_im[y] = 0;
      }
      wave_sync();
      // 2) scatter this row's atoms (np.add.at(grid, (z, y), f), voxelgrids.py:341);
      //    loads for SCATTER_UNROLL atoms are issued before any of them is used
      for (int a = a0 + lane; a < a1; a += GX_WAVE * SCATTER_UNROLL) {
        double xs[SCATTER_UNROLL], ys[SCATTER_UNROLL];
        int2 qs[SCATTER_UNROLL];
#pragma unroll
        for (int u = 0; u < SCATTER_UNROLL; ++u) {
          const int au = a + u * GX_WAVE;
          const bool ok = au < a1;
          xs[u] = ok ? d.ax[au] : 0.0;
          ys[u] = ok ? d.ay[au] : 0.0;
          qs[u] = ok ? reinterpret_cast<const int2 *>(d.aq)[au] : make_int2(0, 0);
        }
#pragma unroll
        for (int u = 0; u < SCATTER_UNROLL; ++u) {
          if (a + u * GX_WAVE >= a1) break;
          const double yy = (xs[u] * s + ys[u] * c) - ymin;
          const int yi = floor_div_in(yy, r, rinv, ya, yb);
          if (yi >= 0) {
            atomicAdd(&acc_re[yi], qs[u].x);
            atomicAdd(&acc_im[yi], qs[u].y);
          }
        }
      }
      wave_sync();
      // 3) first-stage inputs S[y] my[y], exact fixed point -> T
      const T inv = (T)ldexp(1.0, -d.row_exp[z]);
#pragma unroll
      for (int t = 0; t < P / R0; ++t) {
#pragma unroll
        for (int q = 0; q < R0; ++q) {
          const int y = lane + GX_WAVE * t + q * (N / R0);
          const T m = my[t * R0 + q];
          v[t * R0 + q] = C{((T)acc_re[y] * inv) * m, ((T)acc_im[y] * inv) * m};
        }
      }
      wave_sync();
      // 4) FFT along y (outputs of the kept band)
      fft_wave_kept<N, C, PL>(v, wbuf, tw, lane, jlo - N / 2, jhi - N / 2);
    } else {
#pragma unroll
      for (int q = 0; q < P; ++q) v[q] = C{(T)0, (T)0};
    }
    // 5) kept columns [jlo, jhi]: add the background spectrum, stage
    const double mz = d.smooth ? d.col_sz[(int64_t)fb * N + z] : 1.0;
    stage_row_band<N, C, PL>(v, lane, jlo, wf, inside, d, mz, tspec, stage, SP, rb);
    wave_sync();
  }
  __syncthreads();
  // 6) transposed store: inter[fb][w][z0 .. z0+RB)
  C *out = reinterpret_cast<C *>(d.inter) + (int64_t)fb * d.W * d.Z;
  for (int idx = threadIdx.x; idx < wf * RB; idx += BLK) {
    const int w = idx / RB, q = idx - w * RB;
    out[(int64_t)w * d.Z + z0 + q] = stage[w * SP + q];
  }
}

// ------------------------------------------------------------------------------------
// row pass, block-wide variant: a block owns RB consecutive z rows of one frame; all
// 256 threads deposit the atoms of those rows at once (more loads in flight than
// one wave streaming one row), then each wave transforms its rows.  Same arithmetic
// and outputs as k_rowpass; the autotuner picks between the two shapes.
// ------------------------------------------------------------------------------------
template <int N, int RB, typename C, int PL = 0>
__global__ __launch_bounds__(BLK) void k_rowpass_blk(gx_frames d, int64_t f0, int64_t nf) {
  GX_RESOLVE_AVG(d);
  using T = typename Real<C>::T;
  constexpr int P = N / GX_WAVE;
  constexpr int R0 = FftFirst<N, PL>::R;
  constexpr int ROWB = (fft_kept_buf<N>() * (int)sizeof(C) > N * 8) ? fft_kept_buf<N>() * (int)sizeof(C)
                                                                     : N * 8;  // bytes per row
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int SP = RB + 1;                                        // staging row stride (odd)
  C *stage = reinterpret_cast<C *>(smem_raw + (size_t)RB * ROWB);  // [W][SP]
  C *tspec = stage + d.W * SP;                                      // [W]
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int wave = threadIdx.x / GX_WAVE;

  // tile order: frame major (the rows of one frame together), or XCD-aware with the
  // frames of a row block together (row_variant bit 2)
  const int64_t nzb = gridDim.x / nf;
  int fb, zb;
  if (d.row_variant & 4) {
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    fb = (int)(tile % nf);
    zb = (int)(tile / nf);
  } else {
    fb = (int)(blockIdx.x / nzb);
    zb = (int)(blockIdx.x % nzb);
  }
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wf = jhi - jlo + 1;
  if (wf <= 0) return;  // block-uniform
  if (wf > d.W) {
    if (threadIdx.x == 0 && zb == 0) atomicAdd(&d.status[1], 1);
    return;
  }
  const bool dense = d.fill_bkg || d.smooth;
  if (dense) {
    const C *spec = reinterpret_cast<const C *>(d.col_spec) + (int64_t)fb * d.W;
    for (int w = threadIdx.x; w < wf; w += BLK) tspec[w] = spec[w];
  }
  const int z0 = zb * RB;
  for (int i = threadIdx.x; i < RB * 2 * N; i += BLK) {
    const int row = i / (2 * N), k = i - row * 2 * N;
    reinterpret_cast<int32_t *>(smem_raw + (size_t)row * ROWB)[k] = 0;
  }
  __syncthreads();

  const double s = fp[GX_FP_SIN], c = fp[GX_FP_COS];
  const double ymin = dkey_inv(d.ymin_key[3 * fb]);
  const int32_t *bnd = d.idx_bounds + 4 * fb;
  const int ylo = bnd[0], yhi = bnd[1], zlo = bnd[2], zhi = bnd[3];
  const int ya = d.fill_bkg ? (ylo == INT_MAX ? N : ylo + 1) : 0;  // keep ya <= y < yb
  const int yb = d.fill_bkg ? (yhi == INT_MIN ? 0 : yhi) : N;
  const double r = d.r, rinv = 1.0 / d.r;
  // 1) deposit the atoms of rows z0 .. z0+RB-1 that lie inside the filled z range
  const int a_lo = d.row_off[z0], a_hi = d.row_off[z0 + RB];
  for (int a = a_lo + threadIdx.x; a < a_hi; a += BLK) {
    const int v = d.azrow[a];
    if (d.fill_bkg && !(v >= zlo && v < zhi)) continue;
    const double yy = (d.ax[a] * s + d.ay[a] * c) - ymin;
    const int yi = floor_div_in(yy, r, rinv, ya, yb);
    if (yi >= 0) {
      int32_t *acc = reinterpret_cast<int32_t *>(smem_raw + (size_t)(v - z0) * ROWB);
      const int2 q = reinterpret_cast<const int2 *>(d.aq)[a];
      atomicAdd(&acc[yi], q.x);
      atomicAdd(&acc[N + yi], q.y);
    }
  }
  __syncthreads();
  // 2) per row: smoothed-edge mask, FFT along y, kept band
  T my[P];
#pragma unroll
  for (int t = 0; t < P / R0; ++t)
#pragma unroll
    for (int q = 0; q < R0; ++q)
      my[t * R0 + q] = d.smooth ? (T)d.col_sy[(int64_t)fb * N + lane + GX_WAVE * t + q * (N / R0)] : (T)1;
  const C *tw = reinterpret_cast<const C *>(d.twiddle);
  for (int rb = wave; rb < RB; rb += BLK / GX_WAVE) {
    const int z = z0 + rb;
    const bool inside = !d.fill_bkg || (z >= zlo && z < zhi);
    const bool atoms = inside && d.row_off[z + 1] > d.row_off[z];
    unsigned char *rowp = smem_raw + (size_t)rb * ROWB;
    const int32_t *acc = reinterpret_cast<const int32_t *>(rowp);
    C v[P];
    if (atoms) {
      const T inv = (T)ldexp(1.0, -d.row_exp[z]);
#pragma unroll
      for (int t = 0; t < P / R0; ++t)
#pragma unroll
        for (int q = 0; q < R0; ++q) {
          const int y = lane + GX_WAVE * t + q * (N / R0);
          const T m = my[t * R0 + q];
          v[t * R0 + q] = C{((T)acc[y] * inv) * m, ((T)acc[N + y] * inv) * m};
        }
      wave_sync();
      fft_wave_kept<N, C, PL>(v, reinterpret_cast<C *>(rowp), tw, lane, jlo - N / 2, jhi - N / 2);
    } else {
#pragma unroll
      for (int q = 0; q < P; ++q) v[q] = C{(T)0, (T)0};
    }
    const double mz = d.smooth ? d.col_sz[(int64_t)fb * N + z] : 1.0;
    stage_row_band<N, C, PL>(v, lane, jlo, wf, inside, d, mz, tspec, stage, SP, rb);
  }
  __syncthreads();
  C *out = reinterpret_cast<C *>(d.inter) + (int64_t)fb * d.W * d.Z;
  for (int idx = threadIdx.x; idx < wf * RB; idx += BLK) {
    const int w = idx / RB, q = idx - w * RB;
    out[(int64_t)w * d.Z + z0 + q] = stage[w * SP + q];
  }
}

// ------------------------------------------------------------------------------------
// row pass, row-major intermediate (N = 512, 1024, 2048; gx_frames.inter_rm = 1)
//
// Each wave takes RPWV consecutive z rows of one frame: LDS fixed-point deposit of the
// row's atoms, first-stage inputs S[y] my[y] (the same arithmetic as k_rowpass), the
// banded in-wave FFT along y (fft_wave_band: only the kept columns' outputs), and the
// kept band written straight from registers to inter[frame][z][w] (consecutive lanes,
// consecutive columns).  No staging tile and no block barrier: the waves of a block
// are independent.  The background spectrum, the z mask and the rows outside the filled
// z range are applied by the column pass (k_colpass_rm), with the arithmetic that
// stage_row_band applies in the column-major layout, so the bins are the same.  Rows
// inside the filled range without atoms are written as zeros.
// ------------------------------------------------------------------------------------
// Rows per wave: more rows amortise the per-frame prologue (the smoothed y mask in
// registers), but at N = 2048 a loop over rows keeps more than 256 registers live
// (spills), while one row per wave is straight-line code (measured on the 1.3M-atom
// slab: 0.61 ms per 119-frame launch with one row per wave, 1.08 ms with two).
template <int N> struct RowRm {
  static constexpr int RPWV = (N <= 1024) ? 8 : 1;
  static constexpr int RB = (BLK / GX_WAVE) * RPWV;        // rows per block
};

template <int N, typename C, int RPWV, int WPE, int PL = 0>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_rowpass_rm(gx_frames d, int64_t f0, int64_t nf) {
  using T = typename Real<C>::T;
  constexpr int P = N / GX_WAVE;
  constexpr int R0 = FftFirst<N, PL>::R, RL = FftLast<N, PL>::R;
  constexpr int BUF = fft_band_buf<N>();
  constexpr int RB = (BLK / GX_WAVE) * RPWV;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int wave = threadIdx.x / GX_WAVE;
  C *wbuf = reinterpret_cast<C *>(smem_raw) + wave * BUF;
  int32_t *acc_re = reinterpret_cast<int32_t *>(wbuf);
  int32_t *acc_im = acc_re + N;

  // tile order: frame major, or XCD-aware with the frames of a row block together (bit 2)
  const int64_t nzb = gridDim.x / nf;
  int fb, zb;
  if (d.row_variant & 4) {
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    fb = (int)(tile % nf);
    zb = (int)(tile / nf);
  } else {
    fb = (int)(blockIdx.x / nzb);
    zb = (int)(blockIdx.x % nzb);
  }
  const int64_t f = f0 + fb;
  const double *fp = d.fparams + f * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wf = jhi - jlo + 1;
  if (wf <= 0) return;  // block-uniform
  if (wf > d.W) {       // planner guarantees wf <= W; never write past a row
    if (threadIdx.x == 0 && zb == 0) atomicAdd(&d.status[1], 1);
    return;
  }
  const double s = fp[GX_FP_SIN], c = fp[GX_FP_COS];
  const double ymin = dkey_inv(d.ymin_key[3 * fb]);
  const int32_t *bnd = d.idx_bounds + 4 * fb;
  const int ylo = bnd[0], yhi = bnd[1], zlo = bnd[2], zhi = bnd[3];
  const int ya = d.fill_bkg ? (ylo == INT_MAX ? N : ylo + 1) : 0;  // keep ya <= y < yb
  const int yb = d.fill_bkg ? (yhi == INT_MIN ? 0 : yhi) : N;
  const double r = d.r, rinv = 1.0 / d.r;
  const C *tw = reinterpret_cast<const C *>(d.twiddle);
  const int klo = jlo - N / 2, khi = jhi - N / 2;  // kept signed frequencies
  C *out = reinterpret_cast<C *>(d.inter) + (int64_t)fb * d.Z * d.W;

  T my[P];
#pragma unroll
  for (int t = 0; t < P / R0; ++t)
#pragma unroll
    for (int q = 0; q < R0; ++q)
      my[t * R0 + q] = d.smooth ? (T)d.col_sy[(int64_t)fb * N + lane + GX_WAVE * t + q * (N / R0)] : (T)1;

  const int zw = zb * RB + wave * RPWV;
  for (int rr = 0; rr < RPWV; ++rr) {
    const int z = zw + rr;
    if (z >= d.Z) break;
    if (d.fill_bkg && !(z >= zlo && z < zhi)) continue;  // synthesised by the column pass
    C *orow = out + (int64_t)z * d.W;
    const int a0 = d.row_off[z], a1 = d.row_off[z + 1];
    if (a1 <= a0) {
      for (int w = lane; w < wf; w += GX_WAVE) orow[w] = C{(T)0, (T)0};
      continue;
    }
    // 1) zero the fixed-point accumulators (2N int32, 16 bytes per store)
#pragma unroll
    for (int i = 0; i < (2 * N) / (4 * GX_WAVE); ++i)
      reinterpret_cast<int4 *>(acc_re)[lane + GX_WAVE * i] = make_int4(0, 0, 0, 0);
    wave_sync();
    // 2) scatter this row's atoms (np.add.at(grid, (z, y), f), voxelgrids.py:341)
    for (int a = a0 + lane; a < a1; a += GX_WAVE * SCATTER_UNROLL) {
      double xs[SCATTER_UNROLL], ys[SCATTER_UNROLL];
      int2 qs[SCATTER_UNROLL];
#pragma unroll
      for (int u = 0; u < SCATTER_UNROLL; ++u) {
        const int au = a + u * GX_WAVE;
        const bool ok = au < a1;
        xs[u] = ok ? d.ax[au] : 0.0;
        ys[u] = ok ? d.ay[au] : 0.0;
        qs[u] = ok ? reinterpret_cast<const int2 *>(d.aq)[au] : make_int2(0, 0);
      }
#pragma unroll
      for (int u = 0; u < SCATTER_UNROLL; ++u) {
        if (a + u * GX_WAVE >= a1) break;
        const double yy = (xs[u] * s + ys[u] * c) - ymin;
        const int yi = floor_div_in(yy, r, rinv, ya, yb);
        if (yi >= 0) {
          atomicAdd(&acc_re[yi], qs[u].x);
          atomicAdd(&acc_im[yi], qs[u].y);
        }
      }
    }
    wave_sync();
    // 3) first-stage inputs S[y] my[y], exact fixed point -> T
    const T inv = (T)ldexp(1.0, -d.row_exp[z]);
    C v[P];
#pragma unroll
    for (int t = 0; t < P / R0; ++t)
#pragma unroll
      for (int q = 0; q < R0; ++q) {
        const int y = lane + GX_WAVE * t + q * (N / R0);
        const T m = my[t * R0 + q];
        v[t * R0 + q] = C{((T)acc_re[y] * inv) * m, ((T)acc_im[y] * inv) * m};
      }
    wave_sync();
    // 4) FFT along y, outputs of the kept band only
    fft_wave_band<N, C, PL>(v, wbuf, tw, lane, klo, khi);
    // 5) kept band, row-major
#pragma unroll
    for (int t = 0; t < P / RL; ++t)
#pragma unroll
      for (int q = 0; q < RL; ++q) {
        const int k = lane + GX_WAVE * t + q * (N / RL);
        const int w = ((k + N / 2) & (N - 1)) - jlo;
        if (w >= 0 && w < wf) orow[w] = v[t * RL + q];
      }
  }
}

// ------------------------------------------------------------------------------------
// column pass over the row-major intermediate (gx_frames.inter_rm = 1)
//
// A block of WAVES waves takes WAVES consecutive kept columns of one frame.  Phase 1:
// the block reads the [Z][WAVES] tile of the row pass's output (WAVES consecutive
// complex values per row) and turns each element into the column transform's input:
// rows outside the filled z range are the constant A N at the k = 0 column (else 0);
// inside, with fill or smoothing, mz[z] (raw + U[w]) (+ A N at k = 0 when smoothing) --
// stage_row_band's arithmetic.  Phase 2: each wave loads its column (rows >= Z from
// col_input), the tile's LDS becomes the waves' FFT buffers, and the banded FFT along z
// is followed by |F|^2 and the binning of k_colpass.
// ------------------------------------------------------------------------------------
// Four waves per block: with eight, a block's load phase and transform phase alternate in
// step with too few blocks per CU to overlap them (N = 1024 FP32: 0.128 -> 0.116 ms per
// 213-frame launch with four; two tie four).
template <int N, typename C> struct ColRm {
  static constexpr int WAVES = 4;
};

template <int N, typename C, int WAVES, int PL = 0>
__global__ __launch_bounds__(WAVES * GX_WAVE) void k_colpass_rm(gx_frames d, int64_t f0) {
  GX_RESOLVE_AVG(d);
  using T = typename Real<C>::T;
  constexpr int P = N / GX_WAVE;
  constexpr int R0 = FftFirst<N, PL>::R, RL = FftLast<N, PL>::R;
  constexpr int BUF = fft_band_buf<N>();
  constexpr int TS = WAVES + 1;  // tile row stride (odd: a wave's column reads spread over banks)
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *tile = reinterpret_cast<C *>(smem_raw);
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int wave = threadIdx.x / GX_WAVE;
  // work item = (frame, column group), numbered frame-major and handed to the XCDs in
  // contiguous ranges: the groups of a frame share the 128-byte lines of each row of the
  // intermediate, so they run on one XCD and read each line into one L2
  const int ng = (int)((d.W + WAVES - 1) / WAVES);
  const int64_t item = xcd_tile(blockIdx.x, gridDim.x);
  const int fb = (int)(item / ng);
  const int64_t f = f0 + fb;
  const double *fp = d.fparams + f * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wf = jhi - jlo + 1;
  const int w0 = (int)(item - (int64_t)fb * ng) * WAVES;
  if (wf <= 0 || wf > d.W || w0 >= wf) return;  // block-uniform, before any barrier
  const int32_t *bnd = d.idx_bounds + 4 * fb;
  const int zlo = bnd[2], zhi = bnd[3];
  const bool dense = d.fill_bkg || d.smooth;
  const C a0 = C{(T)(d.avg_re * d.max_voxels * (double)N), (T)(d.avg_im * d.max_voxels * (double)N)};
  const C *spec = reinterpret_cast<const C *>(d.col_spec) + (int64_t)fb * d.W;
  const double *sz_row = d.col_sz + (int64_t)fb * N;
  const C *src = reinterpret_cast<const C *>(d.inter) + (int64_t)fb * d.Z * d.W + w0;

  // phase 1: the tile of column inputs for rows z < Z.  Thread t takes column t % WAVES
  // and rows t / WAVES + 64 i; every load is issued before the first is used.
  {
    constexpr int NIT = N / GX_WAVE;  // Z <= N
    const int cl = threadIdx.x % WAVES, zt = threadIdx.x / WAVES;
    const int w = w0 + cl;
    const bool kcol = w < wf;
    const bool k0 = kcol && jlo + w == N / 2;
    const C outside = k0 ? a0 : C{(T)0, (T)0};
    // (branches, not conditional expressions, around the loads: a conditional whose
    // operands are a memory element and a local value may be compiled to a select of
    // addresses and a flat load through the private aperture)
    C u = C{(T)0, (T)0};
    if (kcol && dense) u = spec[w];
    constexpr int CH = NIT < 8 ? NIT : 8;  // rows in flight per thread
    for (int i0 = 0; i0 < NIT && zt + GX_WAVE * i0 < d.Z; i0 += CH) {
      C x[CH];
      double mz[CH];
#pragma unroll
      for (int ii = 0; ii < CH; ++ii) {
        const int z = zt + GX_WAVE * (i0 + ii);
        const bool ld = kcol && z < d.Z && (!d.fill_bkg || (z >= zlo && z < zhi));
        x[ii] = outside;
        if (ld) x[ii] = src[(int64_t)z * d.W + cl];
        mz[ii] = (ld && d.smooth) ? sz_row[z] : 1.0;
      }
#pragma unroll
      for (int ii = 0; ii < CH; ++ii) {
        const int z = zt + GX_WAVE * (i0 + ii);
        if (z >= d.Z) break;
        const bool inside = kcol && (!d.fill_bkg || (z >= zlo && z < zhi));
        if (inside && dense) {
          const T mzt = (T)mz[ii];
          T ore = mzt * (x[ii].x + u.x);
          T oim = mzt * (x[ii].y + u.y);
          if (d.smooth && k0) {
            ore = ore + a0.x;
            oim = oim + a0.y;
          }
          x[ii] = C{ore, oim};
        }
        tile[z * TS + cl] = x[ii];
      }
    }
  }
  __syncthreads();
  // phase 2: this wave's column
  const int wcol = w0 + wave;
  const bool active = wcol < wf;
  const int j = jlo + wcol;
  const bool k0 = (j == N / 2);
  C u = C{(T)0, (T)0};
  if (active && !d.fill_bkg && d.smooth) u = spec[wcol];
  C v[P];
#pragma unroll
  for (int t = 0; t < P / R0; ++t)
#pragma unroll
    for (int q = 0; q < R0; ++q) {
      const int z = lane + GX_WAVE * t + q * (N / R0);
      if (z < d.Z)
        v[t * R0 + q] = tile[z * TS + wave];
      else
        v[t * R0 + q] = col_input(d, (const C *)nullptr, z, k0, u, sz_row);
    }
  __syncthreads();  // the tile is read; its LDS becomes the FFT buffers
  if (!active) return;
  // det_h_qx / det_h_qy of this column (numpy linspace: j*step + start, last = stop)
  const double hx = (j == N - 1) ? fp[GX_FP_HX_STOP] : (double)j * fp[GX_FP_HX_STEP] + fp[GX_FP_HX_START];
  const double hy = (j == N - 1) ? fp[GX_FP_HY_STOP] : (double)j * fp[GX_FP_HY_STEP] + fp[GX_FP_HY_START];
  const double dqinv = 1.0 / d.dq;
  const long long bx = to_int(floor_div(hx - d.qmin, d.dq, dqinv)) - d.c_lo;
  const long long by = to_int(floor_div(hy - d.qmin, d.dq, dqinv)) - d.c_lo;
  if (bx < 0 || bx >= d.c_n || by < 0 || by >= d.c_n) return;
  const C *tw = reinterpret_cast<const C *>(d.twiddle);
  fft_wave_band<N, C, PL>(v, reinterpret_cast<C *>(smem_raw) + wave * BUF, tw, lane, (int)d.i_lo - N / 2,
                      (int)d.i_hi - N / 2);

  const int64_t cbase = (by * d.c_n + bx) * d.c_n;
  const double val = d.q_val;
#pragma unroll
  for (int t = 0; t < P / RL; ++t) {
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      const int k = lane + GX_WAVE * t + q * (N / RL);
      const int i = (k + N / 2) & (N - 1);
      if (i < d.i_lo || i > d.i_hi) continue;
      // det_v_qz[i] = fftshift(fftfreq(N, r) * 2 * pi)[i]
      const int kint = (k < N / 2) ? k : k - N;
      const double qz = (((double)kint * val) * 2.0) * M_PI;
      const long long bz = to_int(floor_div(qz - d.qmin, d.dq, dqinv)) - d.c_lo;
      if (bz < 0 || bz >= d.c_n) continue;
      const double I = (double)cnorm(v[t * RL + q]);
      const unsigned long long fx = (unsigned long long)llrint(I * d.bin_scale);
      atomicAdd(&d.bin_sum[cbase + bz], fx);
      atomicAdd(&d.bin_cnt[cbase + bz], 1);
    }
  }
}

// ------------------------------------------------------------------------------------
// column pass
// ------------------------------------------------------------------------------------
template <int N, typename C, int PL = 0>
__global__ __launch_bounds__(BLK) void k_colpass(gx_frames d, int64_t f0) {
  GX_RESOLVE_AVG(d);
  constexpr int P = N / GX_WAVE;
  constexpr int R0 = FftFirst<N, PL>::R, RL = FftLast<N, PL>::R;
  __shared__ C lds[BLK / GX_WAVE][fft_kept_buf<N>()];
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int wave = threadIdx.x / GX_WAVE;
  const int fb = blockIdx.y;
  const int64_t f = f0 + fb;
  const double *fp = d.fparams + f * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wcol = blockIdx.x * (BLK / GX_WAVE) + wave;
  if (wcol > jhi - jlo || wcol >= d.W) return;  // wave-uniform; no block barrier below
  const int j = jlo + wcol;
  // det_h_qx / det_h_qy of this column (numpy linspace: j*step + start, last = stop)
  const double hx = (j == N - 1) ? fp[GX_FP_HX_STOP] : (double)j * fp[GX_FP_HX_STEP] + fp[GX_FP_HX_START];
  const double hy = (j == N - 1) ? fp[GX_FP_HY_STOP] : (double)j * fp[GX_FP_HY_STEP] + fp[GX_FP_HY_START];
  const double dqinv = 1.0 / d.dq;
  const long long bx = to_int(floor_div(hx - d.qmin, d.dq, dqinv)) - d.c_lo;
  const long long by = to_int(floor_div(hy - d.qmin, d.dq, dqinv)) - d.c_lo;
  if (bx < 0 || bx >= d.c_n || by < 0 || by >= d.c_n) return;

  using T = typename Real<C>::T;
  const C *col = reinterpret_cast<const C *>(d.inter) + ((int64_t)fb * d.W + wcol) * d.Z;
  const bool k0 = (j == N / 2);
  C u = C{(T)0, (T)0};
  if (!d.fill_bkg && d.smooth) u = reinterpret_cast<const C *>(d.col_spec)[(int64_t)fb * d.W + wcol];
  const double *sz_row = d.col_sz + (int64_t)fb * N;
  C v[P];
#pragma unroll
  for (int t = 0; t < P / R0; ++t)
#pragma unroll
    for (int q = 0; q < R0; ++q) v[t * R0 + q] = col_input(d, col, lane + GX_WAVE * t + q * (N / R0), k0, u, sz_row);
  const C *tw = reinterpret_cast<const C *>(d.twiddle);
  fft_wave_kept<N, C, PL>(v, lds[wave], tw, lane, (int)d.i_lo - N / 2, (int)d.i_hi - N / 2);

  const int64_t cbase = (by * d.c_n + bx) * d.c_n;
  const double val = d.q_val;
  // the kept band of this column's transform for the differentiable backward
  C *fband = d.f_band ? reinterpret_cast<C *>(d.f_band) + ((int64_t)fb * d.W + wcol) * (d.i_hi - d.i_lo + 1)
                      : nullptr;
#pragma unroll
  for (int t = 0; t < P / RL; ++t) {
#pragma unroll
    for (int q = 0; q < RL; ++q) {
      const int k = lane + GX_WAVE * t + q * (N / RL);
      const int i = (k + N / 2) & (N - 1);
      if (i < d.i_lo || i > d.i_hi) continue;
      if (fband) fband[i - d.i_lo] = v[t * RL + q];
      // det_v_qz[i] = fftshift(fftfreq(N, r) * 2 * pi)[i]
      const int kint = (k < N / 2) ? k : k - N;
      const double qz = (((double)kint * val) * 2.0) * M_PI;
      const long long bz = to_int(floor_div(qz - d.qmin, d.dq, dqinv)) - d.c_lo;
      if (bz < 0 || bz >= d.c_n) continue;
      const double I = (double)cnorm(v[t * RL + q]);
      const unsigned long long fx = (unsigned long long)llrint(I * d.bin_scale);
      atomicAdd(&d.bin_sum[cbase + bz], fx);
      atomicAdd(&d.bin_cnt[cbase + bz], 1);
    }
  }
}

// ------------------------------------------------------------------------------------
// grid edges that are not powers of two: zoomed chirp-z (Bluestein) transforms
//
// With nk = (n^2 + k^2 - (k - n)^2) / 2 for any integers, X[k] = c[k] sum_n (x[n] c[n])
// conj(c[k - n]), c[m] = exp(-i pi m^2 / N) = c[-m].  Only a band of signed
// frequencies k = k_lo + j, j < K, is kept, so the sum is a linear convolution of the
// N inputs with the lags t = j - n in [-(N-1), K): a circular convolution of length
// M >= N + K - 1 (a power of two) evaluated as IFFT_M(FFT_M(x c) H), H = FFT_M of
// h[t] = conj(c[k_lo + t]) wrapped (host, FP64, already divided by M); the inverse is
// conj(FFT(conj(.))).  Output j of the inverse transform is X[k_lo + j].
// ------------------------------------------------------------------------------------
// per-frame background spectrum on the kept columns by a direct FP64 DFT
__global__ __launch_bounds__(BLK) void k_frames_spectra_dft(gx_frames d, int64_t f0) {
  GX_RESOLVE_AVG(d);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double2 *u = reinterpret_cast<double2 *>(smem_raw);  // [N]
  const int N = (int)d.N;
  const int fb = blockIdx.x;
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wf = jhi - jlo + 1;
  if (wf <= 0 || wf > d.W) return;
  const int32_t *bnd = d.idx_bounds + 4 * fb;
  const int ylo = bnd[0], yhi = bnd[1];
  const double avv_re = d.avg_re * d.max_voxels, avv_im = d.avg_im * d.max_voxels;
  for (int y = threadIdx.x; y < N; y += BLK) {
    const double my = d.smooth ? d.col_sy[(int64_t)fb * N + y] : 1.0;
    double ure = 0.0, uim = 0.0;
    if (d.fill_bkg) {
      double bre = avv_re, bim = avv_im;
      if (y > ylo && y < yhi) {
        const double nm = (double)d.col_nm[(int64_t)fb * N + y];
        bre = nm * d.avg_re;
        bim = nm * d.avg_im;
      }
      ure = bre * my;
      uim = bim * my;
    }
    if (d.smooth) {
      ure = ure - avv_re * my;
      uim = uim - avv_im * my;
    }
    u[y] = make_double2(ure, uim);
  }
  __syncthreads();
  const double2 *tw = reinterpret_cast<const double2 *>(d.twiddle64);
  for (int w = threadIdx.x; w < wf; w += BLK) {
    const int k = (jlo + w + N - N / 2) % N;  // unshifted frequency of column jlo + w
    double re = 0.0, im = 0.0;
    int m = 0;
    for (int y = 0; y < N; ++y) {
      const double2 t = tw[m], x = u[y];
      re = fma(x.x, t.x, fma(-x.y, t.y, re));
      im = fma(x.x, t.y, fma(x.y, t.x, im));
      m += k;
      if (m >= N) m -= N;
    }
    if (d.fft_f64)
      reinterpret_cast<double2 *>(d.col_spec)[(int64_t)fb * d.W + w] = make_double2(re, im);
    else
      reinterpret_cast<float2 *>(d.col_spec)[(int64_t)fb * d.W + w] = make_float2((float)re, (float)im);
  }
}

// Waves per SIMD the FP32 chirp-z row pass is compiled for, up to M = 2048: at 4 (128 VGPRs,
// a 92-byte spill) a fit evaluation at N = 1048 measured ~10% faster than at 3 (155 VGPRs,
// no spill); FP64 and M = 4096 keep the compiler's choice.
// Elements of C at the start of the chirp-z row pass's LDS: the length-M transform
// exchange, which the row's 2 N int32 accumulators alias (N may exceed M when the
// band's convolution is sized to the rows the atoms reach).
template <int M, typename C> __host__ __device__ constexpr int bs_row_fbuf(int N) {
  return (2 * N * 4 + (int)sizeof(C) - 1) / (int)sizeof(C) > M ? (2 * N * 4 + (int)sizeof(C) - 1) / (int)sizeof(C)
                                                                 : M;
}

template <int M, typename C> struct RowBsWaves { static constexpr int value = (sizeof(C) == 8 && M <= 2048) ? 4 : 1; };
template <int M, typename C>
__global__ __launch_bounds__(GX_BLK_FFT) __attribute__((amdgpu_waves_per_eu(RowBsWaves<M, C>::value)))
void k_rowpass_bs(gx_frames d, int64_t f0) {
  GX_RESOLVE_AVG(d);
  using T = typename Real<C>::T;
  constexpr int P = M / GX_BLK_FFT;
  constexpr int R = BlkPlan<M>::R;
  constexpr int RB = 8;  // rows per block, transformed one after another by the whole block
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *fbuf = reinterpret_cast<C *>(smem_raw);  // [M] transform exchange; aliased by the accumulators
  int32_t *acc_re = reinterpret_cast<int32_t *>(fbuf);
  const int N = (int)d.N;
  int32_t *acc_im = acc_re + N;
  C *stage = fbuf + bs_row_fbuf<M, C>(N);  // [W][RB]
  C *tspec = stage + d.W * RB;             // [W]
  const int tid = threadIdx.x;
  const int fb = blockIdx.y;
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wf = jhi - jlo + 1;
  if (wf <= 0) return;
  // the planner sizes W and the chirp-z band to every frame's kept columns, and the
  // band's convolution length to the y rows the frames' atoms reach: inputs at
  // y >= bs_M - bs_rk + 1 would alias (and y >= bs_M would not be read at all)
  const int32_t *bnd = d.idx_bounds + 4 * fb;
  if (wf > d.W || jlo < d.bs_rlo + N / 2 || jhi > d.bs_rlo + d.bs_rk - 1 + N / 2 ||
      (bnd[1] != INT_MIN && bnd[1] > d.bs_M - d.bs_rk)) {
    if (tid == 0 && blockIdx.x == 0) atomicAdd(&d.status[1], 1);
    return;
  }
  const bool dense = d.fill_bkg || d.smooth;
  if (dense) {
    const C *spec = reinterpret_cast<const C *>(d.col_spec) + (int64_t)fb * d.W;
    for (int w = tid; w < wf; w += GX_BLK_FFT) tspec[w] = spec[w];
  }
  const double s = fp[GX_FP_SIN], c = fp[GX_FP_COS];
  const double ymin = dkey_inv(d.ymin_key[3 * fb]);
  const int ylo = bnd[0], yhi = bnd[1], zlo = bnd[2], zhi = bnd[3];
  const int ya = d.fill_bkg ? (ylo == INT_MAX ? N : ylo + 1) : 0;
  const int yb = d.fill_bkg ? (yhi == INT_MIN ? 0 : yhi) : N;
  const double r = d.r, rinv = 1.0 / d.r;
  const double avv_re = d.avg_re * d.max_voxels, avv_im = d.avg_im * d.max_voxels;
  const C *chirp = reinterpret_cast<const C *>(d.bs_chirp);
  const int z0 = blockIdx.x * RB;
  __syncthreads();
  for (int rb = 0; rb < RB; ++rb) {
    const int z = z0 + rb;
    const bool inside = !d.fill_bkg || (z >= zlo && z < zhi);
    const int a0 = d.row_off[z], a1 = d.row_off[z + 1];
    const bool atoms = inside && a1 > a0;  // block-uniform
    // the tables are the same for every row: an opaque copy of their addresses per row
    // keeps the compiler from holding all of a thread's table values in registers
    // across the row loop (204 VGPRs, two waves per SIMD, for M = 2048 in FP32)
    const C *chirp_r = chirp, *kern_r = reinterpret_cast<const C *>(d.bs_kernel),
            *tw_r = reinterpret_cast<const C *>(d.bs_twiddle);
    const double *sy_r = d.col_sy;
    asm volatile("" : "+s"(chirp_r), "+s"(kern_r), "+s"(tw_r), "+s"(sy_r));
    C v[P];
    if (atoms) {
      for (int y = tid; y < N; y += GX_BLK_FFT) {
        acc_re[y] = 0;
        acc_im[y] = 0;
      }
      __syncthreads();
      for (int a = a0 + tid; a < a1; a += GX_BLK_FFT) {
        const double yy = (d.ax[a] * s + d.ay[a] * c) - ymin;
        const int yi = floor_div_in(yy, r, rinv, ya, yb);
        if (yi >= 0) {
          atomicAdd(&acc_re[yi], d.aq[2 * (int64_t)a]);
          atomicAdd(&acc_im[yi], d.aq[2 * (int64_t)a + 1]);
        }
      }
      __syncthreads();
      const T inv = (T)ldexp(1.0, -d.row_exp[z]);
#pragma unroll
      for (int t = 0; t < P / R; ++t)
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int n = tid + GX_BLK_FFT * t + q * (M / R);
          C x = C{(T)0, (T)0};
          if (n < N) {
            const T m = d.smooth ? (T)sy_r[(int64_t)fb * N + n] : (T)1;
            x = cmul(C{((T)acc_re[n] * inv) * m, ((T)acc_im[n] * inv) * m}, chirp_r[n]);
          }
          v[t * R + q] = x;
        }
      __syncthreads();
      chirpz_block<M, C>(v, fbuf, kern_r, tw_r, tid);
    } else {
#pragma unroll
      for (int q = 0; q < P; ++q) v[q] = C{(T)0, (T)0};
    }
    const double mz = d.smooth ? d.col_sz[(int64_t)fb * N + z] : 1.0;
#pragma unroll
    for (int t = 0; t < P / R; ++t)
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int j = tid + GX_BLK_FFT * t + q * (M / R);
        if (j >= (int)d.bs_rk) continue;
        const int k = (int)d.bs_rlo + j;   // signed frequency, -N/2 <= k < N - N/2
        const int w = k + N / 2 - jlo;     // fftshifted column within the frame's band
        if (w < 0 || w >= wf) continue;
        C o = atoms ? cmul(v[t * R + q], chirp_r[k < 0 ? -k : k]) : C{(T)0, (T)0};
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
        stage[w * RB + rb] = o;
      }
    __syncthreads();
  }
  C *out = reinterpret_cast<C *>(d.inter) + (int64_t)fb * d.W * d.Z;
  for (int idx = tid; idx < wf * RB; idx += GX_BLK_FFT) {
    const int w = idx / RB, q = idx - w * RB;
    out[(int64_t)w * d.Z + z0 + q] = stage[idx];
  }
}

template <int M, typename C>
__global__ __launch_bounds__(GX_BLK_FFT) void k_colpass_bs(gx_frames d, int64_t f0) {
  GX_RESOLVE_AVG(d);
  using T = typename Real<C>::T;
  constexpr int P = M / GX_BLK_FFT;
  constexpr int R = BlkPlan<M>::R;
  __shared__ C fbuf[M];
  const int N = (int)d.N;
  const int tid = threadIdx.x;
  const int fb = blockIdx.y;
  const double *fp = d.fparams + (f0 + fb) * GX_FP_STRIDE;
  const int jlo = (int)fp[GX_FP_JLO], jhi = (int)fp[GX_FP_JHI];
  const int wcol = blockIdx.x;
  if (wcol > jhi - jlo || wcol >= d.W) return;  // block-uniform
  const int j = jlo + wcol;
  const double hx = (j == N - 1) ? fp[GX_FP_HX_STOP] : (double)j * fp[GX_FP_HX_STEP] + fp[GX_FP_HX_START];
  const double hy = (j == N - 1) ? fp[GX_FP_HY_STOP] : (double)j * fp[GX_FP_HY_STEP] + fp[GX_FP_HY_START];
  const double dqinv = 1.0 / d.dq;
  const long long bx = to_int(floor_div(hx - d.qmin, d.dq, dqinv)) - d.c_lo;
  const long long by = to_int(floor_div(hy - d.qmin, d.dq, dqinv)) - d.c_lo;
  if (bx < 0 || bx >= d.c_n || by < 0 || by >= d.c_n) return;
  const C *col = reinterpret_cast<const C *>(d.inter) + ((int64_t)fb * d.W + wcol) * d.Z;
  const bool k0 = (j == N / 2);
  C u = C{(T)0, (T)0};
  if (!d.fill_bkg && d.smooth) u = reinterpret_cast<const C *>(d.col_spec)[(int64_t)fb * d.W + wcol];
  const double *sz_row = d.col_sz + (int64_t)fb * N;
  const C *chirp = reinterpret_cast<const C *>(d.bs_chirp);
  // A band shorter than N + bs_ck - 1 takes the rows z < Lc = bs_Mc - bs_ck + 1 (>= Z,
  // checked on the host): rows >= Z are zero in every column but the k = 0 column with
  // fill, whose constant A N on the rows [Lc, N) is added below in closed form.
  const int Lc = (int)min((int64_t)N, d.bs_Mc - d.bs_ck + 1);
  C v[P];
#pragma unroll
  for (int t = 0; t < P / R; ++t)
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int z = tid + GX_BLK_FFT * t + q * (M / R);
      v[t * R + q] = (z < Lc) ? cmul(col_input(d, col, z, k0, u, sz_row), chirp[z]) : C{(T)0, (T)0};
    }
  chirpz_block<M, C>(v, fbuf, reinterpret_cast<const C *>(d.bs_kernel_c),
                     reinterpret_cast<const C *>(d.bs_twiddle_c), tid);
  const bool tail = k0 && d.fill_bkg && Lc < N;
  const int64_t cbase = (by * d.c_n + bx) * d.c_n;
  const double val = d.q_val;
  const int nn = (N - 1) / 2 + 1;  // fftfreq: k < nn are the non-negative frequencies
#pragma unroll
  for (int t = 0; t < P / R; ++t)
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int j = tid + GX_BLK_FFT * t + q * (M / R);
      if (j >= (int)d.bs_ck) continue;
      const int k = (int)d.bs_clo + j;     // signed frequency, -N/2 <= k < N - N/2
      const int i = k + N / 2;             // fftshifted row
      if (i < d.i_lo || i > d.i_hi) continue;
      const int kk = k < 0 ? k + N : k;
      const int kint = (kk < nn) ? kk : kk - N;
      const double qz = (((double)kint * val) * 2.0) * M_PI;
      const long long bz = to_int(floor_div(qz - d.qmin, d.dq, dqinv)) - d.c_lo;
      if (bz < 0 || bz >= d.c_n) continue;
      C F = cmul(v[t * R + q], chirp[k < 0 ? -k : k]);
      if (tail) {
        const double2 gt = chirpz_fill_tail(d, kk, Lc);
        F = C{(T)((double)F.x + gt.x), (T)((double)F.y + gt.y)};
      }
      const double I = (double)cnorm(F);
      const unsigned long long fx = (unsigned long long)llrint(I * d.bin_scale);
      atomicAdd(&d.bin_sum[cbase + bz], fx);
      atomicAdd(&d.bin_cnt[cbase + bz], 1);
    }
}

template <int M, typename C>
int launch_rowpass_bs_t(const gx_frames *d, int64_t f0, int64_t nf, hipStream_t st) {
  constexpr int RB = 8;
  size_t lds = ((size_t)bs_row_fbuf<M, C>((int)d->N) + (size_t)d->W * (RB + 1)) * sizeof(C);
  if (lds > 160 * 1024)
    return set_error(GX_ERR_UNSUPPORTED, "chirp-z row pass needs %zu bytes of LDS (W=%lld)", lds, (long long)d->W);
  static bool attr_set = false;
  if (!attr_set) {
    GX_HIP(hipFuncSetAttribute((const void *)k_rowpass_bs<M, C>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               160 * 1024));
    attr_set = true;
  }
  dim3 grid((unsigned)(d->Z / RB), (unsigned)nf);
  k_rowpass_bs<M, C><<<grid, GX_BLK_FFT, lds, st>>>(*d, f0);
  return check_launch("k_rowpass_bs");
}

template <typename C>
int launch_bs(const gx_frames *d, int64_t f0, int64_t nf, hipStream_t st, bool row) {
  if (row) {
    switch (d->bs_M) {
      case 1024: return launch_rowpass_bs_t<1024, C>(d, f0, nf, st);
      case 2048: return launch_rowpass_bs_t<2048, C>(d, f0, nf, st);
      case 4096: return launch_rowpass_bs_t<4096, C>(d, f0, nf, st);
    }
  } else {
    dim3 grid((unsigned)d->W, (unsigned)nf);
    switch (d->bs_Mc) {
      case 1024: k_colpass_bs<1024, C><<<grid, GX_BLK_FFT, 0, st>>>(*d, f0); return check_launch("k_colpass_bs");
      case 2048: k_colpass_bs<2048, C><<<grid, GX_BLK_FFT, 0, st>>>(*d, f0); return check_launch("k_colpass_bs");
      case 4096: k_colpass_bs<4096, C><<<grid, GX_BLK_FFT, 0, st>>>(*d, f0); return check_launch("k_colpass_bs");
    }
  }
  return set_error(GX_ERR_UNSUPPORTED, "chirp-z length M=%lld", (long long)(row ? d->bs_M : d->bs_Mc));
}

// ------------------------------------------------------------------------------------
// bins
// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLK) void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) p[i] = v;
}

__global__ __launch_bounds__(BLK) void k_bins_clear(unsigned long long *s, int32_t *c, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    s[i] = 0;
    c[i] = 0;
  }
}

struct AffArgs {
  double a[9];
};

__global__ __launch_bounds__(BLK) void k_bins_finalize(const unsigned long long *__restrict__ bsum,
                                                       const int32_t *__restrict__ bcnt, int64_t c_n,
                                                       double inv_scale, int64_t c_lo, int64_t q_num,
                                                       double qa_start, double qa_step, double qa_stop,
                                                       AffArgs aff, double z_elem, double c16pi2,
                                                       double *iq64, float *iq32) {
  const int64_t n = c_n * c_n * c_n;
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    const int64_t iz = i % c_n, ix = (i / c_n) % c_n, iy = i / (c_n * c_n);
    const int32_t cnt = bcnt[i];
    double mean = 0.0;
    if (cnt != 0) mean = ((double)bsum[i] * inv_scale) / (double)cnt;
    const double qx = lin_axis(c_lo + ix, q_num, qa_start, qa_step, qa_stop);
    const double qy = lin_axis(c_lo + iy, q_num, qa_start, qa_step, qa_stop);
    const double qz = lin_axis(c_lo + iz, q_num, qa_start, qa_step, qa_stop);
    const double q2 = (qx * qx + qy * qy) + qz * qz;
    const double v = mean * f0_weight(aff.a, q2, z_elem, c16pi2);
    if (iq64) iq64[i] = v;
    if (iq32) iq32[i] = (float)v;
  }
}

template <int N, typename C, int PL>
int launch_rowpass_t(const gx_frames *d, int64_t f0, int64_t nf, hipStream_t st) {
  constexpr int RPW = (N <= 1024) ? 2 : 1;
  constexpr int RB = (BLK / GX_WAVE) * RPW;
  size_t lds = ((size_t)(BLK / GX_WAVE) * fft_kept_buf<N>() + (size_t)d->W * (RB + 2)) * sizeof(C);
  if (lds > 160 * 1024)
    return set_error(GX_ERR_UNSUPPORTED, "row pass needs %zu bytes of LDS (N=%d, W=%lld, %s)", lds, N,
                     (long long)d->W, sizeof(C) == 16 ? "FP64" : "FP32");
  static bool attr_set = false;
  if (!attr_set) {
    GX_HIP(hipFuncSetAttribute((const void *)k_rowpass<N, RPW, C, PL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               160 * 1024));
    attr_set = true;
  }
  const unsigned nblk = (unsigned)((d->Z / RB) * nf);
  k_rowpass<N, RPW, C, PL><<<nblk, BLK, lds, st>>>(*d, f0, nf);
  return check_launch("k_rowpass");
}

template <int N, int RB, typename C, int PL>
int launch_rowpass_blk_t(const gx_frames *d, int64_t f0, int64_t nf, hipStream_t st) {
  constexpr int ROWB = (fft_kept_buf<N>() * (int)sizeof(C) > N * 8) ? fft_kept_buf<N>() * (int)sizeof(C) : N * 8;
  const size_t lds = (size_t)RB * ROWB + (size_t)d->W * (RB + 2) * sizeof(C);
  if (lds > 160 * 1024)
    return set_error(GX_ERR_UNSUPPORTED, "block-wide row pass needs %zu bytes of LDS (N=%d, RB=%d)", lds, N, RB);
  static bool attr_set = false;
  if (!attr_set) {
    GX_HIP(hipFuncSetAttribute((const void *)k_rowpass_blk<N, RB, C, PL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               160 * 1024));
    attr_set = true;
  }
  const unsigned nblk = (unsigned)((d->Z / RB) * nf);
  k_rowpass_blk<N, RB, C, PL><<<nblk, BLK, lds, st>>>(*d, f0, nf);
  return check_launch("k_rowpass_blk");
}

// row_variant bits 0-1: 0 per-wave rows (2 rows per wave for N <= 1024), 1 block-wide
// deposit over 4 rows, 2 block-wide over 8 rows; bit 2: XCD-aware tile order
template <int N, typename C, int PL>
int launch_rowpass_v(const gx_frames *d, int64_t f0, int64_t nf, hipStream_t st) {
  switch (d->row_variant & 3) {
    case 1: return launch_rowpass_blk_t<N, 4, C, PL>(d, f0, nf, st);
    case 2: return launch_rowpass_blk_t<N, 8, C, PL>(d, f0, nf, st);
    default: return launch_rowpass_t<N, C, PL>(d, f0, nf, st);
  }
}

template <int N, typename C, int RPWV, int WPE, int PL>
int launch_rowpass_rm_x(const gx_frames *d, int64_t f0, int64_t nf, hipStream_t st) {
  constexpr int RB = (BLK / GX_WAVE) * RPWV;
  const size_t lds = (size_t)(BLK / GX_WAVE) * fft_band_buf<N>() * sizeof(C);
  static bool attr_set = false;
  if (!attr_set) {
    GX_HIP(hipFuncSetAttribute((const void *)k_rowpass_rm<N, C, RPWV, WPE, PL>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  const unsigned nblk = (unsigned)(((d->Z + RB - 1) / RB) * nf);
  k_rowpass_rm<N, C, RPWV, WPE, PL><<<nblk, BLK, lds, st>>>(*d, f0, nf);
  return check_launch("k_rowpass_rm");
}

template <int N, typename C, int PL>
int launch_rowpass_rm_t(const gx_frames *d, int64_t f0, int64_t nf, hipStream_t st) {
  return launch_rowpass_rm_x<N, C, RowRm<N>::RPWV, 1, PL>(d, f0, nf, st);
}

template <int N, typename C, int WAVES, int PL>
int launch_colpass_rm_w(const gx_frames *d, int64_t f0, int64_t nf, hipStream_t st) {
  const size_t tile = (size_t)d->Z * (WAVES + 1) * sizeof(C);
  const size_t bufs = (size_t)WAVES * fft_band_buf<N>() * sizeof(C);
  const size_t lds = tile > bufs ? tile : bufs;
  if (lds > 160 * 1024)
    return set_error(GX_ERR_UNSUPPORTED, "column pass needs %zu bytes of LDS (N=%d, Z=%lld)", lds, N,
                     (long long)d->Z);
  static bool attr_set = false;
  if (!attr_set) {
    GX_HIP(hipFuncSetAttribute((const void *)k_colpass_rm<N, C, WAVES, PL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               160 * 1024));
    attr_set = true;
  }
  const unsigned nblk = (unsigned)(grid1d(d->W, WAVES) * nf);
  k_colpass_rm<N, C, WAVES, PL><<<nblk, WAVES * GX_WAVE, lds, st>>>(*d, f0);
  return check_launch("k_colpass_rm");
}

template <int N, typename C, int PL>
int launch_colpass_rm_t(const gx_frames *d, int64_t f0, int64_t nf, hipStream_t st) {
  return launch_colpass_rm_w<N, C, ColRm<N, C>::WAVES, PL>(d, f0, nf, st);
}

template <int N, int PL>
int launch_rowpass_p(const gx_frames *d, int64_t f0, int64_t nf, hipStream_t st) {
  if constexpr (N != 256) {
    if (d->inter_rm) {
      if (d->fft_f64) return launch_rowpass_rm_t<N, double2, PL>(d, f0, nf, st);
      return launch_rowpass_rm_t<N, float2, PL>(d, f0, nf, st);
    }
  }
  if (d->fft_f64) return launch_rowpass_v<N, double2, PL>(d, f0, nf, st);
  return launch_rowpass_v<N, float2, PL>(d, f0, nf, st);
}

template <int N, int PL>
constexpr bool plan_exists() { return PL < gx_fft_plans(N); }

template <int N>
int launch_rowpass(const gx_frames *d, int64_t f0, int64_t nf, hipStream_t st) {
  if constexpr (plan_exists<N, 1>()) {
    if (d->fft_plan_row == 1) return launch_rowpass_p<N, 1>(d, f0, nf, st);
    if (d->fft_plan_row == 2) return launch_rowpass_p<N, 2>(d, f0, nf, st);
  }
  if constexpr (plan_exists<N, 3>()) {
    if (d->fft_plan_row == 3) return launch_rowpass_p<N, 3>(d, f0, nf, st);
  }
  return launch_rowpass_p<N, 0>(d, f0, nf, st);
}

template <int N, int PL>
int launch_colpass_p(const gx_frames *d, int64_t f0, int64_t nf, hipStream_t st) {
  if constexpr (N != 256) {
    if (d->inter_rm) {
      if (d->fft_f64) return launch_colpass_rm_t<N, double2, PL>(d, f0, nf, st);
      return launch_colpass_rm_t<N, float2, PL>(d, f0, nf, st);
    }
  }
  dim3 grid(grid1d(d->W, BLK / GX_WAVE), (unsigned)nf);
  if (d->fft_f64)
    k_colpass<N, double2, PL><<<grid, BLK, 0, st>>>(*d, f0);
  else
    k_colpass<N, float2, PL><<<grid, BLK, 0, st>>>(*d, f0);
  return check_launch("k_colpass");
}

template <int N>
int launch_colpass(const gx_frames *d, int64_t f0, int64_t nf, hipStream_t st) {
  if constexpr (plan_exists<N, 1>()) {
    if (d->fft_plan_col == 1) return launch_colpass_p<N, 1>(d, f0, nf, st);
    if (d->fft_plan_col == 2) return launch_colpass_p<N, 2>(d, f0, nf, st);
  }
  if constexpr (plan_exists<N, 3>()) {
    if (d->fft_plan_col == 3) return launch_colpass_p<N, 3>(d, f0, nf, st);
  }
  return launch_colpass_p<N, 0>(d, f0, nf, st);
}

template <int N>
int launch_spectra(const gx_frames *d, int64_t f0, int64_t nf, hipStream_t st) {
  if (d->fft_f64)
    k_frames_spectra<N, double2><<<(unsigned)nf, GX_WAVE, 0, st>>>(*d, f0);
  else
    k_frames_spectra<N, float2><<<(unsigned)nf, GX_WAVE, 0, st>>>(*d, f0);
  return check_launch("k_frames_spectra");
}

bool pow2_grid(int64_t N) { return N == 256 || N == 512 || N == 1024 || N == 2048; }

int check_frames(const gx_frames *d, int64_t f0, int64_t nf) {
  GX_REQUIRE(d != nullptr, "gx_frames descriptor is NULL");
  GX_REQUIRE(gx_supported_grid(d->N), "grid edge N=%lld outside [16, 2048]", (long long)d->N);
  GX_REQUIRE(pow2_grid(d->N) ? d->bs_M == 0 : (d->bs_M == 1024 || d->bs_M == 2048 || d->bs_M == 4096),
             "grid edge N=%lld needs bs_M = 0 for 256..2048 powers of two, else a chirp-z length "
             "1024/2048/4096 (got %lld)", (long long)d->N, (long long)d->bs_M);
  GX_REQUIRE(!d->bs_M || (d->bs_chirp && d->bs_kernel && d->bs_twiddle && d->bs_kernel_c && d->bs_twiddle_c),
             "chirp-z tables missing");
  GX_REQUIRE(!d->bs_M || (d->bs_Mc == 1024 || d->bs_Mc == 2048 || d->bs_Mc == 4096),
             "chirp-z column length %lld", (long long)d->bs_Mc);
  GX_REQUIRE(!d->bs_M || (d->bs_rk >= 1 && d->bs_rk <= d->N && d->bs_M >= d->bs_rk &&
                          d->bs_rlo >= -(d->N / 2) && d->bs_rlo + d->bs_rk <= d->N - d->N / 2),
             "chirp-z row band [%lld, +%lld) does not fit M=%lld or the frequencies of N=%lld",
             (long long)d->bs_rlo, (long long)d->bs_rk, (long long)d->bs_M, (long long)d->N);
  // a column band shorter than N + bs_ck - 1 must still take every row z < Z, and is exact only
  // where rows >= Z are zero in every column but the fill constant of the k = 0 column
  GX_REQUIRE(!d->bs_M || d->bs_Mc >= d->N + d->bs_ck - 1 ||
                 (d->bs_Mc - d->bs_ck + 1 >= d->Z && (d->fill_bkg || !d->smooth)),
             "chirp-z column band of length %lld covers rows below %lld, Z=%lld (and needs fill_bkg "
             "or no smoothing when shorter than N + bs_ck - 1)", (long long)d->bs_Mc,
             (long long)(d->bs_Mc - d->bs_ck + 1), (long long)d->Z);
  GX_REQUIRE(!d->bs_M || (d->bs_ck >= 1 && d->bs_ck <= d->N && d->bs_Mc >= d->bs_ck &&
                          d->bs_clo >= -(d->N / 2) && d->bs_clo + d->bs_ck <= d->N - d->N / 2 &&
                          d->bs_clo + d->N / 2 <= d->i_lo && d->bs_clo + d->bs_ck - 1 + d->N / 2 >= d->i_hi),
             "chirp-z column band [%lld, +%lld) does not fit M=%lld or cover rows [%lld, %lld]",
             (long long)d->bs_clo, (long long)d->bs_ck, (long long)d->bs_Mc, (long long)d->i_lo,
             (long long)d->i_hi);
  GX_REQUIRE(d->r > 0.0, "r_voxel_size must be positive");
  GX_REQUIRE(nf >= 0 && nf <= d->batch, "batch of %lld frames exceeds workspace batch %lld", (long long)nf,
             (long long)d->batch);
  GX_REQUIRE(f0 >= 0 && f0 + nf <= d->n_frames, "frames [%lld, %lld) outside [0, %lld)", (long long)f0,
             (long long)(f0 + nf), (long long)d->n_frames);
  GX_REQUIRE(d->W >= 1 && d->W <= d->N, "intermediate width W=%lld outside [1, N]", (long long)d->W);
  GX_REQUIRE(d->Z >= 8 && d->Z <= d->N && d->Z % 8 == 0, "row range Z=%lld must be a multiple of 8 in [8, N]",
             (long long)d->Z);
  GX_REQUIRE(d->c_n >= 1, "empty cropped grid");
  GX_REQUIRE(d->row_variant >= 0 && d->row_variant <= 6 && (d->row_variant & 3) != 3,
             "row_variant=%lld is not one of 0, 1, 2, 4, 5, 6", (long long)d->row_variant);
  {
    const int np = d->bs_M ? 1 : gx_fft_plans(d->N);
    GX_REQUIRE(d->fft_plan_row >= 0 && d->fft_plan_row < np && d->fft_plan_col >= 0 && d->fft_plan_col < np,
               "fft_plan_row / fft_plan_col (%lld, %lld) outside [0, %d) for N = %lld", (long long)d->fft_plan_row,
               (long long)d->fft_plan_col, np, (long long)d->N);
  }
  GX_REQUIRE(!d->inter_rm || ((d->N == 512 || d->N == 1024 || d->N == 2048) && (d->row_variant & 3) == 0),
             "the row-major intermediate needs N = 512, 1024 or 2048 and the per-wave row shape "
             "(N=%lld, row_variant=%lld)", (long long)d->N, (long long)d->row_variant);
  GX_REQUIRE(!d->f_band || (!d->bs_M && !d->inter_rm),
             "f_band (the stored kept band) needs a power-of-two N and inter_rm = 0");
  GX_REQUIRE(d->fparams && d->ymin_key && d->idx_bounds && d->inter && d->bin_sum && d->bin_cnt && d->twiddle &&
                 d->status && d->row_off && d->row_exp,
             "a required device buffer of gx_frames is NULL");
  GX_REQUIRE(!d->fill_bkg || d->col_nm, "fill_bkg needs col_nm");
  GX_REQUIRE(!d->smooth || (d->col_sy && d->col_sz && d->gauss_cw), "smooth needs col_sy, col_sz, gauss_cw");
  GX_REQUIRE(!(d->fill_bkg || d->smooth) || (d->col_spec && d->twiddle64),
             "fill_bkg / smooth need col_spec and twiddle64");
  GX_REQUIRE(d->n_atoms >= 0 && (d->n_atoms == 0 || (d->ax && d->ay && d->aq && d->azrow)), "atom arrays");
  return GX_OK;
}

}  // namespace

// =====================================================================================
// C ABI
// =====================================================================================
extern "C" {

int gx_abi_version(void) { return GX_ABI_VERSION; }

int gx_supported_grid(int64_t N) { return N >= 16 && N <= 2048; }

int gx_slab_bounds(const gx_slab *s, double *out_minmax, int64_t *out_elem_count, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(s && out_minmax && out_elem_count, "NULL argument");
  GX_REQUIRE(s->n_atoms > 0, "slab has no atoms");
  GX_REQUIRE(s->x && s->y && s->z && s->elem, "atom arrays are NULL");
  hipStream_t st = as_stream(stream);
  auto *keys = reinterpret_cast<unsigned long long *>(out_minmax);
  auto *cnts = reinterpret_cast<unsigned long long *>(out_elem_count);
  k_slab_bounds_init<<<1, 64, 0, st>>>(keys, cnts, s->n_elem);
  GX_REQUIRE(s->n_elem >= 1 && s->n_elem <= MAX_ELEM, "n_elem=%lld outside [1, %d]", (long long)s->n_elem,
             MAX_ELEM);
  // four atoms per thread: 108 blocks for 110k atoms (16 per thread left 27 blocks
  // waiting on their loads one after another)
  unsigned g = grid1d(s->n_atoms, BLK * 4);
  if (g > 2048) g = 2048;
  k_slab_bounds<4><<<g, BLK, 0, st>>>(s->x, s->y, s->z, s->elem, s->n_atoms, s->n_elem, keys, cnts);
  k_slab_bounds_fin<<<1, 64, 0, st>>>(keys);
  return check_launch("gx_slab_bounds");
}

int gx_slab_sort_rows(const gx_slab *s, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(s != nullptr, "NULL slab");
  GX_REQUIRE(s->n_atoms > 0 && s->n_atoms < (1ll << 31), "atom count %lld outside [1, 2^31)",
             (long long)s->n_atoms);
  GX_REQUIRE(s->N >= 1 && s->N <= 4096, "grid edge N=%lld outside [1, 4096]", (long long)s->N);
  GX_REQUIRE(s->r > 0.0, "r_voxel_size must be positive");
  GX_REQUIRE(s->n_elem >= 1 && s->n_elem <= 128, "n_elem=%lld outside [1, 128]", (long long)s->n_elem);
  GX_REQUIRE(s->x && s->y && s->z && s->elem && s->f_table && s->sx && s->sy && s->sq && s->szrow && s->row_off &&
                 s->row_exp && s->work,
             "a slab buffer is NULL");
  hipStream_t st = as_stream(stream);
  const int64_t nwork = s->N * s->n_elem + 2 * (s->N + 1);
  k_fill_i32<<<grid1d(nwork, BLK) > 4096 ? 4096 : grid1d(nwork, BLK), BLK, 0, st>>>(s->work, nwork, 0);
  constexpr int SPT = 2;  // atoms per thread of the row count and placement
  const unsigned g = grid1d(s->n_atoms, sort_chunk<SPT>());
  const int elem_lds = (s->N + 1) + s->N * s->n_elem <= SORT_LDS_INTS;
  const size_t lds_count = (size_t)((s->N + 1) + (elem_lds ? s->N * s->n_elem : 0)) * sizeof(int);
  k_slab_rowcount<SPT><<<g, BLK, lds_count, st>>>(*s, elem_lds);
  k_slab_scan<<<1, 1024, 0, st>>>(*s);
  if (s->sperm && s->stable_work) {
    const int64_t n_chunks = (s->n_atoms + GX_STABLE_CHUNK - 1) / GX_STABLE_CHUNK;
    const size_t lds = (size_t)(s->N + 1) * sizeof(int);
    k_slab_stable_count<<<(unsigned)n_chunks, BLK, lds, st>>>(*s);
    k_slab_stable_scan<<<grid1d(s->N + 1, BLK), BLK, 0, st>>>(*s, n_chunks);
    k_slab_stable_scatter<<<(unsigned)n_chunks, BLK, lds, st>>>(*s);
  } else {
    k_slab_scatter<SPT><<<g, BLK, 2 * (size_t)(s->N + 1) * sizeof(int), st>>>(*s);
  }
  return check_launch("gx_slab_sort_rows");
}

int gx_slab_avg(const gx_slab *s, const int64_t *elem_count, double volume, double r3, double *out,
                gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(s != nullptr && s->f_table && s->row_exp && elem_count && out, "NULL argument");
  GX_REQUIRE(s->n_elem >= 1 && s->n_elem <= 128, "n_elem=%lld outside [1, 128]", (long long)s->n_elem);
  GX_REQUIRE(volume > 0.0 && r3 > 0.0, "volume and r^3 must be positive");
  k_slab_avg<<<1, GX_WAVE, 0, as_stream(stream)>>>(s->f_table, s->row_exp, elem_count, s->n_elem, volume, r3, out);
  return check_launch("gx_slab_avg");
}

int gx_frames_stats(const gx_frames *d, int64_t f0, int64_t nf, gx_stream_t stream) {
  clear_error();
  int rc = check_frames(d, f0, nf);
  if (rc) return rc;
  if (nf == 0) return GX_OK;
  hipStream_t st = as_stream(stream);
  k_frames_init<<<grid1d(nf, BLK), BLK, 0, st>>>(d->ymin_key, d->idx_bounds, nf);
  if (d->n_atoms > 0) {
    dim3 gy(grid1d(d->n_atoms, YMIN_CHUNK), grid1d(nf, YMIN_FRAMES));
    k_frames_ymin<<<gy, BLK, 0, st>>>(*d, f0, nf);
    dim3 g(grid1d(d->n_atoms, STAT_CHUNK), (unsigned)nf);
    k_frames_bounds<<<g, BLK, 0, st>>>(*d, f0, nf);
  }
  return check_launch("gx_frames_stats");
}

int gx_frames_columns(const gx_frames *d, int64_t f0, int64_t nf, gx_stream_t stream) {
  clear_error();
  int rc = check_frames(d, f0, nf);
  if (rc) return rc;
  if (nf == 0) return GX_OK;
  hipStream_t st = as_stream(stream);
  dim3 g(grid1d(d->N, BLK), (unsigned)nf);
  k_frames_columns<<<g, BLK, 0, st>>>(*d, f0, nf);
  rc = check_launch("gx_frames_columns");
  if (rc || !(d->fill_bkg || d->smooth)) return rc;
  if (d->bs_M) {
    k_frames_spectra_dft<<<(unsigned)nf, BLK, (size_t)d->N * sizeof(double2), st>>>(*d, f0);
    return check_launch("k_frames_spectra_dft");
  }
  switch (d->N) {
    case 256: return launch_spectra<256>(d, f0, nf, st);
    case 512: return launch_spectra<512>(d, f0, nf, st);
    case 1024: return launch_spectra<1024>(d, f0, nf, st);
    case 2048: return launch_spectra<2048>(d, f0, nf, st);
  }
  return set_error(GX_ERR_UNSUPPORTED, "N=%lld", (long long)d->N);
}

int gx_frames_rowpass(const gx_frames *d, int64_t f0, int64_t nf, gx_stream_t stream) {
  clear_error();
  int rc = check_frames(d, f0, nf);
  if (rc) return rc;
  if (nf == 0) return GX_OK;
  hipStream_t st = as_stream(stream);
  if (d->bs_M) return d->fft_f64 ? launch_bs<double2>(d, f0, nf, st, true) : launch_bs<float2>(d, f0, nf, st, true);
  switch (d->N) {
    case 256: return launch_rowpass<256>(d, f0, nf, st);
    case 512: return launch_rowpass<512>(d, f0, nf, st);
    case 1024: return launch_rowpass<1024>(d, f0, nf, st);
    case 2048: return launch_rowpass<2048>(d, f0, nf, st);
  }
  return set_error(GX_ERR_UNSUPPORTED, "N=%lld", (long long)d->N);
}

int gx_frames_colpass(const gx_frames *d, int64_t f0, int64_t nf, gx_stream_t stream) {
  clear_error();
  int rc = check_frames(d, f0, nf);
  if (rc) return rc;
  if (nf == 0) return GX_OK;
  hipStream_t st = as_stream(stream);
  if (d->bs_M) return d->fft_f64 ? launch_bs<double2>(d, f0, nf, st, false) : launch_bs<float2>(d, f0, nf, st, false);
  switch (d->N) {
    case 256: return launch_colpass<256>(d, f0, nf, st);
    case 512: return launch_colpass<512>(d, f0, nf, st);
    case 1024: return launch_colpass<1024>(d, f0, nf, st);
    case 2048: return launch_colpass<2048>(d, f0, nf, st);
  }
  return set_error(GX_ERR_UNSUPPORTED, "N=%lld", (long long)d->N);
}

int gx_frames_forward(const gx_frames *d, int64_t f0, int64_t nf, gx_stream_t stream) {
  int rc = gx_frames_stats(d, f0, nf, stream);
  if (rc == GX_OK && (d->fill_bkg || d->smooth)) rc = gx_frames_columns(d, f0, nf, stream);
  if (rc == GX_OK) rc = gx_frames_rowpass(d, f0, nf, stream);
  if (rc == GX_OK) rc = gx_frames_colpass(d, f0, nf, stream);
  return rc;
}

int gx_bins_clear(unsigned long long *bin_sum, int32_t *bin_cnt, int64_t n_bins, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(bin_sum && bin_cnt && n_bins >= 0, "bins");
  if (n_bins == 0) return GX_OK;
  unsigned g = grid1d(n_bins, BLK);
  if (g > 8192) g = 8192;
  k_bins_clear<<<g, BLK, 0, as_stream(stream)>>>(bin_sum, bin_cnt, n_bins);
  return check_launch("gx_bins_clear");
}

int gx_bins_finalize(const unsigned long long *bin_sum, const int32_t *bin_cnt, int64_t c_n, double bin_scale,
                     int64_t c_lo, int64_t q_num, double qa_start, double qa_step, double qa_stop,
                     const double *host_aff, double z_elem, double *iq64, float *iq32, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(bin_sum && bin_cnt && host_aff, "NULL argument");
  GX_REQUIRE(c_n >= 1 && c_lo >= 0 && c_lo + c_n <= q_num, "crop window [%lld, %lld) outside q axis of %lld",
             (long long)c_lo, (long long)(c_lo + c_n), (long long)q_num);
  GX_REQUIRE(bin_scale > 0.0 && z_elem != 0.0, "bin_scale / Z");
  AffArgs aff;
  for (int i = 0; i < 9; ++i) aff.a[i] = host_aff[i];
  const double c16pi2 = 16 * (M_PI * M_PI);
  const int64_t n = c_n * c_n * c_n;
  unsigned g = grid1d(n, BLK);
  if (g > 16384) g = 16384;
  k_bins_finalize<<<g, BLK, 0, as_stream(stream)>>>(bin_sum, bin_cnt, c_n, 1.0 / bin_scale, c_lo, q_num,
                                                     qa_start, qa_step, qa_stop, aff, z_elem, c16pi2, iq64,
                                                     iq32);
  return check_launch("gx_bins_finalize");
}

}  // extern "C"
