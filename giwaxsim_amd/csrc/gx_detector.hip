// gfx950 kernels of the detector stage (reference: tools/detector.py and
// tools/comparison.py detectormaker_fitting 790-870) and the small
// compatibility kernels behind the reference's per-call helpers.
#include "gx_device.h"
#include "gx_host.h"

#include <math.h>
#include <mutex>
#include <string>

using namespace gx;

// ------------------------------------------------------------------------------------
// error state (one message per host thread)
// ------------------------------------------------------------------------------------
namespace gx {
static thread_local std::string g_err;

int set_error(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
void clear_error() { g_err.clear(); }
}  // namespace gx

extern "C" const char *gx_last_error(void) { return gx::g_err.c_str(); }

namespace {

constexpr int BLK = 256;

// canonical R @ (x, y, z): row i = (R[i,0] x + R[i,1] y) + R[i,2] z, no FMA
__device__ __forceinline__ void rot3(const double *R, double &x, double &y, double &z) {
  const double xr = R[0] * x + R[1] * y + R[2] * z;
  const double yr = R[3] * x + R[4] * y + R[5] * z;
  const double zr = R[6] * x + R[7] * y + R[8] * z;
  x = xr;
  y = yr;
  z = zr;
}

__device__ __forceinline__ long long clip_idx(long long i, long long hi) { return i < 0 ? 0 : (i > hi ? hi : i); }

// Orientation-summed detector image.  One thread per pixel; the orientation
// loop runs in (psi outer, phi, theta inner) order with an FP64 accumulator,
// so the sum is deterministic (the reference's `det_ints += det_int` from
// worker threads is racy, tools/detector.py:298).
// q point of pixel p after the initial rotations (make_detector, detector.py:24-27)
__device__ __forceinline__ void det_pixel(const gx_detector &d, int64_t p, double &x0, double &y0, double &z0) {
  const int64_t row = p / d.n_pix, col = p - row * d.n_pix;
  x0 = 0.0;
  y0 = d.h_axis[col];
  z0 = d.v_axis[row];
  for (int k = 0; k < d.n_init; ++k) rot3(d.init_R + 9 * k, x0, y0, z0);
}

// A rotation by an angle of 0 has the matrix diag(1, 1, 1) with +-0 elsewhere: applied
// to finite coordinates it returns them unchanged up to the sign of a zero, which no
// floor index below depends on, so it can be skipped.  ORIENT_FLAGS orientations get
// a per-block table of which of their three rotations are such identities.
__device__ __forceinline__ bool rot_is_identity(const double *R) {
  return R[0] == 1.0 && R[4] == 1.0 && R[8] == 1.0 && R[1] == 0.0 && R[2] == 0.0 && R[3] == 0.0 &&
         R[5] == 0.0 && R[6] == 0.0 && R[7] == 0.0;
}

constexpr int ORIENT_FLAGS = 16384;  // table entries (bytes of LDS) at most

// Fills idf[o] (bit k: rotation k of orientation o is the identity) for the block;
// every thread of the block must call it.  Returns the table length.
__device__ __forceinline__ int orient_flags(const gx_detector &d, unsigned char *idf) {
  const int n = (int)min(d.n_orient, (int64_t)ORIENT_FLAGS);
  for (int o = threadIdx.x; o < n; o += blockDim.x) {
    const double *R = d.orient_R + 27 * (int64_t)o;
    idf[o] = (unsigned char)((rot_is_identity(R) ? 1 : 0) | (rot_is_identity(R + 9) ? 2 : 0) |
                             (rot_is_identity(R + 18) ? 4 : 0));
  }
  __syncthreads();
  return n;
}

__host__ __forceinline__ size_t orient_flags_lds(const gx_detector *d) {
  return (size_t)(d->n_orient < ORIENT_FLAGS ? d->n_orient : ORIENT_FLAGS);
}

// flat [cy][cx][cz] voxel of the pixel's q point under orientation o: the psi,
// phi, theta rotations (those flagged in idf as identities skipped), then the
// edge-clipped floor indices of intersect_detector
__device__ __forceinline__ int64_t det_voxel(const gx_detector &d, double x0, double y0, double z0, int64_t o,
                                             double dqinv, unsigned idf = 0) {
  const double *R = d.orient_R + 27 * o;
  double x = x0, y = y0, z = z0;
  if (!(idf & 1)) rot3(R, x, y, z);        // rotate_about_normal(psi)
  if (!(idf & 2)) rot3(R + 9, x, y, z);    // rotate_about_vertical(phi)
  if (!(idf & 4)) rot3(R + 18, x, y, z);   // rotate_about_horizontal(theta)
  const long long xi = clip_idx(to_int(floor_div(x - d.qx_min, d.dq, dqinv)), d.cx - 1);
  const long long yi = clip_idx(to_int(floor_div(y - d.qy_min, d.dq, dqinv)), d.cy - 1);
  const long long zi = clip_idx(to_int(floor_div(z - d.qz_min, d.dq, dqinv)), d.cz - 1);
  return (yi * d.cx + xi) * d.cz + zi;
}

// exact floor(a / b) with a one-multiply path: t = a * binv differs from the real
// quotient by at most |a / b| 2^-52 (< 2^-26 while |t| < 2^26), so when t is more than
// 2^-24 from an integer (t - floor(t) is exact for t >= 0 and within 2^-53 otherwise),
// floor(t) is the exact floor; floor_div decides the rest.
__device__ __forceinline__ double floor_div_q(double a, double b, double binv) {
  const double t = a * binv;
  const double q = floor(t);
  const double f = t - q;
  if (__builtin_expect(f > 0x1p-24 && f < 1.0 - 0x1p-24 && fabs(t) < 0x1p26, 1)) return q;
  return floor_div(a, b, binv);
}

// clip of an integral floor result to [0, hi], as clip_idx of its integer
__device__ __forceinline__ int clip_q(double q, double hi) { return (int)fmin(fmax(q, 0.0), hi); }

__device__ __forceinline__ unsigned long long dbits(double v) { return (unsigned long long)__double_as_longlong(v); }

// rot_is_identity on bit patterns (+-0 off the diagonal), for uniform matrices
__device__ __forceinline__ bool rot_is_identity_u(const double *R) {
  constexpr unsigned long long one = 0x3FF0000000000000ull;
  const unsigned long long off = dbits(R[1]) | dbits(R[2]) | dbits(R[3]) | dbits(R[5]) | dbits(R[6]) | dbits(R[7]);
  return dbits(R[0]) == one && dbits(R[4]) == one && dbits(R[8]) == one && (off << 1) == 0;
}

__device__ __forceinline__ bool same_matrix_u(const double *A, const double *B) {
  unsigned long long x = 0;
#pragma unroll
  for (int e = 0; e < 9; ++e) x |= dbits(A[e]) ^ dbits(B[e]);
  return x == 0;
}

// One thread per pixel, orientations in (psi outer, phi, theta inner) order with an FP64
// accumulator; identity rotations from the block's LDS flag table.  I32 selects 32-bit
// voxel indices (cy cx cz < 2^31) and floor_div_q's one-multiply path.
template <typename T, bool I32>
__global__ __launch_bounds__(BLK) void k_detector(gx_detector d) {
  extern __shared__ unsigned char idf[];
  const int nflag = orient_flags(d, idf);
  const T *__restrict__ iq = reinterpret_cast<const T *>(d.iq);
  // the block's pixels: a 16 x 16 tile, each wave 16 columns x 4 rows of it (fewer distinct
  // lines per gather than 64 pixels of one row: 2048^2 x 721 6.3 -> 6.1 ms, 512^2 unchanged)
  const int64_t ntc = (d.n_pix + 15) / 16;
  const int64_t trow = blockIdx.x / ntc, tcol = blockIdx.x - trow * ntc;
  const int lane = threadIdx.x & (GX_WAVE - 1), wv = threadIdx.x / GX_WAVE;
  const int64_t prow = trow * 16 + wv * 4 + lane / 16, pcol = tcol * 16 + lane % 16;
  if (prow >= d.n_pix || pcol >= d.n_pix) return;
  const int64_t p = prow * d.n_pix + pcol;
  double x0, y0, z0;
  det_pixel(d, p, x0, y0, z0);
  const double dqinv = 1.0 / d.dq;
  const double hx = (double)(d.cx - 1), hy = (double)(d.cy - 1), hz = (double)(d.cz - 1);
  double acc = 0.0;
  for (int64_t o = 0; o < d.n_orient; ++o) {
    T v;
    if (I32) {
      const unsigned f = o < nflag ? idf[o] : 0u;
      const double *R = d.orient_R + 27 * o;
      double x = x0, y = y0, z = z0;
      if (!(f & 1)) rot3(R, x, y, z);        // rotate_about_normal(psi)
      if (!(f & 2)) rot3(R + 9, x, y, z);    // rotate_about_vertical(phi)
      if (!(f & 4)) rot3(R + 18, x, y, z);   // rotate_about_horizontal(theta)
      const int xi = clip_q(floor_div_q(x - d.qx_min, d.dq, dqinv), hx);
      const int yi = clip_q(floor_div_q(y - d.qy_min, d.dq, dqinv), hy);
      const int zi = clip_q(floor_div_q(z - d.qz_min, d.dq, dqinv), hz);
      v = iq[(yi * (int)d.cx + xi) * (int)d.cz + zi];
    } else {
      v = iq[det_voxel(d, x0, y0, z0, o, dqinv, o < nflag ? idf[o] : 0u)];
    }
    acc = acc + (double)v * d.orient_w[o];
  }
  d.det[p] = acc;
}

// The same sum for detectors with few pixels and many orientations (a fit's 250^2 x 16k):
// one pixel per lane of 64, COOP_NW waves per block.  The waves gather chunks of
// COOP_K orientations each into one of two LDS buffers (v * w, the same product); after a
// barrier wave 0 adds the chunk in orientation order while the other waves fill the other
// buffer with the next chunk (one barrier per chunk), so the result is bitwise the
// one-thread-per-pixel sum while COOP_NW times more gathers are in flight per pixel
// (4 orientations per wave and chunk, double-buffered: 4.62 -> 4.48 ms per image at the
// fit geometry; 8 per wave single-buffered was the previous form, 8 double-buffered 6.5 ms).  An orientation is the same
// for the whole wave, so its matrices and weight are scalar loads and their tests scalar
// compares of bit patterns: rotations that are identities are skipped, and the
// psi-rotated point is kept from the wave's previous orientation when that one has the
// bitwise same psi matrix (psi is the outer loop of the orientation product).  Indices
// are 32-bit (the launch checks cy cx cz < 2^31) and the floor division takes
// floor_div_q's one-multiply path.
constexpr int COOP_NW = 8;                  // waves per block
constexpr int COOP_K = 4;                   // orientations per wave and chunk
constexpr int COOP_MIN_ORIENT = 64;         // fewer orientations: the per-pixel kernel
constexpr int64_t COOP_MAX_PIX = 1 << 17;   // above this the per-pixel kernel has the occupancy
// A wave's 64 pixels are a 16 x 4 tile: at the fit geometry (250^2 px, 16k orientations) it read
// fewer distinct cache lines than a 64-pixel row (4.8 vs 5.1 ms), 8 x 8 tied the row, and taller
// tiles were slower (4 x 16: 5.4 ms, a 64-pixel column 12.3 ms).
constexpr int COOP_TW = 16, COOP_TH = 4;

template <typename T>
__global__ __launch_bounds__(COOP_NW *GX_WAVE) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_detector_coop(gx_detector d) {
  extern __shared__ __align__(16) unsigned char coop_smem[];
  double *prod = reinterpret_cast<double *>(coop_smem);  // [2][COOP_NW * COOP_K][64]
  const double *__restrict__ oR = d.orient_R;
  const double *__restrict__ ow = d.orient_w;
  const T *__restrict__ iq = reinterpret_cast<const T *>(d.iq);
  double *__restrict__ det = d.det;
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / GX_WAVE);
  // the block's 64 pixels: a tile of COOP_TW columns x COOP_TH rows of the detector
  const int64_t ntc = (d.n_pix + COOP_TW - 1) / COOP_TW;
  const int64_t tb = xcd_tile(blockIdx.x, gridDim.x);  // each XCD a contiguous band of tiles
  const int64_t trow = tb / ntc, tcol = tb - trow * ntc;
  const int64_t prow = trow * COOP_TH + lane / COOP_TW, pcol = tcol * COOP_TW + lane % COOP_TW;
  const bool valid = prow < d.n_pix && pcol < d.n_pix;
  const int64_t p = valid ? prow * d.n_pix + pcol : 0;
  double x0 = 0.0, y0 = 0.0, z0 = 0.0;
  det_pixel(d, valid ? p : 0, x0, y0, z0);
  const double dqinv = 1.0 / d.dq;
  const double hx = (double)(d.cx - 1), hy = (double)(d.cy - 1), hz = (double)(d.cz - 1);
  const int cx = (int)d.cx, cz = (int)d.cz;
  const int64_t n_orient = d.n_orient;
  constexpr int CH = COOP_NW * COOP_K;
  double acc = 0.0;
  double xs = 0.0, ys = 0.0, zs = 0.0;  // pixel after the psi rotation of the previous orientation
  int64_t chunk = 0;
  for (int64_t o0 = 0; o0 < n_orient; o0 += CH, ++chunk) {
    double *pb = prod + (chunk & 1) * (CH * GX_WAVE);
#pragma unroll 2  // two orientations' gathers in flight per wave
    for (int k = 0; k < COOP_K; ++k) {
      const int j = wave * COOP_K + k;
      const int64_t o = o0 + j;
      double pr = 0.0;
      if (o < n_orient) {
        const double *R = oR + 27 * o;
        if (k == 0 || !same_matrix_u(R, R - 27)) {  // rotate_about_normal(psi)
          xs = x0;
          ys = y0;
          zs = z0;
          if (!rot_is_identity_u(R)) rot3(R, xs, ys, zs);
        }
        double x = xs, y = ys, z = zs;
        if (!rot_is_identity_u(R + 9)) rot3(R + 9, x, y, z);    // rotate_about_vertical(phi)
        if (!rot_is_identity_u(R + 18)) rot3(R + 18, x, y, z);  // rotate_about_horizontal(theta)
        const int xi = clip_q(floor_div_q(x - d.qx_min, d.dq, dqinv), hx);
        const int yi = clip_q(floor_div_q(y - d.qy_min, d.dq, dqinv), hy);
        const int zi = clip_q(floor_div_q(z - d.qz_min, d.dq, dqinv), hz);
        if (valid) pr = (double)iq[(yi * cx + xi) * cz + zi] * ow[o];
      }
      pb[j * GX_WAVE + lane] = pr;
    }
    __syncthreads();
    if (wave == 0) {
      const int nj = (int)min((int64_t)CH, n_orient - o0);
      for (int j = 0; j < nj; ++j) acc = acc + pb[j * GX_WAVE + lane];
    }
    // no second barrier: the next chunk fills the other buffer, and this one is refilled
    // only after the next barrier, which wave 0 reaches after summing it
  }
  if (wave == 0 && valid) det[p] = acc;
}

// ------------------------------------------------------------------------------------
// adjoint of the detector stage
// ------------------------------------------------------------------------------------
// t = g_out * d(finalize)/dv elementwise: the floor replaces v (NaN or <= 0) by a
// constant, else v * 1e-6; v is the forward's mirrored value, recomputed bitwise.
// Each block also writes sum |t| over its pixels (fixed order) to part[blockIdx.x].
__global__ __launch_bounds__(BLK) void k_det_fin_adj(const double *__restrict__ q, const double *__restrict__ g,
                                                     int64_t nr, int64_t nc, int64_t flags, double *__restrict__ t,
                                                     double *__restrict__ part) {
  __shared__ double sh[BLK / GX_WAVE];
  const int64_t p = blockIdx.x * (int64_t)BLK + threadIdx.x;
  double tv = 0.0;
  if (p < nr * nc) {
    if (!(flags & GX_FIN_FLOOR_SCALE)) {
      tv = g[p];
    } else {
      const int64_t r = p / nc, c = p - r * nc;
      const int64_t rr = nr - 1 - r, cc = nc - 1 - c;
      double v = q[p];
      if (flags & GX_FIN_MIRROR) {
        v = ((q[r * nc + c] + q[r * nc + cc]) + q[rr * nc + c]) + q[rr * nc + cc];
        if ((nr & 1) && r == nr / 2) v = (q[r * nc + c] + q[r * nc + cc]) * 2.0;
        if ((nc & 1) && c == nc / 2) v = (q[r * nc + c] + q[rr * nc + c]) * 2.0;
        if ((nr & 1) && (nc & 1) && r == nr / 2 && c == nc / 2) v = q[r * nc + c] * 4.0;
      }
      tv = (v == v && v > 0.0) ? g[p] * 1e-6 : 0.0;
    }
    t[p] = tv;
  }
  double a = fabs(tv);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, GX_WAVE);
  if ((threadIdx.x & (GX_WAVE - 1)) == 0) sh[threadIdx.x / GX_WAVE] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < BLK / GX_WAVE; ++i) s += sh[i];
    part[blockIdx.x] = s;
  }
}

// One block: fx_scale = 2^(61 - floor(log2 B)) with B = mult * sum(part) * sum_o |w_o|,
// a bound on every |dL/d iq| and on every partial sum of the scatter below (mult
// = 4 covers the mirror, whose outputs add at most four inputs).
__global__ __launch_bounds__(BLK) void k_det_scale(const double *__restrict__ part, int64_t n_part, double mult,
                                                   const double *w, int64_t n_orient, double *fx_scale) {
  __shared__ double sh[2][BLK / GX_WAVE];
  double a = 0.0, b = 0.0;
  for (int64_t i = threadIdx.x; i < n_part; i += BLK) a += part[i];
  for (int64_t o = threadIdx.x; o < n_orient; o += BLK) b += fabs(w[o]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, GX_WAVE);
    b += __shfl_xor(b, off, GX_WAVE);
  }
  const int lane = threadIdx.x & (GX_WAVE - 1), wv = threadIdx.x / GX_WAVE;
  if (lane == 0) {
    sh[0][wv] = a;
    sh[1][wv] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double A = 0.0, B = 0.0;
    for (int i = 0; i < BLK / GX_WAVE; ++i) {
      A += sh[0][i];
      B += sh[1][i];
    }
    const double bound = ((mult * A) * B) * (1.0 + 1e-9);
    double sc = 1.0;
    if (bound > 0.0 && bound < INFINITY) sc = ldexp(1.0, 62 - 1 - ilogb(bound));
    fx_scale[0] = sc;
  }
}

// dL/d iq: the transpose of the orientation gather, as exact fixed-point sums.
// One lane per pixel; neighbouring pixels of a row mostly land in the same voxel,
// so each wave first sums runs of equal voxel index (segmented scan over lanes)
// and only the last lane of a run issues the atomic.  Integer sums are exact, so
// the result does not depend on the grouping.
__global__ __launch_bounds__(BLK) void k_det_adj(gx_detector d, const double *__restrict__ g, const double *fx_scale,
                                                 unsigned long long *g_fx, int64_t n_rep) {
  extern __shared__ unsigned char idf[];
  const int nflag = orient_flags(d, idf);
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int64_t p = blockIdx.x * (int64_t)BLK + threadIdx.x;
  const bool live = p < d.n_pix * d.n_pix;
  const double gp = live ? g[p] : 0.0;
  if (__all(gp == 0.0)) return;  // wave-uniform
  const double sc = fx_scale[0];
  // blocks spread their atomics over n_rep copies of the grid (summed exactly later)
  unsigned long long *dst = g_fx + (int64_t)(blockIdx.x % n_rep) * (d.cy * d.cx * d.cz);
  double x0 = 0.0, y0 = 0.0, z0 = 0.0;
  if (live) det_pixel(d, p, x0, y0, z0);
  const double dqinv = 1.0 / d.dq;
  for (int64_t o = 0; o < d.n_orient; ++o) {
    int key = -1 - lane;  // unique: never merges with a neighbour
    long long val = 0;
    if (live) {
      key = (int)det_voxel(d, x0, y0, z0, o, dqinv, o < nflag ? idf[o] : 0u);
      val = llrint((gp * d.orient_w[o]) * sc);
    }
    const int up = __shfl_up(key, 1, GX_WAVE);
    int f = (lane == 0 || up != key) ? 1 : 0;
#pragma unroll
    for (int s = 1; s < GX_WAVE; s <<= 1) {
      const long long vv = __shfl_up(val, s, GX_WAVE);
      const int ff = __shfl_up(f, s, GX_WAVE);
      if (lane >= s) {
        if (!f) val += vv;
        f |= ff;
      }
    }
    const int dn = __shfl_down(key, 1, GX_WAVE);
    const bool tail = (lane == GX_WAVE - 1) || dn != key;
    if (tail && key >= 0 && val != 0) atomicAdd(&dst[key], (unsigned long long)val);
  }
}

// ------------------------------------------------------------------------------------
// transpose map of the orientation gather (fixed detector geometry), stored as runs.
// One lane per pixel.  Neighbouring pixels of a row mostly read the same voxel (a
// detector pixel is several times finer than a voxel), so per orientation each wave
// splits its lanes into runs of equal voxel; a run (consecutive pixels p .. p + len - 1,
// len <= 64, one orientation o, one voxel v) is one entry of the map: (o << b) | p, len
// and v.  Only a run's first lane issues the atomic.  The order of a voxel's runs
// depends on the timing of the atomics; the adjoint sums exact integers, so its result
// does not.
// ------------------------------------------------------------------------------------

// heads of the runs of equal key across the wave (bit l: lane l starts a run; lane 0
// always does).  Lanes without a voxel carry a key of their own (-1 - lane).
__device__ __forceinline__ unsigned long long run_heads(int key, int lane) {
  const int up = __shfl_up(key, 1, GX_WAVE);
  return __ballot(lane == 0 || up != key);
}

// length of the run that starts at `lane` (a head)
__device__ __forceinline__ int run_length(unsigned long long heads, int lane) {
  const unsigned long long above = lane == GX_WAVE - 1 ? 0ull : heads >> (lane + 1);
  return above ? __builtin_ctzll(above) + 1 : GX_WAVE - lane;
}

// runs per voxel
__global__ __launch_bounds__(BLK) void k_det_map_count(gx_detector d, int32_t *cnt) {
  extern __shared__ unsigned char idf[];
  const int nflag = orient_flags(d, idf);
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int64_t p = blockIdx.x * (int64_t)BLK + threadIdx.x;
  const bool live = p < d.n_pix * d.n_pix;
  if (__all(!live)) return;  // wave-uniform: the lane exchanges below need the whole wave
  double x0 = 0.0, y0 = 0.0, z0 = 0.0;
  if (live) det_pixel(d, p, x0, y0, z0);
  const double dqinv = 1.0 / d.dq;
  for (int64_t o = 0; o < d.n_orient; ++o) {
    const int key = live ? (int)det_voxel(d, x0, y0, z0, o, dqinv, o < nflag ? idf[o] : 0u) : -1 - lane;
    const unsigned long long heads = run_heads(key, lane);
    if (live && ((heads >> lane) & 1ull)) atomicAdd(&cnt[key], 1);
  }
}

// exclusive scan of int32 counts into int64 offsets, SCAN_B elements per block
constexpr int SCAN_B = 2048;
__global__ __launch_bounds__(BLK) void k_scan_local(const int32_t *cnt, int64_t n, int64_t *off, int64_t *bsum) {
  __shared__ int64_t sh[BLK];
  const int64_t base = (int64_t)blockIdx.x * SCAN_B;
  constexpr int PER = SCAN_B / BLK;
  int64_t v[PER];
  int64_t tot = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int64_t i = base + threadIdx.x * PER + k;
    v[k] = i < n ? cnt[i] : 0;
    tot += v[k];
  }
  sh[threadIdx.x] = tot;
  __syncthreads();
  for (int s = 1; s < BLK; s <<= 1) {
    const int64_t t = threadIdx.x >= s ? sh[threadIdx.x - s] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int64_t run = sh[threadIdx.x] - tot;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int64_t i = base + threadIdx.x * PER + k;
    if (i < n) off[i] = run;
    run += v[k];
  }
  if (threadIdx.x == BLK - 1) bsum[blockIdx.x] = sh[BLK - 1];
}

__global__ void k_scan_blocks(int64_t *bsum, int64_t nb) {
  if (threadIdx.x != 0) return;
  int64_t run = 0;
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t t = bsum[b];
    bsum[b] = run;
    run += t;
  }
  bsum[nb] = run;
}

__global__ __launch_bounds__(BLK) void k_scan_add(int64_t *off, int64_t n, const int64_t *bsum, int64_t nb,
                                                  int32_t *cnt) {
  const int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (i < n) {
    off[i] += bsum[i / SCAN_B];
    cnt[i] = 0;  // becomes the fill cursor
  }
  if (i == 0) off[n] = bsum[nb];
}

// the same runs as k_det_map_count, each written at its voxel's next slot (positions
// at or beyond `cap` are not written)
__global__ __launch_bounds__(BLK) void k_det_map_fill(gx_detector d, const int64_t *off, int32_t *cur, uint32_t *runs,
                                                      int32_t *run_voxel, uint8_t *run_len, int64_t cap, int pbits) {
  extern __shared__ unsigned char idf[];
  const int nflag = orient_flags(d, idf);
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int64_t p = blockIdx.x * (int64_t)BLK + threadIdx.x;
  const bool live = p < d.n_pix * d.n_pix;
  if (__all(!live)) return;  // wave-uniform
  double x0 = 0.0, y0 = 0.0, z0 = 0.0;
  if (live) det_pixel(d, p, x0, y0, z0);
  const double dqinv = 1.0 / d.dq;
  for (int64_t o = 0; o < d.n_orient; ++o) {
    const int key = live ? (int)det_voxel(d, x0, y0, z0, o, dqinv, o < nflag ? idf[o] : 0u) : -1 - lane;
    const unsigned long long heads = run_heads(key, lane);
    if (live && ((heads >> lane) & 1ull)) {
      const int64_t pos = off[key] + atomicAdd(&cur[key], 1);
      if (pos < cap) {
        runs[pos] = ((uint32_t)o << pbits) | (uint32_t)p;
        run_voxel[pos] = key;
        run_len[pos] = (uint8_t)run_length(heads, lane);
      }
    }
  }
}

// per-pixel fixed-point terms g[p] w0 (w0 = the first orientation's weight) as
// inclusive prefix sums within each aligned segment of 64 pixels (the segments the map's
// runs are cut in): a run p .. p + len - 1 at an orientation of weight w0 sums to
// S[p + len - 1] - S[p - 1] (0 for the first pixel of a segment), the same integers the
// per-pixel loop adds.  The scale bounds every partial sum below 2^61.
__global__ __launch_bounds__(BLK) void k_det_adj_prefix(const double *__restrict__ g, const double *orient_w,
                                                        const double *fx_scale, int64_t npx, long long *S) {
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const int64_t p = blockIdx.x * (int64_t)BLK + threadIdx.x;
  const double sc = fx_scale[0], w0 = orient_w[0];
  long long t = 0;
  if (p < npx) t = llrint((g[p] * w0) * sc);
#pragma unroll
  for (int o = 1; o < GX_WAVE; o <<= 1) {
    const long long u = __shfl_up(t, o, GX_WAVE);
    if (lane >= o) t += u;
  }
  if (p < npx) S[p] = t;
}

// dL/d iq per voxel from the transpose map, balanced over runs: a wave takes
// MAP_UNROLL chunks of 64 consecutive runs of the voxel-sorted map (every load of the
// chunks is issued before the first use); each lane sums its run's fixed-point terms
// g[p] w[o] (one per pixel, rounded as the scatter path rounds them: from the prefix
// table when w[o] is the first orientation's weight, the usual case of uniform weights,
// else pixel by pixel), the wave sums runs of equal voxel with a segmented scan across
// the lanes, and the last lane of each segment adds its total to g_fx[voxel] with an
// integer atomic.  Integer sums are exact, so the result does not depend on the order.
// g_fx starts at zero.
constexpr int MAP_UNROLL = 4;

__global__ __launch_bounds__(BLK) void k_det_adj_map(gx_detector d, const double *__restrict__ g,
                                                     const long long *__restrict__ S, const double *fx_scale,
                                                     const int32_t *__restrict__ run_voxel,
                                                     const uint32_t *__restrict__ runs,
                                                     const uint8_t *__restrict__ run_len, int64_t n_runs,
                                                     unsigned long long *g_fx, int pbits) {
  constexpr int SPAN = MAP_UNROLL * GX_WAVE;  // runs per wave and step
  const int lane = threadIdx.x & (GX_WAVE - 1);
  const double sc = fx_scale[0], w0 = d.orient_w[0];
  const uint32_t pmask = (pbits >= 32) ? 0xffffffffu : ((1u << pbits) - 1u);
  const int64_t wave0 = (blockIdx.x * (int64_t)BLK + threadIdx.x) / GX_WAVE;
  const int64_t n_waves = (int64_t)gridDim.x * (BLK / GX_WAVE);
  for (int64_t base = wave0 * SPAN; base < n_runs; base += n_waves * SPAN) {
    int v[MAP_UNROLL], len[MAP_UNROLL];
    uint32_t id[MAP_UNROLL];
#pragma unroll
    for (int u = 0; u < MAP_UNROLL; ++u) {
      const int64_t k = base + u * GX_WAVE + lane;
      v[u] = -1;
      id[u] = 0u;
      len[u] = 0;
      if (k < n_runs) {
        v[u] = run_voxel[k];
        id[u] = runs[k];
        len[u] = run_len[k];
      }
    }
#pragma unroll
    for (int u = 0; u < MAP_UNROLL; ++u) {
      long long t = 0;
      if (v[u] >= 0) {
        const uint32_t p = id[u] & pmask;
        const double w = d.orient_w[id[u] >> pbits];
        if (w == w0) {
          long long lo = 0;
          if (p & (GX_WAVE - 1)) lo = S[p - 1];
          t = S[p + len[u] - 1] - lo;
        } else {
          for (int j = 0; j < len[u]; ++j) t += llrint((g[p + j] * w) * sc);
        }
      }
      // inclusive scan within segments of equal voxel (contiguous: the map is voxel-sorted)
#pragma unroll
      for (int o = 1; o < GX_WAVE; o <<= 1) {
        const long long tu = __shfl_up(t, o, GX_WAVE);
        const int vu = __shfl_up(v[u], o, GX_WAVE);
        if (lane >= o && vu == v[u]) t += tu;
      }
      const int vn = __shfl_down(v[u], 1, GX_WAVE);
      if (v[u] >= 0 && (lane == GX_WAVE - 1 || vn != v[u])) atomicAdd(&g_fx[v[u]], (unsigned long long)t);
    }
  }
}

__global__ __launch_bounds__(BLK) void k_fill_zero_i32(int32_t *p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) p[i] = 0;
}

__global__ __launch_bounds__(BLK) void k_zero_u64(unsigned long long *p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) p[i] = 0;
}

// mirror_vertical_horizontal + NaN/<=0 floor + 1e-6 scale
__global__ __launch_bounds__(BLK) void k_detector_finalize(const double *__restrict__ q, int64_t nr, int64_t nc,
                                                           int64_t flags, double *__restrict__ out) {
  const int64_t p = blockIdx.x * (int64_t)BLK + threadIdx.x;
  if (p >= nr * nc) return;
  const int64_t r = p / nc, c = p - r * nc;
  const int64_t rr = nr - 1 - r, cc = nc - 1 - c;
  double v = q[p];
  if (flags & GX_FIN_MIRROR) {
    v = ((q[r * nc + c] + q[r * nc + cc]) + q[rr * nc + c]) + q[rr * nc + cc];
    if ((nr & 1) && r == nr / 2) v = (q[r * nc + c] + q[r * nc + cc]) * 2.0;
    if ((nc & 1) && c == nc / 2) v = (q[r * nc + c] + q[rr * nc + c]) * 2.0;
    if ((nr & 1) && (nc & 1) && r == nr / 2 && c == nc / 2) v = q[r * nc + c] * 4.0;
  }
  if (flags & GX_FIN_FLOOR_SCALE) {
    if (v != v) v = 1e-6;
    if (v <= 0.0) v = 1e-6;
    v = v * 1e-6;
  }
  out[p] = v;
}

struct Mat9 {
  double m[9];
};

__global__ __launch_bounds__(BLK) void k_rotate_points(Mat9 R, const double *x, const double *y, const double *z,
                                                       int64_t n, double *ox, double *oy, double *oz) {
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    double a = x[i], b = y[i], c = z[i];
    rot3(R.m, a, b, c);
    ox[i] = a;
    oy[i] = b;
    oz[i] = c;
  }
}

__global__ __launch_bounds__(BLK) void k_intersect(const double *iq, int64_t cy, int64_t cx, int64_t cz,
                                                   double qx_min, double qy_min, double qz_min, double dq,
                                                   const double *x, const double *y, const double *z, int64_t n,
                                                   double weight, int64_t accumulate, double *out) {
  const double dqinv = 1.0 / dq;
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    const long long xi = clip_idx(to_int(floor_div(x[i] - qx_min, dq, dqinv)), cx - 1);
    const long long yi = clip_idx(to_int(floor_div(y[i] - qy_min, dq, dqinv)), cy - 1);
    const long long zi = clip_idx(to_int(floor_div(z[i] - qz_min, dq, dqinv)), cz - 1);
    const double v = iq[(yi * cx + xi) * cz + zi] * weight;
    out[i] = accumulate ? out[i] + v : v;
  }
}

struct QLim {
  double v[6];
};

// process_file2's binning (tools/voxelgrids.py:474-503) of one frame into FP64 grids,
// bit-identical to np.add.at and deterministic.  The horizontal axes are linspaces and
// the vertical one is sorted, so every bin index is monotone along its axis: the pixels
// of one bin form a rectangle (a run of rows times a run of columns).  The thread at a
// rectangle's first row and column adds its pixels to the bin in np.add.at's order
// (row-major over the masked frame), starting from the grid's current value; no other
// thread touches that bin in this launch.
struct ColBin {
  bool ok;
  long long x, y;
};
__device__ __forceinline__ ColBin col_bin(const double *hqx, const double *hqy, int64_t j, const QLim &L, double dq,
                                          double dqinv) {
  const double qx = hqx[j], qy = hqy[j];
  ColBin c;
  c.ok = qx <= L.v[1] && qx >= L.v[0] && qy <= L.v[3] && qy >= L.v[2];
  c.x = c.ok ? to_int(floor_div(qx - L.v[0], dq, dqinv)) : 0;
  c.y = c.ok ? to_int(floor_div(qy - L.v[2], dq, dqinv)) : 0;
  return c;
}
__device__ __forceinline__ bool row_bin(const double *vqz, int64_t i, const QLim &L, double dq, double dqinv,
                                        long long *z) {
  const double qz = vqz[i];
  if (!(qz <= L.v[5] && qz >= L.v[4])) return false;
  *z = to_int(floor_div(qz - L.v[4], dq, dqinv));
  return true;
}

// Axes that are not monotone (a caller's own arrays): one atomic add per pixel, in no
// particular order.
__global__ __launch_bounds__(BLK) void k_bin_frame_f64_any(const double *iq2d, int64_t nv, int64_t nh,
                                                           const double *hqx, const double *hqy, const double *vqz,
                                                           QLim L, double dq, int64_t q_num, double *gsum,
                                                           double *gcnt, int32_t *status) {
  const double dqinv = 1.0 / dq;
  for (int64_t p = blockIdx.x * (int64_t)BLK + threadIdx.x; p < nv * nh; p += (int64_t)gridDim.x * BLK) {
    const int64_t i = p / nh, j = p - i * nh;
    long long zi;
    if (!row_bin(vqz, i, L, dq, dqinv, &zi)) continue;
    const ColBin c = col_bin(hqx, hqy, j, L, dq, dqinv);
    if (!c.ok) continue;
    if (c.x < 0 || c.x >= q_num || c.y < 0 || c.y >= q_num || zi < 0 || zi >= q_num) {
      atomicAdd(&status[0], 1);
      continue;
    }
    const int64_t b = (c.y * q_num + c.x) * q_num + zi;
    atomicAdd(&gsum[b], iq2d[p]);
    atomicAdd(&gcnt[b], 1.0);
  }
}

__global__ __launch_bounds__(BLK) void k_bin_frame_f64(const double *iq2d, int64_t nv, int64_t nh, const double *hqx,
                                                       const double *hqy, const double *vqz, QLim L, double dq,
                                                       int64_t q_num, double *gsum, double *gcnt, int32_t *status) {
  const double dqinv = 1.0 / dq;
  for (int64_t p = blockIdx.x * (int64_t)BLK + threadIdx.x; p < nv * nh; p += (int64_t)gridDim.x * BLK) {
    const int64_t i = p / nh, j = p - i * nh;
    long long zi, zp;
    if (!row_bin(vqz, i, L, dq, dqinv, &zi)) continue;
    if (i > 0 && row_bin(vqz, i - 1, L, dq, dqinv, &zp) && zp == zi) continue;  // not the run's first row
    const ColBin c = col_bin(hqx, hqy, j, L, dq, dqinv);
    if (!c.ok) continue;
    if (j > 0) {
      const ColBin cp = col_bin(hqx, hqy, j - 1, L, dq, dqinv);
      if (cp.ok && cp.x == c.x && cp.y == c.y) continue;  // not the run's first column
    }
    if (c.x < 0 || c.x >= q_num || c.y < 0 || c.y >= q_num || zi < 0 || zi >= q_num) {
      atomicAdd(&status[0], 1);
      continue;
    }
    int64_t j1 = j + 1;
    while (j1 < nh) {
      const ColBin cn = col_bin(hqx, hqy, j1, L, dq, dqinv);
      if (!cn.ok || cn.x != c.x || cn.y != c.y) break;
      ++j1;
    }
    const int64_t b = (c.y * q_num + c.x) * q_num + zi;
    double acc = gsum[b];
    int64_t rows = 0;
    for (int64_t ii = i; ii < nv; ++ii) {
      long long zn;
      if (ii > i && !(row_bin(vqz, ii, L, dq, dqinv, &zn) && zn == zi)) break;
      for (int64_t jj = j; jj < j1; ++jj) acc = acc + iq2d[ii * nh + jj];
      ++rows;
    }
    gsum[b] = acc;
    gcnt[b] = gcnt[b] + (double)(rows * (j1 - j));
  }
}

struct FP {
  double v[GX_FP_STRIDE];
};

__global__ __launch_bounds__(BLK) void k_collapse_lengths(FP fp, const double *x, int64_t n, double *out) {
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK)
    out[i] = collapse_length(fp.v, x[i]);
}

struct Aff9 {
  double a[9];
};

// iq * ((f0(|q|) / Z)^2) over a [ny][nx][nz] grid (add_f0_q_3d)
__global__ __launch_bounds__(BLK) void k_apply_f0(const double *iq, int64_t ny, int64_t nx, int64_t nz, const double *qx,
                                                  const double *qy, const double *qz, Aff9 aff, double z_elem,
                                                  double c16pi2, double *out) {
  const int64_t n = ny * nx * nz;
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    const int64_t iz = i % nz, ix = (i / nz) % nx, iy = i / (nz * nx);
    const double a = qx[ix], b = qy[iy], c = qz[iz];
    const double q2 = (a * a + b * b) + c * c;
    out[i] = iq[i] * f0_weight(aff.a, q2, z_elem, c16pi2);
  }
}

// grid_sum += bin_sum / scale, grid_cnt += bin_cnt
__global__ __launch_bounds__(BLK) void k_bins_accumulate_f64(const unsigned long long *bs, const int32_t *bc, int64_t n,
                                                             double inv_scale, double *gs, double *gc) {
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    gs[i] = gs[i] + (double)bs[i] * inv_scale;
    gc[i] = gc[i] + (double)bc[i];
  }
}

// deterministic two-level FP64 reductions for the scale/offset fit
constexpr int MOM_BLOCKS = 256;

__device__ double block_sum(double v, double *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, GX_WAVE);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < BLK / GX_WAVE; ++i) t += sh[i];
  return t;
}

__global__ __launch_bounds__(BLK) void k_mom1(const double *s, const double *y, const uint8_t *mask, int64_t n,
                                              double *part) {
  __shared__ double sh[BLK / GX_WAVE];
  double c = 0, ss = 0, sy = 0;
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)MOM_BLOCKS * BLK)
    if (mask[i] == 0) {
      c += 1.0;
      ss += s[i];
      sy += y[i];
    }
  double a = block_sum(c, sh), b = block_sum(ss, sh), e = block_sum(sy, sh);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = a;
    part[3 * blockIdx.x + 1] = b;
    part[3 * blockIdx.x + 2] = e;
  }
}

__global__ void k_mom1_fin(double *part, double *res) {
  if (threadIdx.x == 0) {
    double c = 0, a = 0, b = 0;
    for (int i = 0; i < MOM_BLOCKS; ++i) {
      c += part[3 * i];
      a += part[3 * i + 1];
      b += part[3 * i + 2];
    }
    res[0] = c;
    res[1] = c > 0 ? a / c : 0.0;
    res[2] = c > 0 ? b / c : 0.0;
  }
}

__global__ __launch_bounds__(BLK) void k_mom2(const double *s, const double *y, const uint8_t *mask, int64_t n,
                                              const double *res, double *part) {
  __shared__ double sh[BLK / GX_WAVE];
  const double ms = res[1], my = res[2];
  double a = 0, b = 0;
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)MOM_BLOCKS * BLK)
    if (mask[i] == 0) {
      const double ds = s[i] - ms, dy = y[i] - my;
      a += ds * ds;
      b += ds * dy;
    }
  double ta = block_sum(a, sh), tb = block_sum(b, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ta;
    part[2 * blockIdx.x + 1] = tb;
  }
}

__global__ void k_mom2_fin(const double *part, double *res) {
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int i = 0; i < MOM_BLOCKS; ++i) {
      a += part[2 * i];
      b += part[2 * i + 1];
    }
    res[3] = a;
    res[4] = b;
    res[5] = 0.0;
  }
}

}  // namespace

extern "C" {

int gx_detector_accumulate(const gx_detector *d, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(d != nullptr, "NULL detector");
  GX_REQUIRE(d->n_pix >= 1 && d->n_pix <= 65536, "n_pix=%lld", (long long)d->n_pix);
  GX_REQUIRE(d->n_init >= 0 && d->n_init <= 3, "n_init=%lld outside [0, 3]", (long long)d->n_init);
  GX_REQUIRE(d->n_orient >= 0, "n_orient");
  GX_REQUIRE(d->cy >= 1 && d->cx >= 1 && d->cz >= 1, "empty iq grid");
  GX_REQUIRE(d->dq > 0.0, "dq must be positive");
  GX_REQUIRE(d->h_axis && d->v_axis && d->iq && d->det && (d->n_init == 0 || d->init_R) &&
                 (d->n_orient == 0 || (d->orient_R && d->orient_w)),
             "a detector buffer is NULL");
  const int64_t npx = d->n_pix * d->n_pix;
  if (npx <= COOP_MAX_PIX && d->n_orient >= COOP_MIN_ORIENT && d->cy * d->cx * d->cz < (1ll << 31)) {
    // few pixels, many orientations: COOP_NW waves gather for the same 64 pixels
    const unsigned nb = (unsigned)(((d->n_pix + COOP_TW - 1) / COOP_TW) * ((d->n_pix + COOP_TH - 1) / COOP_TH));
    const size_t lds = (size_t)2 * COOP_NW * COOP_K * GX_WAVE * sizeof(double);
    if (d->iq_f64)
      k_detector_coop<double><<<nb, COOP_NW * GX_WAVE, lds, as_stream(stream)>>>(*d);
    else
      k_detector_coop<float><<<nb, COOP_NW * GX_WAVE, lds, as_stream(stream)>>>(*d);
    return check_launch("gx_detector_accumulate");
  }
  const bool i32 = d->cy * d->cx * d->cz < (1ll << 31);
  const unsigned g = (unsigned)(((d->n_pix + 15) / 16) * ((d->n_pix + 15) / 16));
  if (d->iq_f64) {
    if (i32)
      k_detector<double, true><<<g, BLK, orient_flags_lds(d), as_stream(stream)>>>(*d);
    else
      k_detector<double, false><<<g, BLK, orient_flags_lds(d), as_stream(stream)>>>(*d);
  } else {
    if (i32)
      k_detector<float, true><<<g, BLK, orient_flags_lds(d), as_stream(stream)>>>(*d);
    else
      k_detector<float, false><<<g, BLK, orient_flags_lds(d), as_stream(stream)>>>(*d);
  }
  return check_launch("gx_detector_accumulate");
}

// bits of a pair id that hold the pixel (the orientation takes the rest)
static int map_pixel_bits(int64_t npx) {
  int b = 0;
  while ((1ll << b) < npx) ++b;
  return b;
}

static int check_map_detector(const gx_detector *d) {
  GX_REQUIRE(d != nullptr, "NULL detector");
  GX_REQUIRE(d->n_pix >= 1 && d->n_pix <= 65536 && d->n_init >= 0 && d->n_init <= 3 && d->n_orient >= 1,
             "detector sizes");
  GX_REQUIRE(d->cy >= 1 && d->cx >= 1 && d->cz >= 1 && d->dq > 0.0, "iq grid");
  GX_REQUIRE(d->h_axis && d->v_axis && (d->n_init == 0 || d->init_R) && d->orient_R && d->orient_w,
             "a detector buffer is NULL");
  GX_REQUIRE(d->cy * d->cx * d->cz < (1ll << 31), "iq grid too large for 32-bit voxel ids");
  const int64_t npx = d->n_pix * d->n_pix;
  const int pbits = map_pixel_bits(npx);
  GX_REQUIRE(pbits < 32 && d->n_orient <= (1ll << (32 - pbits)),
             "%lld orientations x %lld pixels do not pack into 32-bit run ids", (long long)d->n_orient,
             (long long)npx);
  return GX_OK;
}

int gx_detector_map_count(const gx_detector *d, int32_t *counts, int64_t *scan_work, int64_t *voxel_off,
                          gx_stream_t stream) {
  clear_error();
  const int rc = check_map_detector(d);
  if (rc != GX_OK) return rc;
  GX_REQUIRE(counts && scan_work && voxel_off, "NULL map buffer");
  const int64_t npx = d->n_pix * d->n_pix, nv = d->cy * d->cx * d->cz;
  hipStream_t st = as_stream(stream);
  const int64_t nb = (nv + SCAN_B - 1) / SCAN_B;
  k_fill_zero_i32<<<grid1d(nv, BLK) > 8192 ? 8192 : grid1d(nv, BLK), BLK, 0, st>>>(counts, nv);
  k_det_map_count<<<grid1d(npx, BLK), BLK, orient_flags_lds(d), st>>>(*d, counts);
  k_scan_local<<<(unsigned)nb, BLK, 0, st>>>(counts, nv, voxel_off, scan_work);
  k_scan_blocks<<<1, 64, 0, st>>>(scan_work, nb);
  k_scan_add<<<grid1d(nv, BLK), BLK, 0, st>>>(voxel_off, nv, scan_work, nb, counts);
  return check_launch("gx_detector_map_count");
}

int gx_detector_map_fill(const gx_detector *d, const int64_t *voxel_off, int32_t *counts, uint32_t *runs,
                         int32_t *run_voxel, uint8_t *run_len, int64_t capacity, gx_stream_t stream) {
  clear_error();
  const int rc = check_map_detector(d);
  if (rc != GX_OK) return rc;
  GX_REQUIRE(counts && voxel_off && runs && run_voxel && run_len, "NULL map buffer");
  GX_REQUIRE(capacity >= 0, "capacity=%lld", (long long)capacity);
  const int64_t npx = d->n_pix * d->n_pix;
  k_det_map_fill<<<grid1d(npx, BLK), BLK, orient_flags_lds(d), as_stream(stream)>>>(
      *d, voxel_off, counts, runs, run_voxel, run_len, capacity, map_pixel_bits(npx));
  return check_launch("gx_detector_map_fill");
}

int gx_detector_adjoint_mapped(const gx_detector *d, const double *g_out, int64_t flags, double *work,
                               const int32_t *run_voxel, const uint32_t *runs, const uint8_t *run_len, int64_t n_runs,
                               unsigned long long *g_fx, double *fx_scale, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(d != nullptr, "NULL detector");
  GX_REQUIRE(d->n_pix >= 1 && d->n_pix <= 65536 && d->n_orient >= 1, "detector sizes");
  GX_REQUIRE(d->cy >= 1 && d->cx >= 1 && d->cz >= 1, "empty iq grid");
  GX_REQUIRE(d->det && d->orient_w && g_out && work && run_voxel && runs && run_len && g_fx && fx_scale,
             "NULL adjoint buffer");
  const int64_t n = d->n_pix, npx = n * n;
  GX_REQUIRE(n_runs >= 1 && n_runs <= npx * d->n_orient, "n_runs=%lld outside [1, n_pix^2 n_orient]",
             (long long)n_runs);
  hipStream_t st = as_stream(stream);
  const unsigned nblk = grid1d(npx, BLK);
  double *t = work, *g_raw = work + npx;
  long long *S = reinterpret_cast<long long *>(work + 2 * npx);
  double *part = work + 3 * npx;
  k_det_fin_adj<<<nblk, BLK, 0, st>>>(d->det, g_out, n, n, flags, t, part);
  if (flags & GX_FIN_MIRROR)
    k_detector_finalize<<<nblk, BLK, 0, st>>>(t, n, n, GX_FIN_MIRROR, g_raw);
  else
    g_raw = t;
  k_det_scale<<<1, BLK, 0, st>>>(part, nblk, (flags & GX_FIN_MIRROR) ? 4.0 : 1.0, d->orient_w, d->n_orient,
                                 fx_scale);
  k_det_adj_prefix<<<nblk, BLK, 0, st>>>(g_raw, d->orient_w, fx_scale, npx, S);
  const int64_t nv = d->cy * d->cx * d->cz;
  const unsigned gz = grid1d(nv, BLK);
  k_zero_u64<<<gz > 8192 ? 8192 : gz, BLK, 0, st>>>(g_fx, nv);
  const unsigned gp = grid1d(n_runs, BLK * MAP_UNROLL);  // a wave per 64 MAP_UNROLL runs, grid-strided
  k_det_adj_map<<<gp > 65536 ? 65536 : gp, BLK, 0, st>>>(*d, g_raw, S, fx_scale, run_voxel, runs, run_len, n_runs,
                                                         g_fx, map_pixel_bits(npx));
  return check_launch("gx_detector_adjoint_mapped");
}

int gx_detector_adjoint(const gx_detector *d, const double *g_out, int64_t flags, double *work,
                        unsigned long long *g_fx, int64_t n_rep, double *fx_scale, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(d != nullptr, "NULL detector");
  GX_REQUIRE(d->n_pix >= 1 && d->n_pix <= 65536, "n_pix=%lld", (long long)d->n_pix);
  GX_REQUIRE(d->n_init >= 0 && d->n_init <= 3, "n_init=%lld outside [0, 3]", (long long)d->n_init);
  GX_REQUIRE(d->n_orient >= 0, "n_orient");
  GX_REQUIRE(d->cy >= 1 && d->cx >= 1 && d->cz >= 1, "empty iq grid");
  GX_REQUIRE(d->dq > 0.0, "dq must be positive");
  GX_REQUIRE(d->h_axis && d->v_axis && d->det && (d->n_init == 0 || d->init_R) &&
                 (d->n_orient == 0 || (d->orient_R && d->orient_w)),
             "a detector buffer is NULL");
  GX_REQUIRE(g_out && work && g_fx && fx_scale, "NULL adjoint buffer");
  GX_REQUIRE(n_rep >= 1 && n_rep <= 64, "n_rep=%lld outside [1, 64]", (long long)n_rep);
  hipStream_t st = as_stream(stream);
  const int64_t n = d->n_pix, npx = n * n;
  const unsigned nblk = grid1d(npx, BLK);
  double *t = work, *g_raw = work + npx, *part = work + 2 * npx;
  k_det_fin_adj<<<nblk, BLK, 0, st>>>(d->det, g_out, n, n, flags, t, part);
  if (flags & GX_FIN_MIRROR)  // the mirror sum is symmetric: its transpose is itself
    k_detector_finalize<<<nblk, BLK, 0, st>>>(t, n, n, GX_FIN_MIRROR, g_raw);
  else
    g_raw = t;
  const int64_t nv = d->cy * d->cx * d->cz;
  GX_REQUIRE(nv < 2147483647LL, "iq grid of %lld voxels exceeds int32 indexing", (long long)nv);
  unsigned gz = grid1d(nv * n_rep, BLK);
  k_zero_u64<<<gz > 4096 ? 4096 : gz, BLK, 0, st>>>(g_fx, nv * n_rep);
  k_det_scale<<<1, BLK, 0, st>>>(part, nblk, (flags & GX_FIN_MIRROR) ? 4.0 : 1.0, d->orient_w, d->n_orient,
                                 fx_scale);
  k_det_adj<<<nblk, BLK, orient_flags_lds(d), st>>>(*d, g_raw, fx_scale, g_fx, n_rep);
  return check_launch("gx_detector_adjoint");
}

int gx_detector_finalize(const double *det, int64_t n_rows, int64_t n_cols, int64_t flags, double *out,
                         gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(det && out && det != out, "det/out must be distinct non-NULL buffers");
  GX_REQUIRE(n_rows >= 1 && n_cols >= 1, "empty image");
  k_detector_finalize<<<grid1d(n_rows * n_cols, BLK), BLK, 0, as_stream(stream)>>>(det, n_rows, n_cols, flags,
                                                                                   out);
  return check_launch("gx_detector_finalize");
}

int gx_rotate_points(const double *host_R, const double *x, const double *y, const double *z, int64_t n, double *ox,
                     double *oy, double *oz, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(host_R && x && y && z && ox && oy && oz, "NULL argument");
  GX_REQUIRE(n >= 0, "n");
  if (n == 0) return GX_OK;
  Mat9 R;
  for (int i = 0; i < 9; ++i) R.m[i] = host_R[i];
  unsigned g = grid1d(n, BLK);
  if (g > 16384) g = 16384;
  k_rotate_points<<<g, BLK, 0, as_stream(stream)>>>(R, x, y, z, n, ox, oy, oz);
  return check_launch("gx_rotate_points");
}

int gx_intersect(const double *iq, int64_t cy, int64_t cx, int64_t cz, double qx_min, double qy_min, double qz_min,
                 double dq, const double *x, const double *y, const double *z, int64_t n, double weight,
                 int64_t accumulate, double *out, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(iq && x && y && z && out, "NULL argument");
  GX_REQUIRE(cy >= 1 && cx >= 1 && cz >= 1 && dq > 0.0 && n >= 0, "bad sizes");
  if (n == 0) return GX_OK;
  unsigned g = grid1d(n, BLK);
  if (g > 16384) g = 16384;
  k_intersect<<<g, BLK, 0, as_stream(stream)>>>(iq, cy, cx, cz, qx_min, qy_min, qz_min, dq, x, y, z, n, weight,
                                                 accumulate, out);
  return check_launch("gx_intersect");
}

int gx_bin_frame_f64(const double *iq2d, int64_t nv, int64_t nh, const double *hqx, const double *hqy,
                     const double *vqz, const double *host_qlim, double dq, int64_t q_num, double *grid_sum,
                     double *grid_cnt, int32_t *status, int64_t monotone_axes, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(iq2d && hqx && hqy && vqz && host_qlim && grid_sum && grid_cnt && status, "NULL argument");
  GX_REQUIRE(nv >= 0 && nh >= 0 && q_num >= 1 && dq > 0.0, "bad sizes");
  if (nv * nh == 0) return GX_OK;
  QLim L;
  for (int i = 0; i < 6; ++i) L.v[i] = host_qlim[i];
  unsigned g = grid1d(nv * nh, BLK);
  if (g > 16384) g = 16384;
  if (monotone_axes)
    k_bin_frame_f64<<<g, BLK, 0, as_stream(stream)>>>(iq2d, nv, nh, hqx, hqy, vqz, L, dq, q_num, grid_sum, grid_cnt,
                                                       status);
  else
    k_bin_frame_f64_any<<<g, BLK, 0, as_stream(stream)>>>(iq2d, nv, nh, hqx, hqy, vqz, L, dq, q_num, grid_sum,
                                                           grid_cnt, status);
  return check_launch("gx_bin_frame_f64");
}

int gx_collapse_lengths(const double *host_fp, const double *x, int64_t n, double *out, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(host_fp && x && out && n >= 0, "bad argument");
  if (n == 0) return GX_OK;
  FP fp;
  for (int i = 0; i < GX_FP_STRIDE; ++i) fp.v[i] = host_fp[i];
  k_collapse_lengths<<<grid1d(n, BLK), BLK, 0, as_stream(stream)>>>(fp, x, n, out);
  return check_launch("gx_collapse_lengths");
}

int gx_apply_f0(const double *iq, int64_t ny, int64_t nx, int64_t nz, const double *qx, const double *qy,
                const double *qz, const double *host_aff, double z_elem, double *out, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(iq && qx && qy && qz && host_aff && out, "NULL argument");
  GX_REQUIRE(ny >= 0 && nx >= 0 && nz >= 0 && z_elem != 0.0, "bad sizes");
  const int64_t n = ny * nx * nz;
  if (n == 0) return GX_OK;
  Aff9 aff;
  for (int i = 0; i < 9; ++i) aff.a[i] = host_aff[i];
  unsigned g = grid1d(n, BLK);
  if (g > 16384) g = 16384;
  k_apply_f0<<<g, BLK, 0, as_stream(stream)>>>(iq, ny, nx, nz, qx, qy, qz, aff, z_elem, 16 * (M_PI * M_PI), out);
  return check_launch("gx_apply_f0");
}

int gx_bins_accumulate_f64(const unsigned long long *bin_sum, const int32_t *bin_cnt, int64_t n, double bin_scale,
                           double *grid_sum, double *grid_cnt, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(bin_sum && bin_cnt && grid_sum && grid_cnt && n >= 0 && bin_scale > 0.0, "bad argument");
  if (n == 0) return GX_OK;
  unsigned g = grid1d(n, BLK);
  if (g > 16384) g = 16384;
  k_bins_accumulate_f64<<<g, BLK, 0, as_stream(stream)>>>(bin_sum, bin_cnt, n, 1.0 / bin_scale, grid_sum, grid_cnt);
  return check_launch("gx_bins_accumulate_f64");
}

int gx_scale_offset_moments(const double *sim, const double *exp_, const uint8_t *mask, int64_t n, double *work,
                            double *host_out, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(sim && exp_ && mask && work && host_out && n >= 0, "bad argument");
  hipStream_t st = as_stream(stream);
  double *part = work;                   // [3 * MOM_BLOCKS]
  double *res = work + 3 * MOM_BLOCKS;   // [6]
  k_mom1<<<MOM_BLOCKS, BLK, 0, st>>>(sim, exp_, mask, n, part);
  k_mom1_fin<<<1, 64, 0, st>>>(part, res);
  k_mom2<<<MOM_BLOCKS, BLK, 0, st>>>(sim, exp_, mask, n, res, part);
  k_mom2_fin<<<1, 64, 0, st>>>(part, res);
  int rc = check_launch("gx_scale_offset_moments");
  if (rc) return rc;
  GX_HIP(hipMemcpyAsync(host_out, res, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
  GX_HIP(hipStreamSynchronize(st));
  return GX_OK;
}

}  // extern "C"
