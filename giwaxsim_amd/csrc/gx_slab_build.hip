// gfx950 kernels of the slab builder of a fit loop (reference: tools/comparison.py
// slabmaker_fitting 595-671), which fit_slabsize.py calls once per evaluation:
// replicate a unit cell along a, b and c, shift to the origin, keep the atoms inside
// the centred box of the requested size, shift again.
//
//   k_slab_tile     atom o of the replicated set, in the reference's concatenation
//                   order (x copies, then y copies of that set, then z copies),
//                   with its coordinates added in the same order
//   k_crop_count / k_crop_scan / k_crop_scatter
//                   the box mask and an order-preserving compaction (per-block
//                   counts, one scan, ballot ranks within a wave, waves in order)
//   k_shift_min     subtract the kept atoms' minima (from gx_slab_bounds)
// The bounding boxes come from gx_slab_bounds (exact FP64 min / max).
#include "gx_device.h"
#include "gx_host.h"

#include <math.h>

#include <mutex>

using namespace gx;

// Block counts of gx_slab_crop_noalloc, held by the library on every device it runs on
// (until the process ends): one int64 per 2048-atom block plus the total, for up to
// GX_CROP_NOALLOC_MAX_ATOMS atoms (512 KiB).  In host code this name is not a device
// address: the entry point uses the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) int64_t scratch[GX_CROP_NOALLOC_MAX_ATOMS / 2048 + 1];

namespace {

constexpr int BLK = 256;
constexpr int CROP_B = 2048;  // atoms per block in the compaction

__global__ __launch_bounds__(BLK) void k_slab_tile(const double *__restrict__ cell, const int32_t *__restrict__ cell_elem,
                                                   int64_t n_cell, int64_t nx, int64_t ny, int64_t n_out, double a0,
                                                   double b0, double b1, double c0, double c1, double c2,
                                                   double *__restrict__ x, double *__restrict__ y,
                                                   double *__restrict__ z, int32_t *__restrict__ elem) {
  for (int64_t o = blockIdx.x * (int64_t)BLK + threadIdx.x; o < n_out; o += (int64_t)gridDim.x * BLK) {
    const int64_t i = o % n_cell, k = o / n_cell;
    const int64_t kx = k % (nx + 1), ky = (k / (nx + 1)) % (ny + 1), kz = k / ((nx + 1) * (ny + 1));
    double px = cell[3 * i], py = cell[3 * i + 1], pz = cell[3 * i + 2];
    // each expansion adds its translation vector [v0 (k), v1 (k), v2 (k)] to the set
    if (kx) {
      px = px + a0 * (double)kx;
      py = py + 0.0;
      pz = pz + 0.0;
    }
    if (ky) {
      px = px + b0 * (double)ky;
      py = py + b1 * (double)ky;
      pz = pz + 0.0;
    }
    if (kz) {
      px = px + c0 * (double)kz;
      py = py + c1 * (double)kz;
      pz = pz + c2 * (double)kz;
    }
    x[o] = px;
    y[o] = py;
    z[o] = pz;
    elem[o] = cell_elem[i];
  }
}

// lim: x, y, z minima, then [lo, hi] of the shifted coordinate on each axis
__device__ __forceinline__ bool crop_keep(double x, double y, double z, const double *lim) {
  const double sx = x - lim[0], sy = y - lim[1], sz = z - lim[2];
  return (sx >= lim[3]) & (sx <= lim[4]) & (sy >= lim[5]) & (sy <= lim[6]) & (sz >= lim[7]) & (sz <= lim[8]);
}

struct Lim9 {
  double v[9];
};

__global__ __launch_bounds__(BLK) void k_crop_count(const double *__restrict__ x, const double *__restrict__ y,
                                                    const double *__restrict__ z, int64_t n, Lim9 lim,
                                                    int64_t *__restrict__ bcount) {
  __shared__ int part[BLK / GX_WAVE];
  const int64_t a0 = (int64_t)blockIdx.x * CROP_B;
  int c = 0;
  for (int64_t a = a0 + threadIdx.x; a < min(n, a0 + CROP_B); a += BLK) c += crop_keep(x[a], y[a], z[a], lim.v);
  for (int o = GX_WAVE / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, GX_WAVE);
  if ((threadIdx.x & (GX_WAVE - 1)) == 0) part[threadIdx.x / GX_WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int w = 0; w < BLK / GX_WAVE; ++w) t += part[w];
    bcount[blockIdx.x] = t;
  }
}

// exclusive scan of the nb block counts in place; bcount[nb] = total (one block)
__global__ __launch_bounds__(1024) void k_crop_scan(int64_t *bcount, int64_t nb) {
  __shared__ int64_t sh[1024];
  const int T = blockDim.x, t = threadIdx.x;
  const int64_t chunk = (nb + T - 1) / T;
  const int64_t b0 = t * chunk, b1 = min(nb, b0 + chunk);
  int64_t s = 0;
  for (int64_t i = b0; i < b1; ++i) s += bcount[i];
  sh[t] = s;
  __syncthreads();
  for (int off = 1; off < T; off <<= 1) {
    const int64_t v = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += v;
    __syncthreads();
  }
  int64_t run = sh[t] - s;
  for (int64_t i = b0; i < b1; ++i) {
    const int64_t c = bcount[i];
    bcount[i] = run;
    run += c;
  }
  if (t == T - 1) bcount[nb] = sh[T - 1];
}

__global__ __launch_bounds__(BLK) void k_crop_scatter(const double *__restrict__ x, const double *__restrict__ y,
                                                      const double *__restrict__ z, const int32_t *__restrict__ elem,
                                                      int64_t n, Lim9 lim, const int64_t *__restrict__ boff,
                                                      double *__restrict__ ox, double *__restrict__ oy,
                                                      double *__restrict__ oz, int32_t *__restrict__ oelem) {
  __shared__ int wsum[BLK / GX_WAVE];
  const int lane = threadIdx.x & (GX_WAVE - 1), w = threadIdx.x / GX_WAVE;
  const int64_t a0 = (int64_t)blockIdx.x * CROP_B;
  int64_t base = boff[blockIdx.x];
  for (int64_t r0 = a0; r0 < min(n, a0 + CROP_B); r0 += BLK) {  // rounds of BLK atoms, in order
    const int64_t a = r0 + threadIdx.x;
    const bool keep = a < n && crop_keep(x[a], y[a], z[a], lim.v);
    const unsigned long long m = __ballot(keep);
    const int rank = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int before = 0, total = 0;
    for (int k = 0; k < BLK / GX_WAVE; ++k) {
      if (k < w) before += wsum[k];
      total += wsum[k];
    }
    if (keep) {
      const int64_t p = base + before + rank;
      ox[p] = x[a] - lim.v[0];
      oy[p] = y[a] - lim.v[1];
      oz[p] = z[a] - lim.v[2];
      oelem[p] = elem[a];
    }
    base += total;
    __syncthreads();
  }
}

// coordinates -= the minima in mm (gx_slab_bounds layout: xmin, xmax, ymin, ymax, zmin, zmax)
__global__ __launch_bounds__(BLK) void k_shift_min(double *__restrict__ x, double *__restrict__ y,
                                                   double *__restrict__ z, int64_t n, const double *__restrict__ mm) {
  const double mx = mm[0], my = mm[2], mz = mm[4];
  for (int64_t a = blockIdx.x * (int64_t)BLK + threadIdx.x; a < n; a += (int64_t)gridDim.x * BLK) {
    x[a] = x[a] - mx;
    y[a] = y[a] - my;
    z[a] = z[a] - mz;
  }
}

unsigned grid_cap(int64_t n) {
  const unsigned g = grid1d(n, BLK);
  return g > 16384 ? 16384 : (g ? g : 1);
}

}  // namespace

extern "C" {

int gx_slab_tile(const double *cell_xyz, const int32_t *cell_elem, int64_t n_cell, int64_t num_x, int64_t num_y,
                 int64_t num_z, const double *host_vec, double *x, double *y, double *z, int32_t *elem,
                 gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(n_cell >= 1 && num_x >= 0 && num_y >= 0 && num_z >= 0, "cell of %lld atoms, copies %lld %lld %lld",
             (long long)n_cell, (long long)num_x, (long long)num_y, (long long)num_z);
  GX_REQUIRE(cell_xyz && cell_elem && host_vec && x && y && z && elem, "NULL buffer");
  const int64_t n_out = n_cell * (num_x + 1) * (num_y + 1) * (num_z + 1);
  GX_REQUIRE(n_out < (1ll << 31), "replicated slab of %lld atoms exceeds 2^31", (long long)n_out);
  const double *v = host_vec;
  k_slab_tile<<<grid_cap(n_out), BLK, 0, as_stream(stream)>>>(cell_xyz, cell_elem, n_cell, num_x, num_y, n_out, v[0],
                                                               v[1], v[2], v[3], v[4], v[5], x, y, z, elem);
  return check_launch("gx_slab_tile");
}

int gx_slab_crop(const double *x, const double *y, const double *z, const int32_t *elem, int64_t n,
                 const double *host_lim, double *ox, double *oy, double *oz, int32_t *oelem, int64_t *work,
                 int64_t *host_count, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(n >= 1 && n < (1ll << 31), "n=%lld", (long long)n);
  GX_REQUIRE(x && y && z && elem && host_lim && ox && oy && oz && oelem && work && host_count, "NULL buffer");
  hipStream_t st = as_stream(stream);
  Lim9 lim;
  for (int k = 0; k < 9; ++k) lim.v[k] = host_lim[k];
  const int64_t nb = (n + CROP_B - 1) / CROP_B;
  k_crop_count<<<(unsigned)nb, BLK, 0, st>>>(x, y, z, n, lim, work);
  k_crop_scan<<<1, 1024, 0, st>>>(work, nb);
  k_crop_scatter<<<(unsigned)nb, BLK, 0, st>>>(x, y, z, elem, n, lim, work, ox, oy, oz, oelem);
  int rc = check_launch("gx_slab_crop");
  if (rc) return rc;
  GX_HIP(hipMemcpyAsync(host_count, work + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  GX_HIP(hipStreamSynchronize(st));
  return GX_OK;
}

int gx_slab_crop_noalloc(const double *x, const double *y, const double *z, const int32_t *elem, int64_t n,
                         const double *host_lim, double *ox, double *oy, double *oz, int32_t *oelem,
                         int64_t *host_count, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(n >= 1 && n <= GX_CROP_NOALLOC_MAX_ATOMS,
             "gx_slab_crop_noalloc: n=%lld atoms outside [1, %lld] (the library's block-count array holds %lld "
             "blocks of %d atoms); use gx_slab_crop with a caller-owned work buffer",
             (long long)n, (long long)GX_CROP_NOALLOC_MAX_ATOMS, (long long)(GX_CROP_NOALLOC_MAX_ATOMS / CROP_B),
             CROP_B);
  GX_REQUIRE(x && y && z && elem && host_lim && ox && oy && oz && oelem && host_count, "NULL buffer");
  // the array's device address, looked up once per device with that device current
  static std::mutex mu;
  static int64_t *addr[GX_MAX_DEVICES] = {};
  int dev = 0;
  GX_HIP(hipGetDevice(&dev));
  GX_REQUIRE(dev >= 0 && dev < GX_MAX_DEVICES, "device %d beyond the %d the library tracks", dev, GX_MAX_DEVICES);
  int64_t *work = nullptr;
  {
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
      void *p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || p == nullptr)
        return set_error(GX_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
      addr[dev] = static_cast<int64_t *>(p);
    }
    work = addr[dev];
  }
  // every entry a call reads (the nb block counts, then the total) is written by that
  // call before it is read, so nothing needs clearing between calls
  hipStream_t st = as_stream(stream);
  Lim9 lim;
  for (int k = 0; k < 9; ++k) lim.v[k] = host_lim[k];
  const int64_t nb = (n + CROP_B - 1) / CROP_B;
  k_crop_count<<<(unsigned)nb, BLK, 0, st>>>(x, y, z, n, lim, work);
  k_crop_scan<<<1, 1024, 0, st>>>(work, nb);
  k_crop_scatter<<<(unsigned)nb, BLK, 0, st>>>(x, y, z, elem, n, lim, work, ox, oy, oz, oelem);
  int rc = check_launch("gx_slab_crop_noalloc");
  if (rc) return rc;
  GX_HIP(hipMemcpyAsync(host_count, work + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  GX_HIP(hipStreamSynchronize(st));
  return GX_OK;
}

int gx_shift_min(double *x, double *y, double *z, int64_t n, const double *minmax, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(n >= 0, "n=%lld", (long long)n);
  if (n == 0) return GX_OK;
  GX_REQUIRE(x && y && z && minmax, "NULL buffer");
  k_shift_min<<<grid_cap(n), BLK, 0, as_stream(stream)>>>(x, y, z, n, minmax);
  return check_launch("gx_shift_min");
}

}  // extern "C"
