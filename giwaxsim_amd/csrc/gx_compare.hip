// gfx950 kernels of the experimental-comparison step that a fit evaluation runs
// after the detector image (reference: tools/comparison.py trim_sim_data 161-191,
// rebin_qmap_and_mask 406-445, linear_polar 469-497, polar_linear 499-532,
// add_pad 534-557, shift_peak 559-592; fit_slabsize.py fullfit_model 19-47).
//
// Every sample position, pad map and rebin weight depends only on the detector and
// experimental geometry; the host plans them once (numpy, the reference's own
// formulas) and the device applies them to each new image:
//   k_map_linear   scipy.ndimage.map_coordinates(order=1, mode='constant') at
//                  given (row, col) positions: bilinear inside [0, h-1] x [0, w-1],
//                  cval outside (no interpolation beyond the edges)
//   k_polar_pad    add_pad's column duplication as a column map (src_col[c] < w), and
//                  optionally the `pad[polar == 0] = 0` mask of shift_peak
//   k_csr_apply    y = W x for precomputed sparse interpolation weights (griddata
//                  'linear': barycentric weights of the Delaunay triangle)
//   k_fit_residual (exp - (scale sim + offset)) * (mask == 0)
#include "gx_device.h"
#include "gx_host.h"

#include <math.h>

using namespace gx;

namespace {

constexpr int BLK = 256;

__global__ __launch_bounds__(BLK) void k_map_linear(const double *__restrict__ img, int64_t h, int64_t w,
                                                    const double *__restrict__ ys, const double *__restrict__ xs,
                                                    int64_t n, double cval, double *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK) {
    const double y = ys[i], x = xs[i];
    double v = cval;
    if (y >= 0.0 && y <= (double)(h - 1) && x >= 0.0 && x <= (double)(w - 1)) {
      const double fy0 = floor(y), fx0 = floor(x);
      const int64_t y0 = (int64_t)fy0, x0 = (int64_t)fx0;
      const double fy = y - fy0, fx = x - fx0;
      // the far neighbour at the last row / column carries weight 0
      const int64_t y1 = y0 + 1 < h ? y0 + 1 : h - 1, x1 = x0 + 1 < w ? x0 + 1 : w - 1;
      const double top = (1.0 - fx) * img[y0 * w + x0] + fx * img[y0 * w + x1];
      const double bot = (1.0 - fx) * img[y1 * w + x0] + fx * img[y1 * w + x1];
      v = (1.0 - fy) * top + fy * bot;
    }
    out[i] = v;
  }
}

__global__ __launch_bounds__(BLK) void k_polar_pad(const double *__restrict__ polar, int64_t h, int64_t w,
                                                   const int32_t *__restrict__ src, int mask_zero,
                                                   double *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < h * w; i += (int64_t)gridDim.x * BLK) {
    const int64_t r = i / w, c = i - r * w;
    const double v = polar[r * w + src[c]];
    out[i] = (mask_zero && polar[i] == 0.0) ? 0.0 : v;
  }
}

// one row per thread: sum_k val[k] x[col[k]] in the stored order
__global__ __launch_bounds__(BLK) void k_csr_apply(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                                                   const double *__restrict__ val, int64_t n_rows,
                                                   const double *__restrict__ x, double empty, double *__restrict__ y) {
  for (int64_t r = blockIdx.x * (int64_t)BLK + threadIdx.x; r < n_rows; r += (int64_t)gridDim.x * BLK) {
    const int64_t a = row_ptr[r], b = row_ptr[r + 1];
    double s = empty;
    if (b > a) {
      s = 0.0;
      for (int64_t k = a; k < b; ++k) s += val[k] * x[col[k]];
    }
    y[r] = s;
  }
}

__global__ __launch_bounds__(BLK) void k_fit_residual(const double *__restrict__ sim, const double *__restrict__ ref,
                                                      const uint8_t *__restrict__ mask, int64_t n, double s, double b,
                                                      double *__restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)BLK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLK)
    out[i] = (ref[i] - (s * sim[i] + b)) * (mask[i] == 0 ? 1.0 : 0.0);
}

unsigned grid_for(int64_t n) {
  const unsigned g = grid1d(n, BLK);
  return g > 16384 ? 16384 : (g ? g : 1);
}

}  // namespace

extern "C" {

int gx_map_linear(const double *img, int64_t h, int64_t w, const double *ys, const double *xs, int64_t n, double cval,
                  double *out, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(h >= 1 && w >= 1, "image %lld x %lld", (long long)h, (long long)w);
  GX_REQUIRE(n >= 0, "n=%lld", (long long)n);
  if (n == 0) return GX_OK;
  GX_REQUIRE(img && ys && xs && out, "NULL buffer");
  k_map_linear<<<grid_for(n), BLK, 0, as_stream(stream)>>>(img, h, w, ys, xs, n, cval, out);
  return check_launch("gx_map_linear");
}

int gx_polar_pad(const double *polar, int64_t h, int64_t w, const int32_t *src_col, int64_t mask_zero, double *out,
                 gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(h >= 1 && w >= 1, "polar image %lld x %lld", (long long)h, (long long)w);
  GX_REQUIRE(polar && src_col && out && out != polar, "NULL or aliased buffer");
  k_polar_pad<<<grid_for(h * w), BLK, 0, as_stream(stream)>>>(polar, h, w, src_col, mask_zero ? 1 : 0, out);
  return check_launch("gx_polar_pad");
}

int gx_csr_apply(const int64_t *row_ptr, const int32_t *col, const double *val, int64_t n_rows, const double *x,
                 double empty, double *y, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(n_rows >= 0, "n_rows=%lld", (long long)n_rows);
  if (n_rows == 0) return GX_OK;
  GX_REQUIRE(row_ptr && col && val && x && y, "NULL buffer");
  k_csr_apply<<<grid_for(n_rows), BLK, 0, as_stream(stream)>>>(row_ptr, col, val, n_rows, x, empty, y);
  return check_launch("gx_csr_apply");
}

int gx_fit_residual(const double *sim, const double *ref, const uint8_t *mask, int64_t n, double scale,
                    double offset, double *out, gx_stream_t stream) {
  clear_error();
  GX_REQUIRE(n >= 0, "n=%lld", (long long)n);
  if (n == 0) return GX_OK;
  GX_REQUIRE(sim && ref && mask && out, "NULL buffer");
  k_fit_residual<<<grid_for(n), BLK, 0, as_stream(stream)>>>(sim, ref, mask, n, scale, offset, out);
  return check_launch("gx_fit_residual");
}

}  // extern "C"
