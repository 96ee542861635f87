// gx_workspace_size: byte sizes of the device buffers the descriptors of the forward
// path, the detector and the differentiable variant point to, from the geometry alone.
// Host-only arithmetic; each size follows the layout the kernels index (the comments of
// the descriptor fields in giwaxsim_amd.h).

#include "gx_host.h"

using namespace gx;

namespace {

bool pow2_edge(int64_t N) { return N == 256 || N == 512 || N == 1024 || N == 2048; }
bool chirp_len(int64_t M) { return M == 1024 || M == 2048 || M == 4096; }

}  // namespace

extern "C" int gx_workspace_size(const gx_workspace_query *q, gx_workspace_sizes *out) {
  clear_error();
  GX_REQUIRE(q != nullptr && out != nullptr, "NULL argument");
  const int64_t N = q->N, n = q->n_atoms, ne = q->n_elem, B = q->batch;
  GX_REQUIRE(gx_supported_grid(N), "N=%lld outside [16, 2048]", (long long)N);
  const bool p2 = pow2_edge(N);
  GX_REQUIRE(p2 ? (q->bs_M == 0 && q->bs_Mc == 0) : (chirp_len(q->bs_M) && chirp_len(q->bs_Mc)),
             "N=%lld needs bs_M = bs_Mc = 0 (powers of two 256..2048) or chirp-z lengths 1024/2048/4096 "
             "(got %lld, %lld)", (long long)N, (long long)q->bs_M, (long long)q->bs_Mc);
  // caps that keep every size below 2^62 bytes (no int64 overflow)
  GX_REQUIRE(n >= 0 && n <= (1ll << 36), "n_atoms=%lld outside [0, 2^36]", (long long)n);
  GX_REQUIRE(ne >= 1 && ne <= (1 << 16), "n_elem=%lld outside [1, 65536]", (long long)ne);
  GX_REQUIRE(q->n_frames >= 1 && q->n_frames <= (1ll << 31) && B >= 1 && B <= (1 << 20),
             "n_frames=%lld (1..2^31), batch=%lld (1..2^20)", (long long)q->n_frames, (long long)B);
  GX_REQUIRE(q->W >= 1 && q->W <= N, "W=%lld outside [1, N]", (long long)q->W);
  GX_REQUIRE(q->Z >= 8 && q->Z <= N && q->Z % 8 == 0, "Z=%lld must be a multiple of 8 in [8, N]",
             (long long)q->Z);
  GX_REQUIRE(q->c_n >= 1 && q->c_n <= 4 * N, "c_n=%lld outside [1, 4 N]", (long long)q->c_n);
  GX_REQUIRE(q->fill_bkg == 0 || q->fill_bkg == 1, "fill_bkg=%lld", (long long)q->fill_bkg);
  GX_REQUIRE(q->smooth >= 0 && q->smooth <= N, "smooth=%lld outside [0, N]", (long long)q->smooth);
  GX_REQUIRE(q->fft_f64 == 0 || q->fft_f64 == 1, "fft_f64=%lld", (long long)q->fft_f64);
  GX_REQUIRE(!q->store_band || (p2 && q->i_lo >= 0 && q->i_lo <= q->i_hi && q->i_hi < N),
             "store_band needs a power-of-two N and 0 <= i_lo <= i_hi < N (got %lld..%lld)", (long long)q->i_lo,
             (long long)q->i_hi);
  GX_REQUIRE(q->n_pix >= 0 && q->n_pix <= 65536, "n_pix=%lld", (long long)q->n_pix);
  GX_REQUIRE(q->n_pix == 0 || (q->n_init >= 0 && q->n_init <= 3 && q->n_orient >= 1 && q->n_orient <= (1 << 24)),
             "detector: n_init=%lld (0..3), n_orient=%lld (1..2^24)", (long long)q->n_init, (long long)q->n_orient);

  gx_workspace_sizes s = {};
  const int64_t esz = q->fft_f64 ? 16 : 8;   // one complex transform value
  const int64_t part = q->fft_f64 ? 8 : 4;   // one real per-frame partial
  const int64_t nb = q->c_n * q->c_n * q->c_n;
  const bool dense = q->fill_bkg || q->smooth;
  const bool diff = q->differentiable != 0;
  const bool det = q->n_pix > 0;
  const int64_t np2 = q->n_pix * q->n_pix;

  s.sx = 8 * n;
  s.sy = 8 * n;
  s.sq = 4 * 2 * n;
  s.szrow = 4 * n;
  s.row_off = 4 * (N + 2);
  s.row_exp = 4 * N;
  s.work = 4 * (N * ne + 2 * (N + 1));
  if (diff) {
    s.szf = 8 * n;
    s.sperm = 4 * n;
    s.selem = 4 * n;
    s.stable_work = 4 * ((n + GX_STABLE_CHUNK - 1) / GX_STABLE_CHUNK) * (N + 1);
  }

  s.fparams = 8 * GX_FP_STRIDE * q->n_frames;
  s.twiddle = esz * N;
  s.twiddle64 = 16 * N;
  if (q->smooth) s.gauss_cw = 8 * (2 * (int64_t)(4.0 * (double)q->smooth + 0.5) + 2);
  if (!p2) {
    s.bs_chirp = esz * N;
    s.bs_kernel = esz * q->bs_M;
    s.bs_twiddle = esz * q->bs_M;
    s.bs_kernel_c = esz * q->bs_Mc;
    s.bs_twiddle_c = esz * q->bs_Mc;
  }
  s.ymin_key = 8 * 3 * B;
  s.idx_bounds = 4 * 4 * B;
  if (q->fill_bkg) s.col_nm = 4 * B * N;
  if (q->smooth) {
    s.col_sy = 8 * B * N;
    s.col_sz = 8 * B * N;
  }
  s.inter = esz * B * q->W * q->Z;
  if (dense) s.col_spec = esz * B * q->W;
  if (q->store_band) s.f_band = esz * B * q->W * (q->i_hi - q->i_lo + 1);
  s.avg_q = 16;
  s.bin_sum = 8 * nb;
  s.bin_cnt = 4 * nb;
  s.status = 4 * 4;
  s.iq = 8 * nb;

  if (det) {
    s.h_axis = 8 * q->n_pix;
    s.v_axis = 8 * q->n_pix;
    s.init_R = 8 * 9 * q->n_init;
    s.orient_R = 8 * 27 * q->n_orient;
    s.orient_w = 8 * q->n_orient;
    s.det = 8 * np2;
    s.det_out = 8 * np2;
  }

  if (diff) {
    const int64_t red_a = B * (N + 1), red_b = 2 * (1 + ne) * GX_ADJ_RED_BLOCKS;
    s.adj_row_exp = 4 * N;
    s.ainv = 4 * n;
    s.g_bins = 8 * nb;
    s.g_part = part * B * 4 * n;
    s.red = 8 * (red_a > red_b ? red_a : red_b);
    s.amin = 4 * (B + 7);
    s.g_atom = 8 * 5 * n;
    if (dense) {
      s.avg_spec = 16 * B * (N + q->W);
      s.avg_part = 16 * B * q->W;
    }
    s.g_avg = 16;
    if (ne <= GX_ADJ_ELEM_REG) {
      s.e_part = 16 * B * q->Z * ne;
      s.e_sum = 16 * ne;
    }
    if (!p2) {
      s.bs_kernel_adj = esz * q->bs_M;
      s.bs_kernel_adj_c = esz * q->bs_Mc;
    }
    s.g_coords = 8 * 3 * n;
    s.g_f = 8 * 2 * ne;
    if (det) {
      s.det_work = 8 * (3 * np2 + (np2 + 255) / 256);
      s.g_fx = 8 * nb;
      s.fx_scale = 8;
      s.map_counts = 4 * nb;
      s.map_scan_work = 8 * ((nb + 2047) / 2048 + 1);
      s.map_voxel_off = 8 * (nb + 1);
      s.map_run_bytes = 4 + 4 + 1;
      s.map_runs_max = np2 * q->n_orient;
    }
  }

  const int64_t summed[] = {
      s.sx, s.sy, s.sq, s.szrow, s.row_off, s.row_exp, s.work, s.szf, s.sperm, s.selem, s.stable_work,
      s.fparams, s.twiddle, s.twiddle64, s.gauss_cw, s.bs_chirp, s.bs_kernel, s.bs_twiddle, s.bs_kernel_c,
      s.bs_twiddle_c, s.ymin_key, s.idx_bounds, s.col_nm, s.col_sy, s.col_sz, s.inter, s.col_spec, s.f_band,
      s.avg_q, s.bin_sum, s.bin_cnt, s.status, s.iq, s.h_axis, s.v_axis, s.init_R, s.orient_R, s.orient_w,
      s.det, s.det_out, s.adj_row_exp, s.ainv, s.g_bins, s.g_part, s.red, s.amin, s.g_atom, s.avg_spec,
      s.avg_part, s.g_avg, s.e_part, s.e_sum, s.bs_kernel_adj, s.bs_kernel_adj_c, s.g_coords, s.g_f,
      s.det_work, s.g_fx, s.fx_scale, s.map_counts, s.map_scan_work, s.map_voxel_off};
  for (int64_t v : summed)
    GX_REQUIRE(!__builtin_add_overflow(s.total, v, &s.total), "the total workspace exceeds 2^63 bytes");
  *out = s;
  return GX_OK;
}
