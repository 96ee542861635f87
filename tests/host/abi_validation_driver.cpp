// Host-only driver of the C ABI's argument validation, built together with the library
// sources by tests/test_host_sanitizers.py with host-side ASan and UBSan (no device code).
// Every call below must return an error through the ABI before any device work (NULL or
// inconsistent descriptors and sizes), gx_workspace_size must answer valid queries, and
// the error text must be readable after each call.  Prints "checks passed: <n>" and
// exits 0, or names the first call that misbehaved and exits 1.

#include <stdio.h>
#include <string.h>

#include "giwaxsim_amd.h"

static int n_checks = 0;

static int expect(const char *name, int rc, bool want_error) {
  ++n_checks;
  const char *msg = gx_last_error();
  const bool is_error = rc != GX_OK;
  if (is_error != want_error || (is_error && (msg == nullptr || msg[0] == '\0'))) {
    fprintf(stderr, "%s: rc=%d (%s), message '%s'\n", name, rc, want_error ? "error expected" : "success expected",
            msg ? msg : "(null)");
    return 1;
  }
  return 0;
}

#define ERR(call) \
  if (expect(#call, (call), true)) return 1
#define OK(call) \
  if (expect(#call, (call), false)) return 1

int main() {
  if (gx_abi_version() != GX_ABI_VERSION) {
    fprintf(stderr, "ABI version %d, header %d\n", gx_abi_version(), GX_ABI_VERSION);
    return 1;
  }
  if (!gx_supported_grid(1024) || gx_supported_grid(8) || gx_supported_grid(4096)) return 1;

  gx_slab slab;
  memset(&slab, 0, sizeof slab);
  gx_frames fr;
  memset(&fr, 0, sizeof fr);
  gx_detector det;
  memset(&det, 0, sizeof det);
  gx_adjoint adj;
  memset(&adj, 0, sizeof adj);
  double host9[9] = {0}, out6[6] = {0};
  int64_t count = 0;

  // slab preparation
  ERR(gx_slab_bounds(nullptr, nullptr, nullptr, nullptr));
  ERR(gx_slab_sort_rows(nullptr, nullptr));
  ERR(gx_slab_sort_rows(&slab, nullptr));
  ERR(gx_slab_avg(nullptr, nullptr, 1.0, 1.0, nullptr, nullptr));
  ERR(gx_slab_tile(nullptr, nullptr, 1, 1, 1, 1, host9, nullptr, nullptr, nullptr, nullptr, nullptr));
  ERR(gx_slab_crop(nullptr, nullptr, nullptr, nullptr, 1, host9, nullptr, nullptr, nullptr, nullptr, nullptr,
                   &count, nullptr));
  ERR(gx_slab_crop_noalloc(nullptr, nullptr, nullptr, nullptr, 0, host9, nullptr, nullptr, nullptr, nullptr,
                           &count, nullptr));
  ERR(gx_slab_crop_noalloc(nullptr, nullptr, nullptr, nullptr, 1, host9, nullptr, nullptr, nullptr, nullptr,
                           &count, nullptr));
  ERR(gx_shift_min(nullptr, nullptr, nullptr, 1, nullptr, nullptr));

  // frame pipeline: NULL and zeroed descriptors, then a grid edge out of range
  ERR(gx_frames_stats(nullptr, 0, 1, nullptr));
  ERR(gx_frames_stats(&fr, 0, 1, nullptr));
  ERR(gx_frames_columns(&fr, 0, 1, nullptr));
  ERR(gx_frames_rowpass(&fr, 0, 1, nullptr));
  ERR(gx_frames_colpass(&fr, 0, 1, nullptr));
  ERR(gx_frames_forward(&fr, 0, 1, nullptr));
  fr.N = 4096;
  ERR(gx_frames_rowpass(&fr, 0, 1, nullptr));
  fr.N = 1024;
  fr.r = 0.2;
  fr.batch = 4;
  fr.n_frames = 4;
  ERR(gx_frames_colpass(&fr, 0, 8, nullptr));  // more frames than the batch
  ERR(gx_bins_clear(nullptr, nullptr, 1, nullptr));
  ERR(gx_bins_finalize(nullptr, nullptr, 1, 1.0, 0, 1, 0.0, 1.0, 1.0, host9, 6.0, nullptr, nullptr, nullptr));
  ERR(gx_bins_accumulate_f64(nullptr, nullptr, 1, 1.0, nullptr, nullptr, nullptr));

  // detector
  ERR(gx_detector_accumulate(nullptr, nullptr));
  ERR(gx_detector_accumulate(&det, nullptr));
  ERR(gx_detector_finalize(nullptr, 4, 4, 0, nullptr, nullptr));
  ERR(gx_rotate_points(nullptr, nullptr, nullptr, nullptr, 1, nullptr, nullptr, nullptr, nullptr));
  ERR(gx_intersect(nullptr, 1, 1, 1, 0.0, 0.0, 0.0, 0.1, nullptr, nullptr, nullptr, 1, 1.0, 0, nullptr, nullptr));
  ERR(gx_bin_frame_f64(nullptr, 1, 1, nullptr, nullptr, nullptr, host9, 0.1, 4, nullptr, nullptr, nullptr, 0,
                       nullptr));
  ERR(gx_collapse_lengths(nullptr, nullptr, 1, nullptr, nullptr));
  ERR(gx_apply_f0(nullptr, 1, 1, 1, nullptr, nullptr, nullptr, host9, 6.0, nullptr, nullptr));
  ERR(gx_scale_offset_moments(nullptr, nullptr, nullptr, 1, nullptr, out6, nullptr));

  // comparison helpers
  ERR(gx_map_linear(nullptr, 1, 1, nullptr, nullptr, 1, 0.0, nullptr, nullptr));
  ERR(gx_polar_pad(nullptr, 1, 1, nullptr, 0, nullptr, nullptr));
  ERR(gx_csr_apply(nullptr, nullptr, nullptr, 1, nullptr, 0.0, nullptr, nullptr));
  ERR(gx_fit_residual(nullptr, nullptr, nullptr, 1, 1.0, 0.0, nullptr, nullptr));

  // differentiable variant
  ERR(gx_adj_prepare(nullptr, nullptr, nullptr));
  ERR(gx_adj_prepare(&fr, &adj, nullptr));
  ERR(gx_adj_rowpass(&fr, &adj, 0, 1, nullptr));
  ERR(gx_adj_forward(&fr, &adj, 0, 1, nullptr));
  ERR(gx_adj_colpass(&fr, &adj, 0, 1, nullptr));
  ERR(gx_adj_rowgrad(&fr, &adj, 0, 1, nullptr));
  ERR(gx_adj_reduce(&fr, &adj, 0, 1, nullptr));
  ERR(gx_adj_backward(&fr, &adj, 0, 1, nullptr));
  ERR(gx_adj_backward_stored(&fr, &adj, 0, 1, nullptr));
  ERR(gx_adj_finish(&fr, &adj, nullptr, nullptr, nullptr));
  ERR(gx_bins_adjoint(nullptr, 1, nullptr, nullptr, 1, 0, 1, 0.0, 1.0, 1.0, host9, 6.0, nullptr, nullptr));
  ERR(gx_detector_adjoint(nullptr, nullptr, 0, nullptr, nullptr, 1, nullptr, nullptr));
  ERR(gx_detector_adjoint(&det, nullptr, 0, nullptr, nullptr, 1, nullptr, nullptr));
  ERR(gx_detector_map_count(nullptr, nullptr, nullptr, nullptr, nullptr));
  ERR(gx_detector_map_count(&det, nullptr, nullptr, nullptr, nullptr));
  ERR(gx_detector_map_fill(&det, nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr));
  ERR(gx_detector_adjoint_mapped(&det, nullptr, 0, nullptr, nullptr, nullptr, nullptr, 1, nullptr, nullptr,
                                 nullptr));

  // workspace sizes: valid queries, then each kind of invalid one
  gx_workspace_query q;
  memset(&q, 0, sizeof q);
  gx_workspace_sizes s;
  q.N = 2048;
  q.n_atoms = 1296000;
  q.n_elem = 6;
  q.n_frames = 1307;
  q.batch = 1307;
  q.W = 300;
  q.Z = 416;
  q.c_n = 295;
  q.i_lo = 870;
  q.i_hi = 1178;
  q.fill_bkg = 1;
  q.smooth = 20;
  q.store_band = 1;
  q.differentiable = 1;
  q.n_pix = 2048;
  q.n_init = 1;
  q.n_orient = 721;
  OK(gx_workspace_size(&q, &s));
  if (s.total <= s.inter || s.map_runs_max != 2048LL * 2048 * 721) return 1;
  // the largest query the caps allow: every size and the total stay finite
  gx_workspace_query big = q;
  big.n_atoms = 1LL << 36;
  big.n_elem = 1 << 16;
  big.n_frames = 1LL << 31;
  big.batch = 1 << 20;
  big.W = 2048;
  big.Z = 2048;
  big.c_n = 8192;
  big.i_lo = 0;
  big.i_hi = 2047;
  big.smooth = 2048;
  big.n_pix = 65536;
  big.n_init = 3;
  big.n_orient = 1 << 24;
  big.fft_f64 = 1;
  {
    const int rc = gx_workspace_size(&big, &s);  // either sizes or a refusal, never an overflow
    ++n_checks;
    if (rc != GX_OK && strstr(gx_last_error(), "2^63") == nullptr) return 1;
  }
  ERR(gx_workspace_size(nullptr, &s));
  ERR(gx_workspace_size(&q, nullptr));
  gx_workspace_query b = q;
  b.N = 1048;  // chirp-z edge without chirp-z lengths
  ERR(gx_workspace_size(&b, &s));
  b.bs_M = b.bs_Mc = 2048;  // ... with them, but a stored band is power-of-two only
  ERR(gx_workspace_size(&b, &s));
  b.store_band = 0;
  OK(gx_workspace_size(&b, &s));
  b = q;
  b.n_atoms = (1LL << 36) + 1;
  ERR(gx_workspace_size(&b, &s));
  b = q;
  b.batch = (1 << 20) + 1;
  ERR(gx_workspace_size(&b, &s));
  b = q;
  b.Z = 417;
  ERR(gx_workspace_size(&b, &s));
  b = q;
  b.n_orient = (1 << 24) + 1;
  ERR(gx_workspace_size(&b, &s));

  printf("checks passed: %d\n", n_checks);
  return 0;
}
