/*
 * giwaxsim_amd — C ABI of the MI355X (gfx950) forward-scattering path.
 *
 * The reference (tchaney97/giwaxsim) is pure Python/numpy; it has no native
 * FFI.  Each entry point below replaces the reference function(s) cited in
 * its comment; the Python package `giwaxsim_amd.tools` binds them through
 * ctypes behind the reference's own call signatures.
 *
 * Conventions
 *   - Every function returns GX_OK (0) or an error code; gx_last_error()
 *     returns a message for the calling thread's last failure.  No C++
 *     exception crosses this boundary.
 *   - All array pointers are DEVICE pointers unless the name says host_.
 *     The library never allocates device memory: the caller passes every
 *     buffer (sizes given by the comments), so calls are safe to enqueue
 *     into a caller's stream and graph.
 *   - `stream` is a hipStream_t (NULL = the null stream).  Work is only
 *     enqueued; nothing synchronises the host, except the entry points that
 *     return a value in host memory (gx_scale_offset_moments, gx_slab_crop),
 *     which wait for their own stream and say so.
 *   - Structures contain only 8-byte members (int64_t, double, pointers).
 */
#ifndef GIWAXSIM_AMD_H
#define GIWAXSIM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GX_ABI_VERSION 12

#define GX_OK 0
#define GX_ERR_ARG 1         /* invalid argument (message says which) */
#define GX_ERR_UNSUPPORTED 2 /* e.g. a grid edge without a compiled FFT plan */
#define GX_ERR_HIP 3         /* a HIP runtime call failed */

typedef void *gx_stream_t;

/* ------------------------------------------------------------------------------------
 * Per-frame record: GX_FP_STRIDE doubles per phi-frame, prepared by the host
 * planner from the frame angle (tools/voxelgrids.py:238-309, 381-402).
 * ---------------------------------------------------------------------------------- */
#define GX_FP_STRIDE 24
enum {
  GX_FP_SIN = 0,   /* sin(radians(phi))   rotate_coords_z, tools/utilities.py:305-312 */
  GX_FP_COS,       /* cos(radians(phi)) */
  GX_FP_HX_START,  /* det_h_qx = linspace(start, stop, N), tools/voxelgrids.py:398-401 */
  GX_FP_HX_STEP,
  GX_FP_HX_STOP,
  GX_FP_HY_START,  /* det_h_qy = linspace(start, stop, N), tools/voxelgrids.py:396-400 */
  GX_FP_HY_STEP,
  GX_FP_HY_STOP,
  GX_FP_LEN_MODE,  /* rectangular_collapse_lengths branch: 0 general, 1 phi==0, 2 phi==90 */
  GX_FP_HOR,       /* hor_length / ver_length after the phi>90 swap (voxelgrids.py:253-255) */
  GX_FP_VER,
  GX_FP_STOP1,     /* voxelgrids.py:279-285 */
  GX_FP_STOP2,
  GX_FP_S12,       /* stop1 + stop2 */
  GX_FP_MID,       /* mid_length */
  GX_FP_A,         /* sqrt(ver^2 - (ver cos theta)^2) */
  GX_FP_TP,        /* tan(phi) */
  GX_FP_VC,        /* ver cos(theta) */
  GX_FP_TT,        /* tan(theta) */
  GX_FP_CP,        /* cos(phi) */
  GX_FP_CT,        /* cos(theta) */
  GX_FP_JLO,       /* first fftshifted column whose bin lies in the cropped grid */
  GX_FP_JHI,       /* last such column (JHI < JLO: none) */
  GX_FP_UNUSED
};

/* ------------------------------------------------------------------------------------
 * A slab of atoms prepared for the phi-frame loop.
 * ---------------------------------------------------------------------------------- */
typedef struct gx_slab {
  int64_t n_atoms;
  int64_t n_elem;         /* entries of f_table */
  int64_t N;              /* real-space grid edge (grid_size) */
  double r;               /* r_voxel_size (A) */
  double zmin;            /* min z over all atoms */
  const double *x;        /* [n_atoms] */
  const double *y;        /* [n_atoms] */
  const double *z;        /* [n_atoms] */
  const int32_t *elem;    /* [n_atoms] index into f_table */
  const double *f_table;  /* [n_elem][2] f = Z + f1 + i f2 */
  /* outputs, atoms regrouped by projected z row (rows 0..N-1, then rows >= N) */
  double *sx;             /* [n_atoms] */
  double *sy;             /* [n_atoms] */
  int32_t *sq;            /* [2 n_atoms] fixed-point (re, im) of f, scale 2^row_exp[row] */
  int32_t *szrow;         /* [n_atoms] z row of each regrouped atom (N for rows >= N) */
  int32_t *row_off;       /* [N + 2] first atom of each row; row_off[N] = atoms with row < N */
  int32_t *row_exp;       /* [N] fixed-point exponent of each row (gx_slab_sort_rows writes one
                             slab-wide value, the largest that no row's sum can overflow) */
  int32_t *work;          /* [N * n_elem + 2 * (N + 1)] scratch counters; the first N * n_elem
                             entries keep the per-row element counts after the call */
  /* optional outputs (NULL = not written), used by the differentiable variant */
  double *szf;            /* [n_atoms] fractional z offset (z - zmin) / r - row, in [0, 1] */
  int32_t *sperm;         /* [n_atoms] original index of each regrouped atom */
  int32_t *selem;         /* [n_atoms] element id of each regrouped atom */
  int32_t *stable_work;   /* optional [ceil(n_atoms / GX_STABLE_CHUNK)][N + 1] scratch: when given
                             (N <= 4096), atoms keep their original order within each row, so
                             the regrouped order is reproducible */
} gx_slab;
#define GX_STABLE_CHUNK 4096

/* ------------------------------------------------------------------------------------
 * The phi-frame pipeline of one prepared slab into a cropped 3D bin grid.
 * ---------------------------------------------------------------------------------- */
typedef struct gx_frames {
  int64_t N;               /* real-space grid edge: 16..2048 (powers of two 256..2048 use
                              in-wave transforms, other edges chirp-z transforms) */
  double r;                /* r_voxel_size */
  int64_t n_atoms;         /* atoms (ymin runs over all of them) */
  const double *ax;        /* regrouped atoms (gx_slab.sx) */
  const double *ay;        /* gx_slab.sy */
  const int32_t *aq;       /* gx_slab.sq */
  const int32_t *azrow;    /* gx_slab.szrow */
  const int32_t *row_off;  /* gx_slab.row_off */
  const int32_t *row_exp;  /* gx_slab.row_exp */
  /* background fill and edge smoothing (tools/voxelgrids.py:343-379) */
  int64_t fill_bkg;        /* 0/1 */
  int64_t smooth;          /* sigma in voxels, 0 = off */
  double max_voxels;       /* sqrt(x_b^2 + y_b^2) // r */
  double avg_re, avg_im;   /* avg_voxel_f */
  const double *gauss_cw;  /* [2 smooth_radius + 2] prefix sums of the normalised Gaussian
                              kernel w[-R..R]: gauss_cw[t] = sum_{d < t - R} w[d] */
  int64_t smooth_radius;   /* R = int(4 sigma + 0.5) */
  int64_t zrow_max;        /* largest projected z row of any atom if every row is < N, else -1 */
  /* q geometry */
  double qmin;             /* np.min of the full q axis */
  double dq;               /* np.diff(qz)[0] */
  int64_t c_lo;            /* first index of the cropped window on every axis */
  int64_t c_n;             /* cropped edge c */
  double q_val;            /* 1.0 / (N * r), fftfreq */
  int64_t i_lo, i_hi;      /* fftshifted rows whose bin lies in the crop */
  int64_t W;               /* intermediate row width (>= JHI-JLO+1 of every frame) */
  int64_t Z;               /* z rows that can hold atoms: [0, Z), a multiple of 8, <= N;
                              rows >= Z are synthesised by the column pass */
  double bin_scale;        /* fixed-point scale of I in bin sums (a power of two) */
  int64_t fft_f64;         /* 0: FP32 transforms (float2), 1: FP64 transforms (double2) */
  const void *twiddle;     /* [N] exp(-2 pi i m / N) as float2 or double2 (fft_f64) */
  const double *twiddle64; /* [N] the same as double2 (per-frame background spectra) */
  /* grid edges that are not powers of two: zoomed chirp-z (Bluestein) transforms.  The
     row pass evaluates the DFT only at k = bs_rlo + j, j < bs_rk (signed frequencies
     covering every frame's kept columns), as a circular convolution of length bs_M,
     exact for rows whose deposits stay below L = min(N, bs_M - bs_rk + 1) in y: bs_M
     >= N + bs_rk - 1 covers every row; a shorter bs_M (sized to the slab's projected
     width) is checked per frame (y_max of idx_bounds >= L sets status[1] and the frame
     is skipped).  The column pass likewise at k = bs_clo + j, j < bs_ck (the kept rows
     [i_lo, i_hi]), length bs_Mc >= N + bs_ck - 1, or shorter when it still takes every
     row z < Z (bs_Mc - bs_ck + 1 >= Z) and fill_bkg is set or smoothing is off: rows >= Z
     then carry nothing but the k = 0 column's fill constant, whose transform on the rows
     the band leaves out is added in closed form */
  int64_t bs_M;            /* 0: N is a power of two (in-wave transforms); else 1024, 2048 or 4096 */
  const void *bs_chirp;    /* [N] c[n] = exp(-i pi (n^2 mod 2N) / N) as float2 / double2 */
  const void *bs_kernel;   /* [bs_M] FFT of h[t] = conj(c[|bs_rlo + t|]), lags t in
                              [-(L-1), bs_rk) at t mod bs_M, divided by bs_M */
  const void *bs_twiddle;  /* [bs_M] exp(-2 pi i m / bs_M) */
  int64_t n_frames;        /* frames described by fparams */
  const double *fparams;   /* [n_frames][GX_FP_STRIDE] */
  /* per-frame scratch for a batch of up to `batch` frames */
  int64_t batch;
  unsigned long long *ymin_key; /* [batch][3] order keys of min y' (all atoms), min and
                                   max y' (atoms with z row < N) */
  int32_t *idx_bounds;     /* [batch][4] y_min, y_max, z_min, z_max */
  int32_t *col_nm;         /* [batch][N] num_missing_voxels per column (fill_bkg) */
  double *col_sy;          /* [batch][N] smoothed y mask (smooth) */
  double *col_sz;          /* [batch][N] smoothed z mask (smooth) */
  void *inter;             /* [batch][W][Z] (inter_rm = 0) or [batch][Z][W] (inter_rm = 1) float2 or
                              double2: row-transformed, column-pruned frame */
  void *col_spec;          /* [batch][W] float2 or double2: spectrum of the frame's background /
                              smoothing terms on the kept columns (fill_bkg or smooth) */
  /* output: cropped bins [c][c][c] indexed [qy][qx][qz] */
  unsigned long long *bin_sum; /* fixed-point sum of I (scale bin_scale) */
  int32_t *bin_cnt;
  int32_t *status;         /* [4] counters; [0] += frames without any valid atom */
  int64_t row_variant;     /* row pass shape (same results): bits 0-1: 0 one wave per z row,
                              1 / 2 a block deposits 4 / 8 rows at once, then one wave per row
                              transforms; + 4: XCD-aware tile order (the frames of a row block
                              run together on one XCD and share its atoms in that L2) instead
                              of frame-major order */
  /* zoomed chirp-z bands (bs_M != 0 only) */
  int64_t bs_rlo, bs_rk;   /* row pass: first signed frequency and band length */
  int64_t bs_Mc;           /* column pass: transform length, 1024, 2048 or 4096 */
  int64_t bs_clo, bs_ck;   /* column pass: first signed frequency and band length */
  const void *bs_kernel_c; /* [bs_Mc] the column band's kernel (as bs_kernel) */
  const void *bs_twiddle_c; /* [bs_Mc] exp(-2 pi i m / bs_Mc) */
  const double *avg_q;     /* NULL, or device [2] avg_voxel_f of the f values the deposit adds
                              (gx_slab_avg), used in place of avg_re, avg_im so that the fill
                              and smoothing terms match the deposited atoms */
  int64_t inter_rm;        /* 0: inter is [batch][W][Z] (the row pass applies the background and
                              z-mask terms and stores transposed); 1 (N = 512, 1024, 2048 with
                              the per-wave row shape): inter is [batch][Z][W], the row pass writes
                              the raw kept band of each row straight from registers
                              and the column pass applies those terms.  gx_frames_rowpass and
                              gx_frames_colpass of one frame batch must use the same value */
  int64_t fft_plan_row;    /* radix plan of the row / column pass transforms (N = 1024: 0 = 16.16.4, */
  int64_t fft_plan_col;    /* 1 = 16.8.8, 2 = 8.16.8, 3 = 8.8.16; N = 2048: 0 = 16.16.8, 1 = 16.8.16,
                              2 = 8.16.16; other N and chirp-z grids: 0 only); the bins agree to
                              rounding across plans */
  void *f_band;            /* NULL, or [batch][W][i_hi - i_lo + 1] float2 / double2: the column
                              pass also stores each binned column's transform at the kept rows
                              i_lo..i_hi (natural order), which the differentiable backward of
                              the same batch reads back (gx_adj_backward_stored).  Power-of-two
                              N with inter_rm = 0 only */
} gx_frames;

/* ------------------------------------------------------------------------------------
 * Detector orientation sum (tools/detector.py:5-300, tools/comparison.py:790-870).
 * ---------------------------------------------------------------------------------- */
typedef struct gx_detector {
  int64_t n_pix;           /* pixels per edge; image is [n_pix (v)][n_pix (h)] */
  int64_t n_init;          /* 0..3 initial rotations */
  int64_t n_orient;        /* orientations (psi outer, phi, theta inner) */
  const double *h_axis;    /* [n_pix] linspace(-max_q, max_q, n) */
  const double *v_axis;    /* [n_pix] */
  const double *init_R;    /* [n_init][9] row-major */
  const double *orient_R;  /* [n_orient][27] R_psi, R_phi, R_theta row-major */
  const double *orient_w;  /* [n_orient] psi_w * phi_w * theta_w */
  const void *iq;          /* [cy][cx][cz] cropped I(q), float or double */
  int64_t iq_f64;          /* 1: iq is double, 0: float */
  int64_t cy, cx, cz;
  double qx_min, qy_min, qz_min, dq;
  double *det;             /* [n_pix][n_pix] FP64 accumulated image (overwritten) */
} gx_detector;

const char *gx_last_error(void);
int gx_abi_version(void);
/* 1 if grid edge N is supported (16 <= N <= 2048). */
int gx_supported_grid(int64_t N);

/* Slab bounds: min/max of x, y, z (device out[6] = xmin,xmax,ymin,ymax,zmin,zmax)
 * and per-element atom counts (device counts[n_elem], int64).
 * Replaces the reductions of tools/comparison.py:712-714, 742. */
int gx_slab_bounds(const gx_slab *s, double *out_minmax, int64_t *out_elem_count, gx_stream_t stream);

/* Groups atoms by their phi-invariant projected z row
 * (z_indices of tools/voxelgrids.py:329; the rotation leaves z unchanged) and
 * quantises each atom's f to the row's fixed-point scale. */
int gx_slab_sort_rows(const gx_slab *s, gx_stream_t stream);

/* avg_voxel_f = (sum f / (x_b y_b z_b)) r^3 (tools/comparison.py:741-744) of the f values
 * that gx_slab_sort_rows' fixed point deposits, f'_e = rint(f_e 2^S) 2^-S with the slab's
 * exponent S: device out[2] = (re, im) from the device per-element atom counts
 * elem_count[n_elem] (all atoms), volume = x_b y_b z_b and r3 = r_voxel_size^3.
 * gx_frames.avg_q takes it, so the background fill matches the deposit exactly. */
int gx_slab_avg(const gx_slab *s, const int64_t *elem_count, double volume, double r3, double *out,
                gx_stream_t stream);

/* Per-frame statistics for frames [f0, f0+nf) of d->fparams: min rotated y
 * (tools/voxelgrids.py:319-323) and the index bounds y_min, y_max, z_min,
 * z_max over valid atoms (voxelgrids.py:328-349). */
int gx_frames_stats(const gx_frames *d, int64_t f0, int64_t nf, gx_stream_t stream);
/* Per-column tables: num_missing_voxels (voxelgrids.py:350-354, 238-309) and
 * the smoothed edge masks (voxelgrids.py:363-374); then, for fill_bkg or
 * smooth, the FP64 spectrum along y of the frame's background and blend
 * terms on the kept columns (d->col_spec). */
int gx_frames_columns(const gx_frames *d, int64_t f0, int64_t nf, gx_stream_t stream);
/* Atom scatter (voxelgrids.py:337-341) + fill (355-361) + blend (377-379) +
 * FFT along y of every z row; writes only the column band that reaches the
 * cropped bins (d->inter).  By linearity only rows holding atoms are
 * transformed; the fill/blend terms enter through d->col_spec. */
int gx_frames_rowpass(const gx_frames *d, int64_t f0, int64_t nf, gx_stream_t stream);
/* FFT along z of every kept column + |F|^2 (voxelgrids.py:387-392) + binning
 * into the cropped grid with counts (process_file2, voxelgrids.py:464-503). */
int gx_frames_colpass(const gx_frames *d, int64_t f0, int64_t nf, gx_stream_t stream);
/* stats + columns + rowpass + colpass. */
int gx_frames_forward(const gx_frames *d, int64_t f0, int64_t nf, gx_stream_t stream);

/* Zero the cropped bins (c^3 sums and counts). */
int gx_bins_clear(unsigned long long *bin_sum, int32_t *bin_cnt, int64_t n_bins, gx_stream_t stream);
/* iq = (count ? sum / count : 0) * ((f0_C(|q|) / Z_C)^2) over the cropped grid
 * (tools/comparison.py:765-788; tools/voxelgrids.py:16-48, 828-857).
 * q axis value of crop index i: qa_step * (c_lo + i) + qa_start, or qa_stop for
 * the last full-axis index (numpy linspace).  aff = 9 Cromer-Mann coefficients.
 * Writes FP64 iq64 and/or FP32 iq32 (either may be NULL). */
int gx_bins_finalize(const unsigned long long *bin_sum, const int32_t *bin_cnt, int64_t c_n,
                     double bin_scale, int64_t c_lo, int64_t q_num, double qa_start,
                     double qa_step, double qa_stop, const double *host_aff, double z_elem,
                     double *iq64, float *iq32, gx_stream_t stream);

/* Orientation-summed detector image (generate_detector_ints over all
 * orientations, tools/detector.py:234-244, 278-300; intersect_detector
 * 194-232), accumulated in FP64 in (psi, phi, theta) order. */
int gx_detector_accumulate(const gx_detector *d, gx_stream_t stream);
/* flags & GX_FIN_MIRROR: mirror_vertical_horizontal (tools/detector.py:246-275);
 * then flags & GX_FIN_FLOOR_SCALE: NaN -> 1e-6, <= 0 -> 1e-6, * 1e-6
 * (tools/comparison.py:857-868).  det and out must not alias. */
#define GX_FIN_MIRROR 1
#define GX_FIN_FLOOR_SCALE 2
int gx_detector_finalize(const double *det, int64_t n_rows, int64_t n_cols, int64_t flags,
                         double *out, gx_stream_t stream);

/* (R @ vstack([x, y, z])) of n points, rows evaluated left to right without
 * FMA (tools/detector.py:71, 113, 155; tools/utilities.py:315). */
int gx_rotate_points(const double *host_R, const double *x, const double *y, const double *z,
                     int64_t n, double *ox, double *oy, double *oz, gx_stream_t stream);

/* One detector gather (intersect_detector, tools/detector.py:194-232) of
 * n points into out (FP64), scaled by `weight` and added to out if
 * accumulate != 0, else stored. */
int gx_intersect(const double *iq, int64_t cy, int64_t cx, int64_t cz, double qx_min, double qy_min,
                 double qz_min, double dq, const double *x, const double *y, const double *z,
                 int64_t n, double weight, int64_t accumulate, double *out, gx_stream_t stream);

/* process_file2 (tools/voxelgrids.py:464-503) on one FP64 frame image
 * iq2d [nv][nh] with axes det_h_qx/det_h_qy [nh] and det_v_qz [nv]:
 * masks against host_qlim = {qx_min, qx_max, qy_min, qy_max, qz_min, qz_max},
 * bins (q - q_min) // dq and adds I and 1 into FP64 grids
 * [q_num][q_num][q_num] indexed [qy][qx][qz].  Pixels whose bin falls
 * outside the grid are skipped and counted in status[0] (the reference would
 * raise IndexError).  monotone_axes = 1 (each of the three axes sorted, as the
 * frame axes of tools/voxelgrids.py:394-402 are): every bin's pixels form one
 * rectangle and are added in np.add.at's order, so the grids are bit-identical to
 * the reference's and deterministic; 0: any axes, one FP64 atomic add per pixel. */
int gx_bin_frame_f64(const double *iq2d, int64_t nv, int64_t nh, const double *hqx,
                     const double *hqy, const double *vqz, const double *host_qlim, double dq,
                     int64_t q_num, double *grid_sum, double *grid_cnt, int32_t *status,
                     int64_t monotone_axes, gx_stream_t stream);

/* rectangular_collapse_lengths (tools/voxelgrids.py:238-309) for n x values
 * using a frame record (host_fp, GX_FP_STRIDE doubles) from the planner. */
int gx_collapse_lengths(const double *host_fp, const double *x, int64_t n, double *out,
                        gx_stream_t stream);

/* add_f0_q_3d (tools/voxelgrids.py:828-857) on an FP64 grid iq [ny][nx][nz]
 * with axes qx [nx], qy [ny], qz [nz]: out = iq * ((f0(|q|)/z_elem)^2),
 * host_aff = 9 Cromer-Mann coefficients.  out may alias iq. */
int gx_apply_f0(const double *iq, int64_t ny, int64_t nx, int64_t nz, const double *qx, const double *qy,
                const double *qz, const double *host_aff, double z_elem, double *out, gx_stream_t stream);

/* Fold fixed-point bins into FP64 sum/count grids of the same shape:
 * grid_sum += bin_sum / bin_scale, grid_cnt += bin_cnt (the shared-memory
 * accumulators of tools/voxelgrids.py:471-472, 502-503). */
int gx_bins_accumulate_f64(const unsigned long long *bin_sum, const int32_t *bin_cnt, int64_t n, double bin_scale,
                           double *grid_sum, double *grid_cnt, gx_stream_t stream);

/* Reductions for optimize_scale_offset (tools/comparison.py:873-882): over
 * pixels with mask == 0, host_out[0..5] = n, mean_s, mean_y, Sss, Ssy, (FP64,
 * centred second moments).  Blocks until done (host output). */
int gx_scale_offset_moments(const double *sim, const double *exp_, const uint8_t *mask, int64_t n,
                            double *work /* [GX_MOMENTS_WORK] */, double *host_out, gx_stream_t stream);
#define GX_MOMENTS_WORK 774

/* ------------------------------------------------------------------------------------
 * Slab builder of a fit evaluation (tools/comparison.py slabmaker_fitting 595-671).
 * ---------------------------------------------------------------------------------- */
/* Replicate a unit cell: atom o = ((kz (num_y + 1) + ky) (num_x + 1) + kx) n_cell + i of
 * the reference's concatenations (x copies, then y copies of that set, then z copies)
 * at cell[i] + kx a + ky b + kz c, added in that order with a = (v0, 0, 0),
 * b = (v1, v2, 0), c = (v3, v4, v5) (host_vec[6]); elem[o] = cell_elem[i].  Outputs
 * [n_cell (num_x + 1)(num_y + 1)(num_z + 1)] (< 2^31).  Replaces comparison.py:604-633. */
int gx_slab_tile(const double *cell_xyz, const int32_t *cell_elem, int64_t n_cell, int64_t num_x, int64_t num_y,
                 int64_t num_z, const double *host_vec, double *x, double *y, double *z, int32_t *elem,
                 gx_stream_t stream);
/* The centred crop (comparison.py:648-663): keep atom a when lo_k <= c_k - min_k <= hi_k
 * on every axis (host_lim[9] = x, y, z minima, then lo, hi per axis), written shifted
 * (c_k - min_k) in the original order.  work [ceil(n / 2048) + 1].  Blocks until the
 * kept count is in *host_count. */
int gx_slab_crop(const double *x, const double *y, const double *z, const int32_t *elem, int64_t n,
                 const double *host_lim, double *ox, double *oy, double *oz, int32_t *oelem, int64_t *work,
                 int64_t *host_count, gx_stream_t stream);

/* No counterpart in the reference.  gx_slab_crop for callers that cannot allocate device
 * memory (code under HIP graph capture, tight fit loops, harnesses that forbid hipMalloc):
 * the same outputs, with the block counts in a library-owned device array instead of the
 * caller's `work` (GX_CROP_NOALLOC_MAX_ATOMS / 2048 + 1 int64, 512 KiB per device, held until
 * the process ends).  n <= GX_CROP_NOALLOC_MAX_ATOMS, else GX_ERR_ARG.  Calls share that
 * array: two calls that overlap on one device (other threads or streams) race.  Allocates
 * nothing; blocks until the kept count is in *host_count. */
#define GX_CROP_NOALLOC_MAX_ATOMS 134217728
#define GX_MAX_DEVICES 64
int gx_slab_crop_noalloc(const double *x, const double *y, const double *z, const int32_t *elem, int64_t n,
                         const double *host_lim, double *ox, double *oy, double *oz, int32_t *oelem,
                         int64_t *host_count, gx_stream_t stream);
/* x -= minmax[0], y -= minmax[2], z -= minmax[4] (gx_slab_bounds' device output):
 * the final shift to the origin (comparison.py:667-669). */
int gx_shift_min(double *x, double *y, double *z, int64_t n, const double *minmax, gx_stream_t stream);

/* ------------------------------------------------------------------------------------
 * Experimental comparison of a fit evaluation (tools/comparison.py, fit_slabsize.py).
 * Sample positions, pad maps and rebin weights depend on geometry only and are
 * planned on the host; these entry points apply them to device images.
 * ---------------------------------------------------------------------------------- */
/* scipy.ndimage.map_coordinates(img, (ys, xs), order=1, mode='constant', cval) on a
 * row-major [h][w] FP64 image: bilinear inside [0, h-1] x [0, w-1], cval outside.
 * Replaces the map_coordinates calls of linear_polar / polar_linear
 * (tools/comparison.py:496, 531). */
int gx_map_linear(const double *img, int64_t h, int64_t w, const double *ys, const double *xs, int64_t n, double cval,
                  double *out, gx_stream_t stream);
/* add_pad (tools/comparison.py:534-557) as a column map, out[r][c] =
 * polar[r][src_col[c]] (0 <= src_col[c] < w) on [h][w] FP64 images, and with
 * mask_zero = 1 shift_peak's `polar_image_pad[polar_image == 0] = 0` (577): out is 0
 * wherever polar is.  out must not alias polar. */
int gx_polar_pad(const double *polar, int64_t h, int64_t w, const int32_t *src_col, int64_t mask_zero, double *out,
                 gx_stream_t stream);
/* y[r] = sum_k val[k] x[col[k]] over row r of a CSR matrix (row_ptr [n_rows + 1]);
 * rows without entries get `empty`.  With the barycentric weights of
 * scipy.interpolate.griddata(method='linear') it replaces the two griddata calls of
 * rebin_qmap_and_mask (tools/comparison.py:431, 436). */
int gx_csr_apply(const int64_t *row_ptr, const int32_t *col, const double *val, int64_t n_rows, const double *x,
                 double empty, double *y, gx_stream_t stream);
/* Residual of fullfit_model (fit_slabsize.py:42-45):
 * out = (ref - (scale sim + offset)) * (mask == 0). */
int gx_fit_residual(const double *sim, const double *ref, const uint8_t *mask, int64_t n, double scale,
                    double offset, double *out, gx_stream_t stream);

/* ------------------------------------------------------------------------------------
 * Differentiable variant of the phi-frame path and its hand-written adjoint.
 *
 * The reference has no gradients; these entry points have no counterpart in it.
 * Its nearest-grid-point deposit (np.add.at at (z // r, y' // r),
 * tools/voxelgrids.py:328-341) is piecewise constant in the atom positions, so
 * this variant deposits with cloud-in-cell weights instead: an atom at grid
 * coordinates u = (y' - min y') / r, v = (z - min z) / r adds
 *   f (1 - tv)(1 - tu), f (1 - tv) tu, f tv (1 - tu), f tv tu
 * to nodes (v0, u0), (v0, u0 + 1), (v0 + 1, u0), (v0 + 1, u0 + 1) with
 * u0 = u // 1, tu = u - u0 (likewise v).  Everything after the deposit (FFT,
 * |F|^2, binning with counts, mean, f0, detector gather, mirror, floor) is the
 * forward path's, and every buffer of gx_frames keeps its meaning.
 * fill_bkg / smooth act on the deposited grid S exactly as in the forward path;
 * almost everywhere in the positions the frame grid is then
 *   G = M S + avg_voxel_f p,  M = (mz inz) x (my iny),  p = Pz x Py + max_voxels
 * with real masks (smoothed edges my, mz; filled region iny, inz) and a real
 * pattern p, so dL/dS = M dL/dG and dL/d avg_voxel_f = sum_k dL/dF_k conj(FFT2(p)_k);
 * avg_voxel_f = sum f r^3 / (x_bound y_bound z_bound) passes that on to f and to
 * the six atoms that fix the bounding box.
 * Requirements: power-of-two N, every atom's u0 + 1 and v0 + 1 inside the grid,
 * and d->Z > the largest v0 + 1.
 * ---------------------------------------------------------------------------------- */
#define GX_ADJ_RED_BLOCKS 16
#define GX_ADJ_ELEM_REG 6   /* up to this many elements, dL/df is reduced per element in-block */
typedef struct gx_adjoint {
  int64_t n_elem;
  const double *f_table;   /* [n_elem][2] per-element f (re, im) */
  const double *afz;       /* gx_slab.szf */
  const int32_t *aperm;    /* gx_slab.sperm */
  const int32_t *aelem;    /* gx_slab.selem */
  const int32_t *ecnt;     /* gx_slab.work: [N][n_elem] per-row element counts */
  const double *z;         /* [n_atoms] original z (for the argmin of z) */
  double zmin;             /* gx_slab.zmin */
  int32_t *row_exp;        /* [N] deposit exponents (written by gx_adj_prepare) */
  int32_t *ainv;           /* [n_atoms] regrouped index of each original atom (gx_adj_prepare) */
  /* backward */
  const double *g_bins;    /* [c^3] dL / d(bin sum of I) (gx_bins_adjoint) */
  void *g_part;            /* [batch][4][n_atoms] per-frame dL/dy', dL/dz, dL/df re, im; float
                              with FP32 transforms (fft_f64 = 0), double with FP64; with
                              n_elem <= GX_ADJ_ELEM_REG only dL/dy', dL/dz are written (dL/df
                              goes to e_part) */
  double *red;             /* [max(batch (N + 1), 2 (1 + n_elem) GX_ADJ_RED_BLOCKS)] reduction scratch */
  int32_t *amin;           /* [batch + 7] argmin / argmax atoms (original index) */
  double *g_atom;          /* [5][n_atoms] accumulated dL/dx, dL/dy, dL/dz, dL/df re, im per
                              regrouped atom (zeroed by gx_adj_prepare) */
  /* fill_bkg / smooth only (may be NULL otherwise) */
  const double *x;         /* [n_atoms] original x */
  const double *y;         /* [n_atoms] original y */
  const double *ext;       /* [6] xmin, xmax, ymin, ymax, zmin, zmax (gx_slab_bounds) */
  const int64_t *ecount;   /* [n_elem] atoms per element (gx_slab_bounds) */
  double vol_r3;           /* r^3 / (x_bound y_bound z_bound) */
  void *avg_spec;          /* [batch][N + W] double2: FFT of Pz (all k), FFT of Py (kept columns) */
  double *avg_part;        /* [batch][W][2] per-column share of dL/d avg_voxel_f */
  double *g_avg;           /* [2] dL/d avg_voxel_f (re, im), accumulated (zeroed by gx_adj_prepare) */
  /* n_elem <= GX_ADJ_ELEM_REG only (may be NULL otherwise): dL/df summed per element
     inside each gradient block instead of per atom */
  double *e_part;          /* [batch][Z][n_elem][2] per frame and row block (at most Z blocks) */
  double *e_sum;           /* [n_elem][2] accumulated over batches (zeroed by gx_adj_prepare) */
  /* grid edges that are not powers of two (gx_frames.bs_M != 0): kernel spectra of the
     adjoint band transforms (band values -> grid points), same lengths and layouts as
     gx_frames.bs_kernel / bs_kernel_c: FFT_M(h) / M, h[t mod M] = exp(+i pi ((k_lo - t)^2 mod 2N) / N)
     for t in [-(K-1), L), L = min(N, M - K + 1), k_lo / K the row band (bs_rlo, bs_rk) or column
     band (bs_clo, bs_ck); with a row band shorter than N + bs_rk - 1 every frame's cloud-in-cell
     nodes (y_max + 1) must stay below bs_M - bs_rk + 1 (checked per frame: status[1]); a shorter
     column band follows the forward path's rule (rows z < Z covered, fill_bkg or no smoothing) */
  const void *bs_kernel_adj;
  const void *bs_kernel_adj_c;
} gx_adjoint;

/* Deposit exponents for the current f_table, inverse permutation, zero g_atom. */
int gx_adj_prepare(const gx_frames *d, const gx_adjoint *g, gx_stream_t stream);
/* Cloud-in-cell deposit + FFT along y of frames [f0, f0+nf) into d->inter (needs
 * gx_frames_stats of the same frames first). */
int gx_adj_rowpass(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream);
/* stats + columns + gx_adj_rowpass + gx_frames_colpass: the variant's forward. */
int gx_adj_forward(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream);
/* Adjoint of the column pass, in place on d->inter: FFT along z of the recomputed
 * column, cotangent 2 g_bins[bin] F on binned pixels, adjoint DFT back to z. */
int gx_adj_colpass(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream);
/* Adjoint DFT along y of every row, gathered at the atoms' four nodes into g_part;
 * also each frame's sum of dL/dy' per row block and argmin of y' (for gx_adj_reduce). */
int gx_adj_rowgrad(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream);
/* Fold g_part of frames [f0, f0+nf) into g_atom (chain rule through
 * y' = x sin phi + y cos phi and the per-frame origin min y'); follows gx_adj_rowgrad
 * of the same frames. */
int gx_adj_reduce(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream);
/* columns + rowpass (recomputed) + colpass adjoint + rowgrad + reduce.  d->ymin_key and
 * d->idx_bounds must hold the statistics of the same frames and atoms (gx_frames_stats,
 * run by gx_adj_forward). */
int gx_adj_backward(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream);
/* gx_adj_backward without the row-pass recompute: d->inter must still hold the row
 * pass that gx_adj_forward wrote for the same frames and atoms (nothing has run on
 * d->inter since).  Saves one row pass when the caller keeps the forward's frames. */
int gx_adj_backward_stored(const gx_frames *d, const gx_adjoint *g, int64_t f0, int64_t nf, gx_stream_t stream);
/* Origin min z term, the avg_voxel_f terms (fill_bkg / smooth), per-element f
 * gradients and the coordinate gradients in the original atom order:
 * g_coords [n_atoms][3], g_f [n_elem][2]. */
int gx_adj_finish(const gx_frames *d, const gx_adjoint *g, double *g_coords, double *g_f, gx_stream_t stream);

/* Adjoint of gx_bins_finalize for the FP64 iq: g_bins = g_iq * fn^2 / count
 * (0 where count = 0), with g_iq = (sum of the n_rep copies of g_fx) / fx_scale[0]
 * (gx_detector_adjoint's fixed-point output).  Same axis / Cromer-Mann arguments
 * as gx_bins_finalize. */
int gx_bins_adjoint(const unsigned long long *g_fx, int64_t n_rep, const double *fx_scale,
                    const int32_t *bin_cnt, int64_t c_n, int64_t c_lo, int64_t q_num, double qa_start,
                    double qa_step, double qa_stop, const double *host_aff, double z_elem,
                    double *g_bins, gx_stream_t stream);
/* Adjoint of gx_detector_accumulate + gx_detector_finalize(flags): g_out [n][n] is
 * dL/d(finalized image); d->det must hold the accumulated image of the forward.
 * Writes dL/d(iq) as fixed point with the power-of-two scale fx_scale[0], spread over
 * n_rep (1..64) copies g_fx [n_rep][cy][cx][cz] whose sum is the result (fewer
 * atomics on one address; integer sums are exact, hence deterministic).
 * work: [2 n^2 + ceil(n^2 / 256)] doubles. */
int gx_detector_adjoint(const gx_detector *d, const double *g_out, int64_t flags, double *work,
                        unsigned long long *g_fx, int64_t n_rep, double *fx_scale, gx_stream_t stream);

/* Transpose map of the orientation gather for a fixed detector geometry (no
 * reference counterpart), as runs: a run is up to 64 consecutive pixels p .. p + len - 1
 * whose gather at one orientation o reads the same voxel v of d's [cy][cx][cz] grid
 * (a detector pixel is usually several times finer than a voxel).  Built in two calls:
 *
 * gx_detector_map_count writes voxel_off [nv + 1], the exclusive prefix sum of the
 * runs per voxel (voxel_off[nv] = the number of runs), and leaves counts zeroed for the
 * fill.  Sizes: counts [nv] int32 scratch, scan_work [ceil(nv / 2048) + 1],
 * nv = cy cx cz < 2^31; n_orient must fit the 32 - b bits above the pixel,
 * b = ceil(log2(n_pix^2)).  d->iq and d->det are not used.
 *
 * gx_detector_map_fill then writes, for every voxel v, its runs in
 * [voxel_off[v], voxel_off[v + 1]) (order within a voxel unspecified): runs = (o << b) | p
 * of the run's first pixel, run_voxel = v, run_len = len.  The three arrays hold
 * `capacity` entries (>= voxel_off[nv]); positions beyond it are not written. */
int gx_detector_map_count(const gx_detector *d, int32_t *counts, int64_t *scan_work, int64_t *voxel_off,
                          gx_stream_t stream);
int gx_detector_map_fill(const gx_detector *d, const int64_t *voxel_off, int32_t *counts, uint32_t *runs,
                         int32_t *run_voxel, uint8_t *run_len, int64_t capacity, gx_stream_t stream);
/* gx_detector_adjoint through the map (n_runs = voxel_off[nv] of the count): each run's
 * per-pixel fixed-point terms are summed in a lane (from a table of their prefix sums
 * for orientations of the first orientation's weight), runs of equal voxel across a
 * wave, then one integer atomic per voxel segment; writes one copy of g_fx (n_rep = 1
 * for gx_bins_adjoint), bit for bit the sums of gx_detector_adjoint.
 * work: [3 n^2 + ceil(n^2 / 256)] doubles. */
int gx_detector_adjoint_mapped(const gx_detector *d, const double *g_out, int64_t flags, double *work,
                               const int32_t *run_voxel, const uint32_t *runs, const uint8_t *run_len, int64_t n_runs,
                               unsigned long long *g_fx, double *fx_scale, gx_stream_t stream);

/* ------------------------------------------------------------------------------------
 * Workspace sizes (no reference counterpart: the reference allocates inside each call;
 * for callers that allocate the device memory themselves, e.g. a C, Go or Rust binding
 * or a loop under graph capture).  The library allocates nothing on the paths above:
 * every device buffer comes from the caller through the descriptors.  Fill a query with
 * the geometry; gx_workspace_size writes the byte size of each buffer, named after the
 * descriptor field that points to it (0: not needed for this geometry, the field may
 * be NULL).  Per-frame buffers hold `batch` frames (a caller that keeps the statistics
 * of all frames, as the package's engine does, sizes those with batch = n_frames and
 * offsets the descriptor pointers per batch).
 * ---------------------------------------------------------------------------------- */
typedef struct gx_workspace_query {
  int64_t N;               /* real-space grid edge (16..2048) */
  int64_t n_atoms;
  int64_t n_elem;          /* entries of f_table */
  int64_t n_frames;        /* phi-frames (fparams) */
  int64_t batch;           /* frames per call (gx_frames.batch) */
  int64_t W, Z;            /* intermediate width and row range (gx_frames.W, .Z) */
  int64_t c_n;             /* cropped grid edge (bins [c][c][c]) */
  int64_t i_lo, i_hi;      /* kept rows of the column transforms (gx_frames.i_lo, .i_hi) */
  int64_t fill_bkg, smooth; /* as gx_frames (smooth: sigma in voxels) */
  int64_t fft_f64;         /* 0: float2 transforms, 1: double2 */
  int64_t bs_M, bs_Mc;     /* chirp-z transform lengths (0 for power-of-two N) */
  int64_t store_band;      /* 1: gx_frames.f_band is kept (differentiable path) */
  int64_t differentiable;  /* 1: also the slab's optional outputs and gx_adjoint's buffers */
  int64_t n_pix;           /* detector pixels per edge (0: no detector) */
  int64_t n_init, n_orient; /* detector initial rotations (0..3) and orientations */
} gx_workspace_query;

typedef struct gx_workspace_sizes {
  /* gx_slab outputs */
  int64_t sx, sy, sq, szrow, row_off, row_exp, work;
  int64_t szf, sperm, selem, stable_work;           /* differentiable only */
  /* gx_frames */
  int64_t fparams, twiddle, twiddle64, gauss_cw;
  int64_t bs_chirp, bs_kernel, bs_twiddle, bs_kernel_c, bs_twiddle_c;
  int64_t ymin_key, idx_bounds, col_nm, col_sy, col_sz, inter, col_spec, f_band, avg_q;
  int64_t bin_sum, bin_cnt, status;
  int64_t iq;              /* gx_bins_finalize's FP64 I(q) [c][c][c] (the detector's input) */
  /* gx_detector and gx_detector_finalize */
  int64_t h_axis, v_axis, init_R, orient_R, orient_w, det, det_out;
  /* gx_adjoint (differentiable only) */
  int64_t adj_row_exp, ainv, g_bins, g_part, red, amin, g_atom, avg_spec, avg_part, g_avg;
  int64_t e_part, e_sum, bs_kernel_adj, bs_kernel_adj_c;
  int64_t g_coords, g_f;   /* gx_adj_finish outputs */
  /* detector adjoint (differentiable with a detector): work of gx_detector_adjoint_mapped
     (that of gx_detector_adjoint is smaller), one copy of g_fx, fx_scale, and the
     transpose map's count step; its run arrays take map_run_bytes per run for the
     voxel_off[nv] runs the count reports, at most map_runs_max */
  int64_t det_work, g_fx, fx_scale, map_counts, map_scan_work, map_voxel_off;
  int64_t map_run_bytes, map_runs_max;
  int64_t total;           /* sum of every size above except map_run_bytes and map_runs_max */
} gx_workspace_sizes;

/* Byte sizes for the query's geometry; GX_ERR_ARG for an invalid query (message says
 * which field).  Host-only arithmetic: no device call. */
int gx_workspace_size(const gx_workspace_query *q, gx_workspace_sizes *out);

#ifdef __cplusplus
}
#endif

#endif /* GIWAXSIM_AMD_H */
