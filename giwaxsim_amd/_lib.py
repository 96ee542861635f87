"""ctypes binding of the C ABI declared in ``include/giwaxsim_amd.h``.

The library is ``giwaxsim_amd/lib/libgiwaxsim_amd.so`` (built by
``giwaxsim_amd/build.py``).  There is no fallback: if the library is missing
or fails to load, every entry point raises.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int64, c_void_p

from .build import LIB

GX_OK = 0
GX_ABI_VERSION = 12  # structure layouts below match this version of the header
GX_FP_STRIDE = 24
GX_MOMENTS_WORK = 774
GX_CROP_NOALLOC_MAX_ATOMS = 134217728
GX_FIN_MIRROR = 1
GX_FIN_FLOOR_SCALE = 2

# frame record offsets (enum in include/giwaxsim_amd.h)
(FP_SIN, FP_COS, FP_HX_START, FP_HX_STEP, FP_HX_STOP, FP_HY_START, FP_HY_STEP, FP_HY_STOP,
 FP_LEN_MODE, FP_HOR, FP_VER, FP_STOP1, FP_STOP2, FP_S12, FP_MID, FP_A, FP_TP, FP_VC, FP_TT,
 FP_CP, FP_CT, FP_JLO, FP_JHI, FP_UNUSED) = range(24)


class GxError(RuntimeError):
    """A C-ABI call returned a non-zero status."""


class GxSlab(ctypes.Structure):
    _fields_ = [
        ("n_atoms", c_int64), ("n_elem", c_int64), ("N", c_int64), ("r", c_double),
        ("zmin", c_double),
        ("x", c_void_p), ("y", c_void_p), ("z", c_void_p), ("elem", c_void_p),
        ("f_table", c_void_p),
        ("sx", c_void_p), ("sy", c_void_p), ("sq", c_void_p), ("szrow", c_void_p),
        ("row_off", c_void_p), ("row_exp", c_void_p), ("work", c_void_p),
        ("szf", c_void_p), ("sperm", c_void_p), ("selem", c_void_p), ("stable_work", c_void_p),
    ]


class GxFrames(ctypes.Structure):
    _fields_ = [
        ("N", c_int64), ("r", c_double), ("n_atoms", c_int64),
        ("ax", c_void_p), ("ay", c_void_p), ("aq", c_void_p), ("azrow", c_void_p),
        ("row_off", c_void_p), ("row_exp", c_void_p),
        ("fill_bkg", c_int64), ("smooth", c_int64), ("max_voxels", c_double),
        ("avg_re", c_double), ("avg_im", c_double), ("gauss_cw", c_void_p),
        ("smooth_radius", c_int64), ("zrow_max", c_int64),
        ("qmin", c_double), ("dq", c_double), ("c_lo", c_int64), ("c_n", c_int64),
        ("q_val", c_double), ("i_lo", c_int64), ("i_hi", c_int64), ("W", c_int64),
        ("Z", c_int64),
        ("bin_scale", c_double), ("fft_f64", c_int64), ("twiddle", c_void_p),
        ("twiddle64", c_void_p), ("bs_M", c_int64), ("bs_chirp", c_void_p),
        ("bs_kernel", c_void_p), ("bs_twiddle", c_void_p),
        ("n_frames", c_int64),
        ("fparams", c_void_p), ("batch", c_int64),
        ("ymin_key", c_void_p), ("idx_bounds", c_void_p), ("col_nm", c_void_p),
        ("col_sy", c_void_p), ("col_sz", c_void_p), ("inter", c_void_p), ("col_spec", c_void_p),
        ("bin_sum", c_void_p), ("bin_cnt", c_void_p), ("status", c_void_p),
        ("row_variant", c_int64),
        ("bs_rlo", c_int64), ("bs_rk", c_int64), ("bs_Mc", c_int64), ("bs_clo", c_int64), ("bs_ck", c_int64),
        ("bs_kernel_c", c_void_p), ("bs_twiddle_c", c_void_p),
        ("avg_q", c_void_p), ("inter_rm", c_int64), ("fft_plan_row", c_int64), ("fft_plan_col", c_int64),
        ("f_band", c_void_p),
    ]


class GxDetector(ctypes.Structure):
    _fields_ = [
        ("n_pix", c_int64), ("n_init", c_int64), ("n_orient", c_int64),
        ("h_axis", c_void_p), ("v_axis", c_void_p), ("init_R", c_void_p),
        ("orient_R", c_void_p), ("orient_w", c_void_p), ("iq", c_void_p), ("iq_f64", c_int64),
        ("cy", c_int64), ("cx", c_int64), ("cz", c_int64),
        ("qx_min", c_double), ("qy_min", c_double), ("qz_min", c_double), ("dq", c_double),
        ("det", c_void_p),
    ]


GX_ADJ_RED_BLOCKS = 16
GX_ADJ_ELEM_REG = 6
GX_STABLE_CHUNK = 4096


class GxAdjoint(ctypes.Structure):
    _fields_ = [
        ("n_elem", c_int64), ("f_table", c_void_p), ("afz", c_void_p), ("aperm", c_void_p),
        ("aelem", c_void_p), ("ecnt", c_void_p), ("z", c_void_p), ("zmin", c_double),
        ("row_exp", c_void_p), ("ainv", c_void_p),
        ("g_bins", c_void_p), ("g_part", c_void_p), ("red", c_void_p), ("amin", c_void_p),
        ("g_atom", c_void_p),
        ("x", c_void_p), ("y", c_void_p), ("ext", c_void_p), ("ecount", c_void_p),
        ("vol_r3", c_double), ("avg_spec", c_void_p), ("avg_part", c_void_p), ("g_avg", c_void_p),
        ("e_part", c_void_p), ("e_sum", c_void_p), ("bs_kernel_adj", c_void_p), ("bs_kernel_adj_c", c_void_p),
    ]


# gx_workspace_query / gx_workspace_sizes: every field int64, in the header's order
WORKSPACE_QUERY_FIELDS = (
    "N", "n_atoms", "n_elem", "n_frames", "batch", "W", "Z", "c_n", "i_lo", "i_hi", "fill_bkg", "smooth",
    "fft_f64", "bs_M", "bs_Mc", "store_band", "differentiable", "n_pix", "n_init", "n_orient")
WORKSPACE_SIZE_FIELDS = (
    "sx", "sy", "sq", "szrow", "row_off", "row_exp", "work", "szf", "sperm", "selem", "stable_work",
    "fparams", "twiddle", "twiddle64", "gauss_cw", "bs_chirp", "bs_kernel", "bs_twiddle", "bs_kernel_c",
    "bs_twiddle_c", "ymin_key", "idx_bounds", "col_nm", "col_sy", "col_sz", "inter", "col_spec", "f_band",
    "avg_q", "bin_sum", "bin_cnt", "status", "iq", "h_axis", "v_axis", "init_R", "orient_R", "orient_w", "det",
    "det_out", "adj_row_exp", "ainv", "g_bins", "g_part", "red", "amin", "g_atom", "avg_spec", "avg_part",
    "g_avg", "e_part", "e_sum", "bs_kernel_adj", "bs_kernel_adj_c", "g_coords", "g_f", "det_work", "g_fx",
    "fx_scale", "map_counts", "map_scan_work", "map_voxel_off", "map_run_bytes", "map_runs_max", "total")


class GxWorkspaceQuery(ctypes.Structure):
    _fields_ = [(f, c_int64) for f in WORKSPACE_QUERY_FIELDS]


class GxWorkspaceSizes(ctypes.Structure):
    _fields_ = [(f, c_int64) for f in WORKSPACE_SIZE_FIELDS]


_P = c_void_p
_D = c_double
_I = c_int64
# name -> argtypes (all return int status except the first three)
SIGNATURES = {
    "gx_slab_bounds": [POINTER(GxSlab), _P, _P, _P],
    "gx_slab_sort_rows": [POINTER(GxSlab), _P],
    "gx_slab_avg": [POINTER(GxSlab), _P, _D, _D, _P, _P],
    "gx_frames_stats": [POINTER(GxFrames), _I, _I, _P],
    "gx_frames_columns": [POINTER(GxFrames), _I, _I, _P],
    "gx_frames_rowpass": [POINTER(GxFrames), _I, _I, _P],
    "gx_frames_colpass": [POINTER(GxFrames), _I, _I, _P],
    "gx_frames_forward": [POINTER(GxFrames), _I, _I, _P],
    "gx_bins_clear": [_P, _P, _I, _P],
    "gx_bins_finalize": [_P, _P, _I, _D, _I, _I, _D, _D, _D, POINTER(c_double), _D, _P, _P, _P],
    "gx_detector_accumulate": [POINTER(GxDetector), _P],
    "gx_detector_finalize": [_P, _I, _I, _I, _P, _P],
    "gx_rotate_points": [POINTER(c_double), _P, _P, _P, _I, _P, _P, _P, _P],
    "gx_intersect": [_P, _I, _I, _I, _D, _D, _D, _D, _P, _P, _P, _I, _D, _I, _P, _P],
    "gx_bin_frame_f64": [_P, _I, _I, _P, _P, _P, POINTER(c_double), _D, _I, _P, _P, _P, _I, _P],
    "gx_collapse_lengths": [POINTER(c_double), _P, _I, _P, _P],
    "gx_scale_offset_moments": [_P, _P, _P, _I, _P, POINTER(c_double), _P],
    "gx_slab_tile": [_P, _P, _I, _I, _I, _I, POINTER(c_double), _P, _P, _P, _P, _P],
    "gx_slab_crop": [_P, _P, _P, _P, _I, POINTER(c_double), _P, _P, _P, _P, _P, POINTER(c_int64), _P],
    "gx_slab_crop_noalloc": [_P, _P, _P, _P, _I, POINTER(c_double), _P, _P, _P, _P, POINTER(c_int64), _P],
    "gx_shift_min": [_P, _P, _P, _I, _P, _P],
    "gx_map_linear": [_P, _I, _I, _P, _P, _I, _D, _P, _P],
    "gx_polar_pad": [_P, _I, _I, _P, _I, _P, _P],
    "gx_csr_apply": [_P, _P, _P, _I, _P, _D, _P, _P],
    "gx_fit_residual": [_P, _P, _P, _I, _D, _D, _P, _P],
    "gx_apply_f0": [_P, _I, _I, _I, _P, _P, _P, POINTER(c_double), _D, _P, _P],
    "gx_bins_accumulate_f64": [_P, _P, _I, _D, _P, _P, _P],
    "gx_adj_prepare": [POINTER(GxFrames), POINTER(GxAdjoint), _P],
    "gx_adj_rowpass": [POINTER(GxFrames), POINTER(GxAdjoint), _I, _I, _P],
    "gx_adj_forward": [POINTER(GxFrames), POINTER(GxAdjoint), _I, _I, _P],
    "gx_adj_colpass": [POINTER(GxFrames), POINTER(GxAdjoint), _I, _I, _P],
    "gx_adj_rowgrad": [POINTER(GxFrames), POINTER(GxAdjoint), _I, _I, _P],
    "gx_adj_reduce": [POINTER(GxFrames), POINTER(GxAdjoint), _I, _I, _P],
    "gx_adj_backward": [POINTER(GxFrames), POINTER(GxAdjoint), _I, _I, _P],
    "gx_adj_backward_stored": [POINTER(GxFrames), POINTER(GxAdjoint), _I, _I, _P],
    "gx_adj_finish": [POINTER(GxFrames), POINTER(GxAdjoint), _P, _P, _P],
    "gx_bins_adjoint": [_P, _I, _P, _P, _I, _I, _I, _D, _D, _D, POINTER(c_double), _D, _P, _P],
    "gx_detector_adjoint": [POINTER(GxDetector), _P, _I, _P, _P, _I, _P, _P],
    "gx_detector_map_count": [POINTER(GxDetector), _P, _P, _P, _P],
    "gx_detector_map_fill": [POINTER(GxDetector), _P, _P, _P, _P, _P, _I, _P],
    "gx_detector_adjoint_mapped": [POINTER(GxDetector), _P, _I, _P, _P, _P, _P, _I, _P, _P, _P],
    "gx_workspace_size": [POINTER(GxWorkspaceQuery), POINTER(GxWorkspaceSizes)],
}
EXPORTS = ["gx_last_error", "gx_abi_version", "gx_supported_grid"] + list(SIGNATURES)

_lib = None


def load(path: str | None = None):
    """Load (once) and return the ctypes library with argtypes applied."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB
    if not os.path.exists(path):
        raise GxError(f"{path} not found: build it with `python -m giwaxsim_amd.build` "
                      "(or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    lib.gx_last_error.restype = c_char_p
    lib.gx_last_error.argtypes = []
    lib.gx_abi_version.restype = c_int
    lib.gx_abi_version.argtypes = []
    lib.gx_supported_grid.restype = c_int
    lib.gx_supported_grid.argtypes = [c_int64]
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = c_int
        fn.argtypes = args
    if lib.gx_abi_version() != GX_ABI_VERSION:
        raise GxError(f"{path} has ABI version {lib.gx_abi_version()}, the binding expects "
                      f"{GX_ABI_VERSION}: rebuild it")
    _lib = lib
    return lib


def call(name: str, *args):
    """Call an entry point and raise GxError with the library's message on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != GX_OK:
        msg = lib.gx_last_error().decode(errors="replace")
        raise GxError(f"{name} failed ({rc}): {msg}")
    return rc


def workspace_size(**query):
    """gx_workspace_size for a geometry given as keyword arguments (the field names of
    gx_workspace_query; missing fields are 0): {buffer name: bytes}.  Host-only: needs no
    device."""
    unknown = sorted(set(query) - set(WORKSPACE_QUERY_FIELDS))
    if unknown:
        raise ValueError(f"unknown gx_workspace_query fields {unknown}")
    q = GxWorkspaceQuery(**{k: int(v) for k, v in query.items()})
    out = GxWorkspaceSizes()
    call("gx_workspace_size", ctypes.byref(q), ctypes.byref(out))
    return {f: int(getattr(out, f)) for f in WORKSPACE_SIZE_FIELDS}


def dbl_array(values):
    arr = (c_double * len(values))(*[float(v) for v in values])
    return arr
