"""gx_workspace_size, host-only (no device call): the ctypes structures follow the
header field by field, the total adds the buffer sizes, optional buffers appear exactly
for the geometries that use them, per-frame and per-atom buffers scale with the batch and
the atom count, and invalid queries are refused with the field named."""

import os
import re

import pytest

from giwaxsim_amd import _lib

HEADER = os.path.join(os.path.dirname(__file__), "..", "include", "giwaxsim_amd.h")

# the polymer slab of 100^3 A (110k atoms), N = 1024, fill + smooth 10, 581 frames in
# batches of 213, 512^2 detector x 361 orientations
POLY = dict(N=1024, n_atoms=110000, n_elem=6, n_frames=581, batch=213, W=186, Z=512, c_n=131, i_lo=400,
            i_hi=620, fill_bkg=1, smooth=10, n_pix=512, n_init=1, n_orient=361)
ADJ_ONLY = ("szf", "sperm", "selem", "stable_work", "adj_row_exp", "ainv", "g_bins", "g_part", "red", "amin",
            "g_atom", "avg_spec", "avg_part", "g_avg", "e_part", "e_sum", "g_coords", "g_f", "det_work", "g_fx",
            "fx_scale", "map_counts", "map_scan_work", "map_voxel_off", "map_run_bytes", "map_runs_max")
CHIRP = ("bs_chirp", "bs_kernel", "bs_twiddle", "bs_kernel_c", "bs_twiddle_c")


def _header_fields(name):
    text = open(HEADER).read()
    body = re.search(r"typedef struct " + name + r" \{(.*?)\} " + name + ";", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if decl:
            assert decl.startswith("int64_t "), decl
            fields += [f.strip() for f in decl[len("int64_t "):].split(",")]
    return tuple(fields)


def test_ctypes_structures_follow_the_header():
    assert _header_fields("gx_workspace_query") == _lib.WORKSPACE_QUERY_FIELDS
    assert _header_fields("gx_workspace_sizes") == _lib.WORKSPACE_SIZE_FIELDS


def _total_ok(s):
    return s["total"] == sum(v for k, v in s.items() if k not in ("total", "map_run_bytes", "map_runs_max"))


def test_forward_and_differentiable_sets():
    fwd = _lib.workspace_size(**POLY)
    assert _total_ok(fwd)
    assert all(fwd[k] == 0 for k in ADJ_ONLY + CHIRP + ("f_band",))
    assert all(fwd[k] > 0 for k in ("sx", "inter", "col_nm", "col_sy", "col_spec", "gauss_cw", "bin_sum", "iq",
                                     "det", "orient_R"))
    adj = _lib.workspace_size(**POLY, differentiable=1, store_band=1)
    assert _total_ok(adj)
    assert all(adj[k] > 0 for k in ADJ_ONLY + ("f_band",))
    assert all(adj[k] == fwd[k] for k in ("sx", "inter", "bin_sum", "det"))
    assert adj["map_run_bytes"] == 9 and adj["map_runs_max"] == 512 * 512 * 361
    # more elements than the gather sums in registers: dL/df per atom, no e_part / e_sum
    many = _lib.workspace_size(**{**POLY, "n_elem": _lib.GX_ADJ_ELEM_REG + 1}, differentiable=1)
    assert many["e_part"] == 0 and many["e_sum"] == 0 and many["g_part"] == adj["g_part"]
    # no fill, no smoothing: no column masks, spectra or kernel prefix sums
    plain = _lib.workspace_size(**{**POLY, "fill_bkg": 0, "smooth": 0}, differentiable=1)
    assert all(plain[k] == 0 for k in ("col_nm", "col_sy", "col_sz", "col_spec", "gauss_cw", "avg_spec",
                                       "avg_part"))
    # no detector
    nodet = _lib.workspace_size(**{**POLY, "n_pix": 0}, differentiable=1)
    assert all(nodet[k] == 0 for k in ("det", "det_out", "h_axis", "orient_R", "det_work", "g_fx", "map_counts"))


def test_scaling_with_batch_atoms_and_precision():
    a = _lib.workspace_size(**POLY, differentiable=1, store_band=1)
    b = _lib.workspace_size(**{**POLY, "batch": 2 * POLY["batch"]}, differentiable=1, store_band=1)
    per_frame = ("ymin_key", "idx_bounds", "col_nm", "col_sy", "col_sz", "inter", "col_spec", "f_band", "g_part",
                 "avg_spec", "avg_part", "e_part")
    assert all(b[k] == 2 * a[k] for k in per_frame)
    assert all(b[k] == a[k] for k in ("sx", "fparams", "bin_sum", "det", "g_atom"))
    c = _lib.workspace_size(**{**POLY, "n_atoms": 2 * POLY["n_atoms"]}, differentiable=1, store_band=1)
    assert all(c[k] == 2 * a[k] for k in ("sx", "sy", "sq", "szrow", "szf", "sperm", "ainv", "g_part", "g_atom",
                                          "g_coords"))
    assert all(c[k] == a[k] for k in ("inter", "bin_sum", "row_off"))
    d = _lib.workspace_size(**POLY, differentiable=1, store_band=1, fft_f64=1)
    assert all(d[k] == 2 * a[k] for k in ("inter", "col_spec", "f_band", "twiddle", "g_part"))
    assert all(d[k] == a[k] for k in ("twiddle64", "bin_sum", "avg_spec"))


def test_chirpz_geometry():
    q = dict(POLY, N=1048, W=200, Z=512, bs_M=2048, bs_Mc=2048)
    s = _lib.workspace_size(**q, differentiable=1)
    assert all(s[k] > 0 for k in CHIRP + ("bs_kernel_adj", "bs_kernel_adj_c"))
    assert s["bs_kernel"] == s["bs_kernel_adj"] == 2048 * 8 and s["bs_chirp"] == 1048 * 8
    p2 = _lib.workspace_size(**POLY, differentiable=1)
    assert all(p2[k] == 0 for k in CHIRP + ("bs_kernel_adj", "bs_kernel_adj_c"))


@pytest.mark.parametrize("change,text", [
    (dict(N=4096), "N=4096"),
    (dict(N=1048), "bs_M"),
    (dict(bs_M=2048, bs_Mc=2048), "bs_M"),
    (dict(Z=12), "Z=12"),
    (dict(W=2048), "W=2048"),
    (dict(batch=0), "batch=0"),
    (dict(n_elem=0), "n_elem=0"),
    (dict(c_n=0), "c_n=0"),
    (dict(store_band=1, i_lo=700), "store_band"),
    (dict(n_init=4), "n_init=4"),
    (dict(n_orient=0), "n_orient=0"),
])
def test_invalid_queries_are_refused(change, text):
    with pytest.raises(_lib.GxError, match=text):
        _lib.workspace_size(**{**POLY, **change})
    with pytest.raises(ValueError, match="unknown"):
        _lib.workspace_size(**POLY, n_atom=3)
