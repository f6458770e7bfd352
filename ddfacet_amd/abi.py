"""ctypes mirror of include/ddfgpu.h.

The structs describe one gridding / degridding call in the layout of the reference argument
lists of ``pyGridderWPol`` / ``pyDeGridderWPol`` (DDFacet/Gridder/GridderSmearPols.cc:76-97,
212-233).  They carry raw addresses only, so the same description is used with device
addresses (the HIP library) and with host addresses (the CPU oracle used by the tests).
"""
import ctypes as C

DDF_ABI_VERSION = 1

DDF_OK = 0
DDF_EINVAL = -1
DDF_ERUNTIME = -2
DDF_EHIP = -3
DDF_ENOMEM = -4

P = C.c_void_p


class ddf_cf(C.Structure):
    _fields_ = [
        ("data", P),
        ("offset", P),
        ("support", P),
        ("nw", C.c_int32),
        ("overs", C.c_int32),
        ("ref_wave", C.c_double),
        ("wmax", C.c_double),
        ("max_support", C.c_int32),
    ]


class ddf_jones(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32),
        ("mode", C.c_int32),
        ("interp_mode", C.c_int32),
        ("kms", P),
        ("kms_dims", C.c_int32 * 4),
        ("kms_time", P),
        ("kms_chan", P),
        ("kms_dir", C.c_int32),
        ("beam", P),
        ("beam_dims", C.c_int32 * 4),
        ("beam_time", P),
        ("beam_chan", P),
        ("beam_dir", C.c_int32),
        ("a0", P),
        ("a1", P),
        ("coefs", P),
        ("ncoefs", C.c_int32),
        ("apply_amp", C.c_int32),
        ("apply_phase", C.c_int32),
        ("scale_jones", C.c_int32),
        ("calib_error", C.c_double),
        ("sum_jones", P),
        ("sum_jones_chan", P),
        ("reweight_snr", C.c_double),
    ]


class ddf_decorr(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32),
        ("uvw_dt", P),
        ("dt", C.c_double),
        ("dnu", C.c_double),
        ("tsmear", C.c_int32),
        ("fsmear", C.c_int32),
        ("l0", C.c_double),
        ("m0", C.c_double),
    ]


class ddf_vis(C.Structure):
    _fields_ = [
        ("vis", P),
        ("uvw", P),
        ("flags", P),
        ("weights", P),
        ("freqs", P),
        ("chan_map", P),
        ("nrow", C.c_int64),
        ("nchan", C.c_int32),
        ("ncorr", C.c_int32),
        ("chan_equidistant", C.c_int32),
        ("bda", P),
        ("bda_len", C.c_int64),
        ("nblocks", C.c_int64),
        ("bda_start", P),
        ("sparsification", P),
        ("corr_codes", C.c_uint16 * 4),
        ("ncorr_codes", C.c_int32),
        ("stokes_codes", C.c_uint16 * 4),
        ("nstokes", C.c_int32),
        ("freq_mean", C.c_double),
        ("wavelength_mean", C.c_double),
    ]


class ddf_facet(C.Structure):
    _fields_ = [
        ("grid", P),
        ("sumwt", P),
        ("cf", ddf_cf),
        ("cu", C.c_double),
        ("cv", C.c_double),
        ("l0", C.c_double),
        ("m0", C.c_double),
        ("jones", ddf_jones),
        ("decorr", ddf_decorr),
    ]


class ddf_grid_desc(C.Structure):
    _fields_ = [
        ("grid_f64", C.c_int32),
        ("nband", C.c_int32),
        ("npol", C.c_int32),
        ("ny", C.c_int32),
        ("nx", C.c_int32),
        ("incr", C.c_double * 2),
        ("dopsf", C.c_int32),
    ]


STOKES_CODES = {"undef": 0, "I": 1, "Q": 2, "U": 3, "V": 4, "RR": 5, "RL": 6, "LR": 7, "LL": 8,
                "XX": 9, "XY": 10, "YX": 11, "YY": 12}


class ddf_plan_opts(C.Structure):
    _fields_ = [
        ("tile", C.c_int32),
        ("chunk", C.c_int32),
        ("blocks_per_item", C.c_int32),
        ("group", C.c_int32),
        ("precision", C.c_int32),
        ("workgroups", C.c_int32),
    ]


class ddf_image_facet(C.Structure):
    _fields_ = [
        ("xc", C.c_int32),
        ("yc", C.c_int32),
        ("sphe", P),
        ("inv_sphe", P),
        ("sw", P),
        ("dirty", P),
        ("model", P),
        ("sumwt", P),
        ("sum_jones", P),
    ]


class ddf_cf_spec(C.Structure):
    _fields_ = [
        ("nw", C.c_int32),
        ("overs", C.c_int32),
        ("nfacets", C.c_int32),
        ("pad_", C.c_int32),
        ("sups", P),
        ("wlist", P),
        ("coefs", P),
        ("tapers", P),
        ("ref_wave", C.c_double),
        ("lrad", C.c_double),
    ]


PRECISION = {"mixed": 0, "fp64": 1, "fp32": 2}


class ddf_flag_select(C.Structure):
    _fields_ = [("na", C.c_int32), ("use_uv_range", C.c_int32), ("use_time_range", C.c_int32), ("pad", C.c_int32),
                ("uv_min_m2", C.c_double), ("uv_max_m2", C.c_double), ("time_min", C.c_double),
                ("time_max", C.c_double), ("time_offset", C.c_double), ("ant_threshold", C.c_double),
                ("named_antennas", C.POINTER(C.c_uint8))]


class ddf_bda_spec(C.Structure):
    _fields_ = [("mode", C.c_int32), ("nchan", C.c_int32), ("na", C.c_int32), ("per_row", C.c_int32),
                ("duv_cut", C.c_double), ("k_dnu", C.c_double), ("l", C.c_double), ("dfreq", C.c_double),
                ("k_phase", C.c_double), ("lmax", C.c_double), ("nmax", C.c_double), ("k_span", C.c_double),
                ("dphi", C.c_double)]

class ddf_beam_desc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nchan", C.c_int32), ("ncoef", C.c_int32), ("nchan_jones", C.c_int32),
                ("freqs", C.c_void_p), ("coef0", C.c_void_p), ("coef1", C.c_void_p), ("band_f0", C.c_void_p),
                ("x_null", C.c_void_p), ("y_null", C.c_void_p), ("chan_to_jones", C.c_void_p)]


class ddf_jones_merge_desc(C.Structure):
    _fields_ = [("j0", C.c_void_p), ("j1", C.c_void_p), ("out", C.c_void_p), ("it0", C.c_void_p),
                ("it1", C.c_void_p), ("ich0", C.c_void_p), ("ich1", C.c_void_p), ("nt0", C.c_int32),
                ("nch0", C.c_int32), ("nt1", C.c_int32), ("nch1", C.c_int32), ("nd", C.c_int32), ("na", C.c_int32),
                ("nt", C.c_int32), ("nch", C.c_int32)]


EXTRA_SIGNATURES = {
    "ddf_beam_jones": [C.POINTER(ddf_beam_desc), C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_double,
                       C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p],
    "ddf_jones_merge": [C.POINTER(ddf_jones_merge_desc), C.c_void_p],
    "ddf_weights_index": [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_int64,
                          C.c_int64, C.c_int, C.c_int64, C.c_void_p, C.c_void_p],
    "ddf_prepare_flags": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_void_p],
    "ddf_weights_accumulate": [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p],
    "ddf_weights_briggs": [C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_void_p, C.c_int64, C.c_void_p],
    "ddf_weights_finalize": [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p],
    "ddf_weights_briggs_noalloc": [C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_void_p],
    "ddf_plan_create": [C.POINTER(ddf_grid_desc), C.POINTER(ddf_vis), C.POINTER(ddf_vis), C.POINTER(ddf_facet),
                        C.POINTER(C.c_void_p), C.c_int32, C.POINTER(ddf_plan_opts), C.POINTER(C.c_void_p),
                        C.c_void_p],
    "ddf_plan_grid": [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p],
    "ddf_plan_degrid": [C.c_void_p, C.c_void_p, C.c_void_p],
    "ddf_plan_info": [C.c_void_p, C.POINTER(C.c_int64), C.c_int32],
    "ddf_plan_placements": [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p],
    "ddf_bda_build_create": [C.POINTER(ddf_bda_spec), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                             C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int64), C.c_void_p],
    "ddf_bda_build_write": [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p],
    "ddf_bda_build_destroy": [C.c_void_p],
    "ddf_sort_baseline_time": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p],
    "ddf_gather_rows": [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p],
    "ddf_flag_selection": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32,
                           C.POINTER(ddf_flag_select), C.c_void_p, C.c_void_p, C.c_void_p],
    "ddf_cf_build": [C.c_void_p, C.c_void_p, C.POINTER(C.c_int64), C.c_void_p],
    "ddf_plan_timing": [C.c_void_p, C.c_int32],
    "ddf_plan_pass_ms": [C.c_void_p, C.POINTER(C.c_float), C.c_int32],
    "ddf_plan_destroy": [C.c_void_p],
    "ddf_plan_degrid_adjoint": [C.c_void_p, C.c_void_p, C.c_void_p],
    "ddf_model_to_grid_adjoint": [C.POINTER(ddf_image_facet), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p],
    "ddf_fft2c": [C.c_void_p, C.c_int32, C.c_int64, C.c_int32, C.c_double, C.c_void_p],
    "ddf_facet_norm": [C.POINTER(ddf_image_facet), C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p],
    "ddf_stitch": [C.POINTER(ddf_image_facet), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                   C.c_void_p, C.c_void_p],
    "ddf_model_to_grid": [C.POINTER(ddf_image_facet), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                          C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p],
}

# pass order of ddf_plan_pass_ms (DDF_PASS_* in include/ddfgpu.h)
PASS_NAMES = ("grid_pass_a", "grid_pass_b", "degrid_pass_a", "degrid_pass_b", "degrid_adj_b", "degrid_adj_a")
