"""ctypes bindings of the native libraries (``lib/libgvm_hip.so``, ``lib/libgvm.so``).

The product path is native: these bindings only marshal arguments.  Loading
fails loudly when a library has not been built -- there is no Python or CPU
fallback for any device computation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"

GVM_OK = 0
GVM_ERR_ARG = 1
# include/gvm_hip.h: largest n of gvm_vec_dot_dev_noalloc
GVM_NOALLOC_DOT_MAX_N = 1 << 37
GVM_BEAM_AIRYDISK = 0
GVM_BEAM_GAUSSIAN = 1
ADJOINT_MODES = {"f32": 0, "f16x3": 1, "bf16x3": 2, "direct": 3}
WEIGHTING_SCHEMES = {"natural": 0, "uniform": 1, "briggs": 2, "radial": 3}


class GvmError(RuntimeError):
    def __init__(self, what: str, status: int, detail: str = ""):
        super().__init__(f"{what} failed with status {status}: {detail}")
        self.status = status


class Geometry(C.Structure):
    _fields_ = [
        ("M", C.c_int), ("N", C.c_int), ("image_count", C.c_int),
        ("deltax", C.c_double), ("deltay", C.c_double),
        ("x0_ref", C.c_float), ("y0_ref", C.c_float),
        ("x0_phs", C.c_float), ("y0_phs", C.c_float),
        ("nu_0", C.c_float), ("fg_scale", C.c_float), ("noise_cut", C.c_float),
        ("minpix", C.c_float), ("eta", C.c_float), ("threshold", C.c_float),
        ("beam_kind", C.c_int), ("antenna_diameter", C.c_float),
        ("pb_factor", C.c_float), ("pb_cutoff", C.c_float),
    ]


class CKernel(C.Structure):
    """gvm_ckernel: m x n taps (host pointer) with supports."""
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("support_x", C.c_int), ("support_y", C.c_int),
                ("taps", C.c_void_p)]


class CKernelParams(C.Structure):
    """gvm_ckernel_params (include/gvm/ckernel.h)."""
    _fields_ = [("kind", C.c_int), ("m", C.c_int), ("n", C.c_int), ("amp", C.c_float),
                ("w", C.c_float), ("w2", C.c_float), ("alpha", C.c_float)]


CKERNEL_KINDS = {"pillbox": 0, "gaussian": 1, "sinc": 2, "gaussian_sinc": 3, "pswf": 4}


class UVTaper(C.Structure):
    _fields_ = [("amplitude", C.c_float), ("u_0", C.c_double), ("v_0", C.c_double),
                ("sigma_maj", C.c_float), ("sigma_min", C.c_float), ("bpa", C.c_float)]


class SimConfig(C.Structure):
    _fields_ = [
        ("npix", C.c_int), ("cell_arcsec", C.c_double), ("n_antennas", C.c_int),
        ("layout", C.c_int), ("array_diameter_m", C.c_double), ("layout_sigma_m", C.c_double),
        ("latitude_deg", C.c_double), ("declination_deg", C.c_double),
        ("ha_start_hours", C.c_double), ("ha_end_hours", C.c_double),
        ("dump_seconds", C.c_double), ("freq_hz", C.c_double),
        ("dish_diameter_m", C.c_double), ("pb_factor", C.c_double),
        ("n_gaussians", C.c_int), ("n_points", C.c_int),
        ("sigma_min_pix", C.c_double), ("sigma_max_pix", C.c_double),
        ("flux_law", C.c_int), ("flux_index", C.c_double),
        ("flux_min", C.c_double), ("flux_max", C.c_double),
        ("point_flux_min", C.c_double), ("point_flux_max", C.c_double),
        ("position_radius_frac", C.c_double),
        ("alpha_min", C.c_double), ("alpha_max", C.c_double),
        ("noise_sigma", C.c_double),
        ("seed_layout", C.c_ulonglong), ("seed_sky", C.c_ulonglong), ("seed_noise", C.c_ulonglong),
        ("nu0_hz", C.c_double), ("flag_fraction", C.c_double), ("seed_flag", C.c_ulonglong),
    ]


class SimComponent(C.Structure):
    _fields_ = [("x_pix", C.c_double), ("y_pix", C.c_double), ("flux", C.c_double),
                ("sigma_pix", C.c_double), ("alpha", C.c_double)]


class FitsHeader(C.Structure):
    _fields_ = [("crpix1", C.c_double), ("crpix2", C.c_double), ("crval1", C.c_double),
                ("crval2", C.c_double), ("cdelt1", C.c_double), ("cdelt2", C.c_double),
                ("crval3", C.c_double), ("cdelt3", C.c_double), ("equinox", C.c_float),
                ("radesys", C.c_char_p), ("bunit", C.c_char_p), ("object", C.c_char_p),
                ("niter", C.c_int)]


class FwTerm(C.Structure):
    _fields_ = [("name", C.c_char_p), ("penalizator", C.c_int), ("image_index", C.c_int),
                ("image_to_add", C.c_int), ("prior", C.c_float), ("normalize", C.c_int)]


class FwConfig(C.Structure):
    _fields_ = [("nterms", C.c_int), ("terms", FwTerm * 8), ("nlambdas", C.c_int),
                ("lambdas", C.c_float * 8), ("initial_values", C.c_float * 2),
                ("eta", C.c_float), ("nopositivity", C.c_int), ("flag_opt", C.c_int)]


class FwResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("exit_reason", C.c_int), ("final_value", C.c_double),
                ("function_calls", C.c_long), ("gradient_calls", C.c_long),
                ("seconds", C.c_double)]


_P = C.c_void_p
_I = C.c_int
_L = C.c_long
_D = C.c_double
_F = C.c_float
_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int)

# C-ABI of include/gvm_hip.h: name -> argtypes (restype int unless noted)
HIP_API = {
    "gvm_create": [C.POINTER(Geometry), _P, C.POINTER(_P)],
    "gvm_destroy": [_P],
    "gvm_add_visibilities": [_P, _F, _L, _P, _P, _P, _IP],
    "gvm_num_slices": [_P, _IP],
    "gvm_get_geometry": [_P, C.POINTER(Geometry)],
    "gvm_set_adjoint": [_P, _I],
    "gvm_chi2": [_P, _P, _I, _I, _DP],
    "gvm_chi2_enqueue": [_P, _P, _I, _I, _I],
    "gvm_get_visibilities_out": [_P, _I, _P, _P, _P],
    "gvm_add_visibilities_dev": [_P, _F, _L, _P, _P, _P, _IP],
    "gvm_grid_visibilities_dev": [_P, C.POINTER(CKernel), _F, _L, _P, _P, _P, _P, _P, _P, _L,
                                  C.POINTER(_L)],
    "gvm_add_visibilities_gridded_dev": [_P, C.POINTER(CKernel), _F, _L, _P, _P, _P,
                                         C.POINTER(_L), _IP],
    "gvm_set_gcf": [_P, _P],
    "gvm_values_fetch_enqueue": [_P, _I],
    "gvm_values_host": [_P, _I, _DP],
    "gvm_capture_begin": [_P],
    "gvm_capture_end": [_P, C.POINTER(_P)],
    "gvm_graph_launch": [_P, _P],
    "gvm_graph_destroy": [_P],
    "gvm_epoch": [_P, C.POINTER(_L)],
    "gvm_nn_conv_fwd": [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I],
    "gvm_nn_conv_bwd": [_P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I],
    "gvm_nn_conv_fwd_bf16": [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I],
    "gvm_nn_conv_bwd_bf16": [_P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I],
    "gvm_nn_to_bf16": [_P, _P, _P, _L],
    "gvm_nn_pool_fwd": [_P, _P, _P, _I, _I, _I, _I],
    "gvm_nn_pool_bwd": [_P, _P, _P, _I, _I, _I, _I],
    "gvm_nn_up_fwd": [_P, _P, _P, _I, _I, _I, _I, _I, _I],
    "gvm_nn_up_bwd": [_P, _P, _P, _I, _I, _I, _I, _I, _I],
    "gvm_nn_copy_channels": [_P, _P, _P, _I, _I, _I, _I, _I, _I, _I],
    "gvm_nn_add": [_P, _P, _P, _L],
    "gvm_nn_mse": [_P, _P, _P, _P, _L, _I],
    "gvm_nn_adam": [_P, _P, _P, _P, _P, _L, _F, _F, _F, _F, _I],
    "gvm_set_split_vis": [_P, _L],
    "gvm_set_fft_mode": [_P, _I],
    "gvm_get_split_vis": [_P, C.POINTER(_L)],
    "gvm_adjoint_computed_pixels": [_P, C.POINTER(_L)],
    "gvm_autotune_adjoint": [_P, _P, _I, C.POINTER(_L), _DP, _IP],
    "gvm_weighting_apply_dev": [_P, _I, _F, C.POINTER(UVTaper), _I, _P, _P, _P, _P],
    "gvm_setup_beam_noise_dev": [_I, _P, _P, _P, _P, _D, _DP, _DP, _DP, _DP, _DP, _DP],
    "gvm_setup_beam_from_moments": [_D, _D, _D, _D, _D, _DP, _DP, _DP, _DP, _DP, _DP],
    "gvm_setup_noise_image_dev": [_I, _D, _D, _D, _D, _I, _I, _D, _D, _F, _F, _D, _D, _P,
                                  C.POINTER(C.c_float), C.POINTER(C.c_float)],
    "gvm_entropy_value_enqueue": [_P, _P, _I, _F, _F, _I],
    "gvm_tsv_value_enqueue": [_P, _P, _I, _I],
    "gvm_laplacian_value_enqueue": [_P, _P, _I, _I],
    "gvm_values_read": [_P, _I, _DP],
    "gvm_dchi2": [_P, _P, _P, _I, _I],
    "gvm_get_residuals": [_P, _I, _P],
    "gvm_entropy_value": [_P, _P, _I, _F, _F, _DP],
    "gvm_entropy_grad": [_P, _P, _P, _I, _F, _F, _F],
    "gvm_tsv_value": [_P, _P, _I, _DP],
    "gvm_tsv_grad": [_P, _P, _P, _I, _F],
    "gvm_laplacian_value": [_P, _P, _I, _DP],
    "gvm_laplacian_grad": [_P, _P, _P, _I, _F],
    "gvm_add_to_dphi": [_P, _P, _P, _I],
    "gvm_priors_value_enqueue": [_P, _P, _P, _I, _I],
    "gvm_values_enqueue": [_P, _P, _I, _I, _I, _P, _I, _I],
    "gvm_dphi_assemble": [_P, _P, _P, _P, _P, _P, _I],
    "gvm_vec_zero": [_P, _P, _L],
    "gvm_vec_copy": [_P, _P, _P, _L],
    "gvm_cg_search_direction": [_P, _P, _P, _P, _I],
    "gvm_cg_grad_condition": [_P, _P, _P, _D, _I, _DP],
    "gvm_cg_gg_dgg": [_P, _P, _P, _I, _DP, _DP],
    "gvm_cg_new_xi": [_P, _P, _P, _P, _D, _I],
    "gvm_evaluate_xt": [_P, _P, _P, _P, _D, _I, _I, _F, _F],
    "gvm_new_p": [_P, _P, _P, _D, _I, _I, _F, _F],
    "gvm_vec_absmax": [_P, _P, _L, _DP],
    "gvm_vec_scale": [_P, _P, _P, _D, _L],
    "gvm_vec_dot": [_P, _P, _P, _L, _DP],
    "gvm_vec_axpy": [_P, _P, _P, _D, _L],
    "gvm_vec_dot_dev": [_P, _P, _P, _L, _P],
    "gvm_vec_dot_dev_noalloc": [_P, _P, _L, _P, _P],
    "gvm_scalar_rho_times": [_P, _P, _P, _P, _P],
    "gvm_scalar_h0": [_P, _P, _P, _I, _P],
    "gvm_vec_axpy_dev": [_P, _P, _P, _P, _I, _L],
    "gvm_vec_scale_dev": [_P, _P, _P, _P, _L],
    "gvm_lbfgs_s_and_y": [_P, _P, _P, _P, _P, _P, _P, _L],
    "gvm_primary_beam": [_I, _D, _D, _D, _D, _I, _I, _F, _F, _D, _D, _P],
    "gvm_sync": [_P],
    "gvm_malloc": [C.POINTER(_P), C.c_size_t],
    "gvm_free": [_P],
    "gvm_memcpy_h2d": [_P, _P, C.c_size_t],
    "gvm_memcpy_d2h": [_P, _P, C.c_size_t],
    "gvm_last_error": None,  # const char*
    "gvm_device_info": [C.c_char_p, _I, _IP, _IP],
    "gvm_timing_reset": [_P],
    "gvm_timing_adjoint_ms": [_P, _DP, C.POINTER(C.c_long)],
}

FW_API = {
    "gvm_ckernel_defaults": ([_I, C.POINTER(CKernelParams)], _I),
    "gvm_ckernel_build": ([C.POINTER(CKernelParams), C.c_float, C.c_float, _P, _IP, _IP], _I),
    "gvm_ckernel_gcf": ([C.POINTER(CKernelParams), _I, _I, C.c_float, C.c_float, _P], _I),
    "gvm_fits_write_image": ([C.c_char_p, _P, _I, _L, _L, _I, C.c_float, C.POINTER(FitsHeader)],
                             _I),
    "gvm_fits_last_error": ([], C.c_char_p),
    "gvm_sim_default_config": ([_I, C.POINTER(SimConfig)], None),
    "gvm_sim_count_visibilities": ([C.POINTER(SimConfig)], C.c_long),
    "gvm_sim_generate": ([C.POINTER(SimConfig), _P, _P, _P, _P, C.POINTER(SimComponent), _IP], _I),
    "gvm_sim_generate_flagged": ([C.POINTER(SimConfig), _P, _P, _P, _P, C.POINTER(SimComponent),
                                  _IP, C.POINTER(C.c_long)], _I),
    "gvm_setup_hermitian_lambda": ([_L, _P, _P, _F], None),
    "gvm_setup_beam_noise": ([_I, _FP, C.POINTER(_L), C.POINTER(_P), C.POINTER(_P), _D, _DP, _DP,
                              _DP, _DP, _DP, _DP], _I),
    "gvm_setup_pb_cutoff": ([_D, _D, _D], _D),
    "gvm_setup_noise_image": ([_I, _D, _D, _D, _D, _I, _I, _D, _D, _F, _F, _D, _D, _P, _FP, _FP],
                              _I),
    "gvm_fw_objective_create": ([_P, C.POINTER(FwConfig), C.POINTER(_P)], _I),
    "gvm_fw_objective_destroy": ([_P], _I),
    "gvm_fw_value": ([_P, _P, _DP], _I),
    "gvm_fw_gradient": ([_P, _P, _P, _I], _I),
    "gvm_fw_term_values": ([_P, _DP, _I, _IP], _I),
    "gvm_fw_optimize": ([_P, C.c_char_p, _I, _D, _D, _I, _P, C.POINTER(FwResult), _DP, _I], _I),
    "gvm_fw_optimize_timed": ([_P, C.c_char_p, _I, _D, _D, _I, _P, C.POINTER(FwResult), _DP, _DP,
                               _I], _I),
    "gvm_fw_set_flag": ([_P, _I], _I),
    "gvm_fw_set_graphs": ([_P, _I], _I),
    "gvm_fw_set_lbfgs_host_scalars": ([_P, _I], _I),
    "gvm_fw_set_fused_terms": ([_P, _I], _I),
    "gvm_fw_set_line_search_replay": ([_P, C.POINTER(C.c_double), _I], _I),
    "gvm_fw_line_search_steps": ([_P, C.POINTER(C.c_double), _I, C.POINTER(_I)], _I),
    "gvm_fw_graph_launches": ([_P, C.POINTER(_L)], _I),
    "gvm_fw_last_error": ([], C.c_char_p),
    "gvm_fw_registered": ([C.c_char_p, C.c_char_p], _I),
}

_hip = None
_fw = None


def load_hip() -> C.CDLL:
    """libgvm_hip.so with argtypes set.  Raises if it has not been built."""
    global _hip
    if _hip is None:
        path = LIB_DIR / "libgvm_hip.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: run gpuvmem_amd.build.build() first")
        lib = C.CDLL(str(path))
        for name, args in HIP_API.items():
            fn = getattr(lib, name)
            if args is None:
                fn.restype = C.c_char_p
                fn.argtypes = []
            else:
                fn.restype = C.c_int
                fn.argtypes = args
        _hip = lib
    return _hip


def load_fw() -> C.CDLL:
    """libgvm.so (framework, set-up, synthetic data) with argtypes set."""
    global _fw
    if _fw is None:
        load_hip()
        path = LIB_DIR / "libgvm.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: run gpuvmem_amd.build.build() first")
        lib = C.CDLL(str(path))
        for name, (args, res) in FW_API.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _fw = lib
    return _fw


def check(status: int, what: str) -> None:
    if status != GVM_OK:
        detail = load_hip().gvm_last_error().decode(errors="replace")
        raise GvmError(what, status, detail)


def check_fw(status: int, what: str) -> None:
    if status != GVM_OK:
        detail = load_fw().gvm_fw_last_error().decode(errors="replace")
        raise GvmError(what, status, detail)
