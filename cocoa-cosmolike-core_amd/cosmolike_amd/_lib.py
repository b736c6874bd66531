"""ctypes binding of libcosmolike_hip.so (the C ABI in include/cosmolike_hip.h).

The library is required: importing a compute path without it raises, there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libcosmolike_hip.so")

CLH_OK = 0
ERRORS = {-1: "CLH_ERR_ARG", -2: "CLH_ERR_NAN", -3: "CLH_ERR_STATE", -4: "CLH_ERR_DOMAIN",
          -5: "CLH_ERR_HIP", -6: "CLH_ERR_ALLOC"}

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)
VP = ctypes.c_void_p


class clh_draw(ctypes.Structure):
    _fields_ = [
        ("omega_m", ctypes.c_double),
        ("chi_z", D), ("chi", D), ("chi_n", ctypes.c_int),
        ("G_z", D), ("G", D), ("G_n", ctypes.c_int),
        ("log10k", D), ("pk_z", D), ("lnP", D), ("nk", ctypes.c_int), ("npz", ctypes.c_int),
        ("src_dz", D), ("lens_dz", D), ("lens_stretch", D),
        ("b1", D), ("bmag", D), ("shear_m", D),
        ("ia_A", ctypes.c_double), ("ia_eta", ctypes.c_double), ("ia_bin", D),
        ("pm", D),
        ("ia2_A", ctypes.c_double), ("ia2_eta", ctypes.c_double), ("ia2_bin", D),
        ("bta_A", ctypes.c_double), ("bta_bin", D),
        ("fpt", D), ("fpt_n", ctypes.c_int), ("fpt_kmin", ctypes.c_double), ("fpt_kmax", ctypes.c_double),
        ("h", ctypes.c_double),
        ("plin_log10k", D), ("plin_z", D), ("lnPL", D), ("plin_nk", ctypes.c_int), ("plin_nz", ctypes.c_int),
        ("b2", D), ("bs2", D), ("b3", D), ("bk", D),
        ("fptb", D), ("fptb_n", ctypes.c_int), ("fptb_kmin", ctypes.c_double), ("fptb_kmax", ctypes.c_double),
        ("fptb_sigma4", ctypes.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(python -m cosmolike_amd.build)")
        # torch bundles its own HIP runtime with the same soname
        # (libamdhip64.so.7); loading torch first makes this library bind to
        # that copy, so the process holds a single HIP runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        i32 = ctypes.c_int
        f64 = ctypes.c_double
        sig = {
            "clh_create": (i32, [ctypes.POINTER(VP), i32]),
            "clh_destroy": (None, [VP]),
            "clh_last_error": (ctypes.c_char_p, [VP]),
            "clh_default": (VP, []),
            "clh_set_legendre_mode": (i32, [VP, i32]),
            "clh_init_ntable": (i32, [VP, i32, i32, i32, i32]),
            "clh_init_binning_real_space": (i32, [VP, i32, f64, f64]),
            "clh_init_ia": (i32, [VP, i32, i32]),
            "clh_set_cosmological_parameters": (i32, [VP, f64, f64]),
            "clh_set_distances": (i32, [VP, D, D, i32]),
            "clh_set_growth": (i32, [VP, D, D, i32]),
            "clh_set_non_linear_power_spectrum": (i32, [VP, D, i32, D, i32, D]),
            "clh_set_source_sample": (i32, [VP, D, i32, i32]),
            "clh_set_lens_sample": (i32, [VP, D, i32, i32]),
            "clh_set_nuisance_shear_calib": (i32, [VP, D, i32]),
            "clh_set_nuisance_shear_photoz": (i32, [VP, D, i32]),
            "clh_set_nuisance_clustering_photoz": (i32, [VP, D, i32]),
            "clh_set_nuisance_clustering_photoz_stretch": (i32, [VP, D, i32]),
            "clh_set_nuisance_linear_bias": (i32, [VP, D, i32]),
            "clh_set_nuisance_magnification_bias": (i32, [VP, D, i32]),
            "clh_set_nuisance_ia": (i32, [VP, D, i32]),
            "clh_set_nuisance_ia_tatt": (i32, [VP, D, D, D, i32]),
            "clh_set_IA_PS": (i32, [VP, D, f64, f64, f64, i32]),
            "clh_set_linear_power_spectrum": (i32, [VP, D, i32, D, i32, D]),
            "clh_init_probes": (i32, [VP, ctypes.c_char_p]),
            "clh_init_cmb_cross_correlation": (i32, [VP, i32, i32, f64, D, i32]),
            "clh_compute_6x2pt_real": (i32, [VP, ctypes.POINTER(ctypes.c_int), D, i32]),
            "clh_init_cmb_auto_bandpower": (i32, [VP, i32, i32, i32, D, D, f64]),
            "clh_cmb_kk_bandpower_info": (i32, [VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                                ctypes.POINTER(ctypes.c_int), D]),
            "clh_set_nuisance_nonlinear_bias": (i32, [VP, D, D, i32]),
            "clh_set_nuisance_nonlocal_bias": (i32, [VP, D, D, i32]),
            "clh_set_bias_PS": (i32, [VP, D, f64, f64, f64, f64, i32]),
            "clh_set_adopt_limber_gg": (i32, [VP, i32]),
            "clh_cl_nonlimber": (i32, [VP, D, ctypes.POINTER(ctypes.c_int)]),
            "clh_set_nuisance_pointmass": (i32, [VP, D, i32]),
            "clh_init_ggl_exclude": (i32, [VP, I, i32]),
            "clh_test_zoverlap": (i32, [VP, i32, i32]),
            "clh_int_for_limber": (i32, [VP, i32, D, D, i32, D]),
            "clh_set_vjp_checkpoint": (i32, [VP, i32]),
            "clh_set_vjp_lens_limits": (i32, [VP, i32]),
            "clh_init_binning_fourier": (i32, [VP, i32, i32, i32, i32]),
            "clh_datavector_size_fourier": (i32, [VP]),
            "clh_compute_3x2pt_fourier": (i32, [VP, I, I, D, i32]),
            "clh_like_set_mask": (i32, [VP, I, i32, I]),
            "clh_like_set_data": (i32, [VP, D, i32, I]),
            "clh_like_set_inv_cov": (i32, [VP, D, i32, I]),
            "clh_chi2": (i32, [VP, D, i32, D]),
            "clh_batch_chi2": (i32, [VP, I, VP, VP, VP]),
            "clh_datavector_size": (i32, [VP]),
            "clh_compute_3x2pt_real": (i32, [VP, I, D, i32]),
            "clh_xi_pm": (i32, [VP, D, D]),
            "clh_gammat": (i32, [VP, D]),
            "clh_wtheta": (i32, [VP, D]),
            "clh_cl_tables": (i32, [VP, D, D, D]),
            "clh_cl_tables_bb": (i32, [VP, D]),
            "clh_cl_nointerp": (i32, [VP, i32, D, i32, D]),
            "clh_batch_upload": (i32, [VP, i32, ctypes.POINTER(clh_draw)]),
            "clh_layout_capture": (i32, [VP, ctypes.POINTER(VP)]),
            "clh_layout_destroy": (None, [VP]),
            "clh_layout_datavector_size": (i32, [VP]),
            "clh_batch_upload_layouts": (i32, [VP, i32, ctypes.POINTER(clh_draw), ctypes.POINTER(VP)]),
            "clh_batch_datavector_offset": (ctypes.c_longlong, [VP, i32]),
            "clh_set_precision": (i32, [VP, i32]),
            "clh_batch_forward": (i32, [VP, I, VP]),
            "clh_synchronize": (i32, [VP]),
            "clh_set_stream": (i32, [VP, VP]),
            "clh_batch_use_graph": (i32, [VP, i32]),
            "clh_batch_set_params": (i32, [VP, VP, VP, VP, VP, VP, VP]),
            "clh_batch_vjp": (i32, [VP, I, VP, VP, VP, VP, VP, VP, VP]),
            "clh_last_kernel_times": (i32, [VP, ctypes.POINTER(ctypes.c_char_p),
                                            ctypes.POINTER(ctypes.c_float), i32]),
            "clh_enable_timing": (i32, [VP, i32]),
            "clh_set_tile_rows": (i32, [VP, I]),
            "clh_get_tile_rows": (i32, [VP, I]),
            "clh_autotune": (i32, [VP, I, I]),
            "xi_pm_tomo": (f64, [i32, i32, i32, i32, i32]),
            "w_gammat_tomo": (f64, [i32, i32, i32, i32]),
            "w_gg_tomo": (f64, [i32, i32, i32, i32]),
            "C_ss_tomo_limber": (f64, [f64, i32, i32, i32]),
            "C_gs_tomo_limber": (f64, [f64, i32, i32]),
            "C_gg_tomo_limber": (f64, [f64, i32, i32]),
            "C_ss_tomo_limber_nointerp": (f64, [f64, i32, i32, i32, i32]),
            "C_gs_tomo_limber_nointerp": (f64, [f64, i32, i32, i32]),
            "C_gg_tomo_limber_nointerp": (f64, [f64, i32, i32, i32]),
            "test_zoverlap": (i32, [i32, i32]),
            "N_ggl": (i32, [i32, i32]),
            "ZL": (i32, [i32]),
            "ZS": (i32, [i32]),
            "w_gk_tomo": (f64, [i32, i32, i32]),
            "w_ks_tomo": (f64, [i32, i32, i32]),
            "C_gk_tomo_limber": (f64, [f64, i32]),
            "C_ks_tomo_limber": (f64, [f64, i32]),
            "C_kk_limber": (f64, [f64]),
            "C_gk_tomo_limber_nointerp": (f64, [f64, i32, i32]),
            "C_ks_tomo_limber_nointerp": (f64, [f64, i32, i32]),
            "C_kk_limber_nointerp": (f64, [f64, i32]),
            "int_for_C_ss_tomo_limber": (f64, [f64, D]),
            "int_for_C_gs_tomo_limber": (f64, [f64, D]),
            "int_for_C_gg_tomo_limber": (f64, [f64, D]),
            "int_for_C_gk_tomo_limber": (f64, [f64, D]),
            "int_for_C_ks_tomo_limber": (f64, [f64, D]),
            "int_for_C_kk_limber": (f64, [f64, D]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
