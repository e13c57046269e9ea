"""ctypes binding of libmgkmc_hip.so (C ABI declared in include/mgkmc_hip.h).

The library is loaded from this package directory only.  There is no CPU
fallback: if the library or a GPU is missing, every entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmgkmc_hip.so")

MGK_OK = 0
MGK_NOT_CONVERGED = 1
MGK_ERR_ARG = -1
MGK_ERR_HIP = -2
MGK_ERR_WORKSPACE = -3
MGK_ERR_NONFINITE = -4
MGK_ERR_SOLVER = -5
MGK_KMC_RUN_TIME, MGK_KMC_RUN_COUNT, MGK_KMC_RUN_SEQ, MGK_KMC_RUN_UNSTABLE = 0, 1, 2, -6
(MGK_CASCADE_DONE, MGK_CASCADE_TRUNCATED, MGK_CASCADE_EPS_TARGET, MGK_CASCADE_UNSTABLE,
 MGK_CASCADE_CAPACITY) = 0, 1, 2, 3, 4

_lib = None


class MgkLsInfo(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("status", ctypes.c_int), ("diff", ctypes.c_double),
                ("eps_mean", ctypes.c_double * 9), ("sig_mean", ctypes.c_double * 9)]


class MgkLsMember(ctypes.Structure):
    _fields_ = [("lam", ctypes.c_void_p), ("mu", ctypes.c_void_p), ("lam0", ctypes.c_double),
                ("mu0", ctypes.c_double), ("eps_bar", ctypes.POINTER(ctypes.c_double)),
                ("eps_plastic", ctypes.c_void_p), ("eps", ctypes.c_void_p), ("sig", ctypes.c_void_p),
                ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_size_t), ("lam32", ctypes.c_void_p),
                ("mu32", ctypes.c_void_p), ("eps_plastic32", ctypes.c_void_p), ("eps32", ctypes.c_void_p),
                ("workspace32", ctypes.c_void_p), ("workspace32_bytes", ctypes.c_size_t)]


MGK_LS_MAX_BATCH = 8


class MgkLandauParams(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("v1", "v2", "v3", "g1", "g2", "g3", "g4")] + \
               [("capping", ctypes.c_int), ("cap_smooth", ctypes.c_int), ("cap_limit", ctypes.c_double),
                ("cap_power", ctypes.c_double), ("cap_ratio", ctypes.c_double)]


class MgkSecantParams(ctypes.Structure):
    _fields_ = [("law", ctypes.c_int), ("plane_stress", ctypes.c_int), ("p1", ctypes.c_double),
                ("p2", ctypes.c_double)]


class MgkKmcParams(ctypes.Structure):
    _fields_ = [("volume", ctypes.c_double), ("rate_volume", ctypes.c_double), ("nu0", ctypes.c_double),
                ("temperature", ctypes.c_double), ("cap", ctypes.c_double), ("tau", ctypes.c_double),
                ("time", ctypes.c_double), ("scheme", ctypes.c_int), ("cascade", ctypes.c_int)]


class MgkKmcSel(ctypes.Structure):
    _fields_ = [("item", ctypes.c_longlong), ("mode", ctypes.c_int), ("fallthrough", ctypes.c_int),
                ("r", ctypes.c_double), ("cum_before", ctypes.c_double), ("cum_at", ctypes.c_double),
                ("q", ctypes.c_double)]


class MgkMtState(ctypes.Structure):
    _fields_ = [("key", ctypes.c_uint32 * 624), ("pos", ctypes.c_int), ("has_gauss", ctypes.c_int),
                ("gauss", ctypes.c_double)]


class MgkKmcLoop(ctypes.Structure):
    pass


class MgkKmcEvent(ctypes.Structure):
    _fields_ = [("v", ctypes.c_longlong), ("m", ctypes.c_int), ("is_instab", ctypes.c_int), ("time", ctypes.c_double),
                ("ls_iters", ctypes.c_int), ("ls_iters32", ctypes.c_int), ("eps_macro", ctypes.c_double * 9),
                ("sig_mean", ctypes.c_double * 9), ("substeps", ctypes.c_int)]


class MgkCascadeFlip(ctypes.Structure):
    _fields_ = [("batch", ctypes.c_int), ("v", ctypes.c_longlong), ("m", ctypes.c_int)]


MgkKmcLoop._fields_ = [
    ("ndim", ctypes.c_int), ("shape", ctypes.c_int64 * 3), ("M", ctypes.c_int),
    ("Q0", ctypes.c_void_p), ("cat", ctypes.c_void_p), ("soft", ctypes.c_void_p), ("tlast", ctypes.c_void_p),
    ("prev", ctypes.c_void_p), ("epsp", ctypes.c_void_p), ("eps_precision", ctypes.c_int),
    ("epsp_shadow", ctypes.c_void_p), ("kmc_ws", ctypes.c_void_p), ("redraw_dev", ctypes.c_void_p),
    ("kp", MgkKmcParams), ("jp", ctypes.c_double), ("jt", ctypes.c_double), ("jn", ctypes.c_double),
    ("stability_threshold", ctypes.c_double), ("redraw_directions", ctypes.c_int), ("redraw_barriers", ctypes.c_int),
    ("gamma0", ctypes.c_double), ("q_mean", ctypes.c_double), ("q_std", ctypes.c_double), ("q_min", ctypes.c_double),
    ("q_max", ctypes.c_double), ("q_has_min", ctypes.c_int), ("q_has_max", ctypes.c_int),
    ("ls_precision", ctypes.c_int), ("sig_kind", ctypes.c_int), ("lam", ctypes.c_void_p), ("mu", ctypes.c_void_p),
    ("lam32", ctypes.c_void_p), ("mu32", ctypes.c_void_p), ("lam0", ctypes.c_double), ("mu0", ctypes.c_double),
    ("eps", ctypes.c_void_p), ("sig", ctypes.c_void_p), ("eps32", ctypes.c_void_p), ("ws", ctypes.c_void_p),
    ("ws_bytes", ctypes.c_size_t), ("ws32", ctypes.c_void_p), ("ws32_bytes", ctypes.c_size_t), ("warm", ctypes.c_int),
    ("max_iter", ctypes.c_int), ("tol", ctypes.c_double), ("tol32", ctypes.c_double),
    ("check_unstable", ctypes.c_int), ("strain_rate", ctypes.c_double), ("unit", ctypes.c_double * 9),
    ("patch", ctypes.c_void_p), ("fs", ctypes.c_void_p), ("law", ctypes.c_void_p)]


class MgkKmcPatchCfg(ctypes.Structure):
    _fields_ = [("radius", ctypes.c_int * 3), ("nbasis", ctypes.c_int), ("kernels", ctypes.c_void_p),
                ("missing_mean", (ctypes.c_double * 6) * 5), ("sig_unit", ctypes.c_void_p),
                ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_size_t), ("sync_interval", ctypes.c_int),
                ("flips_since_sync", ctypes.POINTER(ctypes.c_int)), ("unit_live", ctypes.POINTER(ctypes.c_int)),
                ("sig_mean", ctypes.POINTER(ctypes.c_double)), ("mean_from_patch", ctypes.POINTER(ctypes.c_int))]


def mt_from(rs):
    """MgkMtState of a numpy RandomState (or the np.random module)."""
    name, key, pos, has_gauss, gauss = rs.get_state()
    if name != "MT19937":
        raise ValueError("an MT19937 RandomState is required")
    st = MgkMtState()
    ctypes.memmove(st.key, np.ascontiguousarray(key, dtype=np.uint32).ctypes.data, 624 * 4)
    st.pos, st.has_gauss, st.gauss = int(pos), int(has_gauss), float(gauss)
    return st


def mt_to(rs, st):
    """Hand an MgkMtState back to a numpy RandomState (or the np.random module)."""
    key = np.frombuffer(bytes(st.key), dtype=np.uint32).copy()
    rs.set_state(("MT19937", key, int(st.pos), int(st.has_gauss), float(st.gauss)))


class MgkFsMaterial(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int), ("capping", ctypes.c_int), ("cap_smooth", ctypes.c_int),
                ("plane", ctypes.c_int)] + \
               [(k, ctypes.c_double) for k in ("v1", "v2", "v3", "g1", "g2", "g3", "g4", "Am", "Bm", "Cm",
                                               "cap_limit", "cap_power", "cap_ratio")] + \
               [("landau_field", ctypes.c_void_p)]


class MgkFsInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("fail_kind", ctypes.c_int), ("newton_iters", ctypes.c_int),
                ("krylov_iters", ctypes.c_int), ("macro_iters", ctypes.c_int), ("max_err", ctypes.c_double),
                ("F_mean", ctypes.c_double * 9), ("sig_mean", ctypes.c_double * 9), ("F_bar", ctypes.c_double * 9)]


class MgkFsStepper(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int), ("precision", ctypes.c_int), ("mixed", ctypes.c_int),
                ("shape", ctypes.c_int64 * 3), ("F", ctypes.c_void_p), ("Fp", ctypes.c_void_p),
                ("lam", ctypes.c_void_p), ("mu", ctypes.c_void_p), ("lam32", ctypes.c_void_p),
                ("mu32", ctypes.c_void_p), ("mat", ctypes.c_void_p), ("P", ctypes.c_void_p),
                ("sig", ctypes.c_void_p), ("ws", ctypes.c_void_p), ("ws_bytes", ctypes.c_size_t),
                ("ws32", ctypes.c_void_p), ("ws32_bytes", ctypes.c_size_t), ("L0", ctypes.c_void_p),
                ("E_avg", ctypes.c_double), ("nu_avg", ctypes.c_double), ("tol_NW", ctypes.c_double),
                ("tol_CG", ctypes.c_double), ("tol_macro", ctypes.c_double), ("pixel", ctypes.c_double),
                ("max_NW", ctypes.c_int), ("max_iter_macro", ctypes.c_int)]


class MgkKmcFsSubstep(ctypes.Structure):
    _fields_ = [("kmc", ctypes.c_void_p), ("step", ctypes.c_void_p), ("drv_i", ctypes.c_int),
                ("drv_j", ctypes.c_int), ("n_targets", ctypes.c_int), ("target_k", ctypes.c_int * 3),
                ("target_val", ctypes.c_double * 3), ("nlive", ctypes.c_int), ("live", ctypes.c_void_p * 12),
                ("live_bytes", ctypes.c_size_t * 12), ("backup", ctypes.c_void_p), ("backup_bytes", ctypes.c_size_t),
                ("sig2", ctypes.c_void_p), ("flag", ctypes.c_void_p), ("F_macro", ctypes.c_void_p),
                ("info", ctypes.c_void_p)]


class MgkLawStepper(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("shape", ctypes.c_int64 * 2), ("lam", ctypes.c_void_p),
                ("mu", ctypes.c_void_p), ("lam0", ctypes.c_double), ("mu0", ctypes.c_double),
                ("landau", ctypes.c_void_p), ("secant", ctypes.c_void_p), ("plane_stress", ctypes.c_int),
                ("eps", ctypes.c_void_p), ("sig", ctypes.c_void_p), ("e33", ctypes.c_void_p),
                ("e33_have", ctypes.c_void_p), ("tau", ctypes.c_void_p), ("anderson_m", ctypes.c_int),
                ("aa", ctypes.c_void_p), ("ws", ctypes.c_void_p), ("ws_bytes", ctypes.c_size_t),
                ("max_iter", ctypes.c_int), ("tol", ctypes.c_double)]


class MgkKmcLawSubstep(ctypes.Structure):
    _fields_ = [("step", ctypes.c_void_p), ("nlive", ctypes.c_int), ("live", ctypes.c_void_p * 12),
                ("live_bytes", ctypes.c_size_t * 12), ("backup", ctypes.c_void_p), ("backup_bytes", ctypes.c_size_t)]


_c_p = ctypes.c_void_p
_c_i = ctypes.c_int
_c_d = ctypes.c_double
_c_sz = ctypes.c_size_t
_c_i64p = ctypes.POINTER(ctypes.c_int64)
_c_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); every symbol of include/mgkmc_hip.h
SIGNATURES = {
    "mgk_version": (_c_i, []),
    "mgk_last_error": (ctypes.c_char_p, []),
    "mgk_device_info": (_c_i, [ctypes.POINTER(_c_i), ctypes.POINTER(_c_i)]),
    "mgk_profile_enable": (None, [_c_i]),
    "mgk_profile_reset": (None, []),
    "mgk_profile_read": (_c_i, [_c_i, _c_dp, _c_dp, ctypes.POINTER(ctypes.c_longlong)]),
    "mgk_ls_workspace_bytes": (_c_i, [_c_i, _c_i64p, _c_i, ctypes.POINTER(_c_sz)]),
    "mgk_ls_solve": (_c_i, [_c_i, _c_i64p, _c_i, _c_p, _c_p, _c_d, _c_d, _c_dp, _c_p, _c_p, _c_p,
                            _c_i, _c_i, _c_d, _c_p, _c_sz, _c_p, ctypes.POINTER(MgkLsInfo)]),
    "mgk_stream_copy": (_c_i, [_c_p, _c_p, _c_sz, _c_p]),
    "mgk_ls_solve_noalloc": (_c_i, [_c_i, _c_i64p, _c_i, _c_p, _c_p, _c_d, _c_d, _c_dp, _c_p, _c_p, _c_p,
                                    _c_i, _c_i, _c_d, _c_p, ctypes.POINTER(MgkLsInfo)]),
    "mgk_ls_solve_refined": (_c_i, [_c_i, _c_i64p, _c_p, _c_p, _c_p, _c_p, _c_d, _c_d, _c_dp, _c_p, _c_p, _c_p,
                                    _c_p, _c_p, _c_i, _c_i, _c_d, _c_d, _c_p, _c_sz, _c_p, _c_sz, _c_p,
                                    ctypes.POINTER(MgkLsInfo), ctypes.POINTER(_c_i)]),
    "mgk_ls_solve_batch": (_c_i, [_c_i, _c_i64p, _c_i, _c_i, ctypes.POINTER(MgkLsMember), _c_i, _c_i, _c_d, _c_p,
                                  ctypes.POINTER(MgkLsInfo)]),
    "mgk_ls_solve_refined_batch": (_c_i, [_c_i, _c_i64p, _c_i, ctypes.POINTER(MgkLsMember), _c_i, _c_i, _c_d, _c_d,
                                          _c_p, ctypes.POINTER(MgkLsInfo), ctypes.POINTER(_c_i)]),
    "mgk_ls_solve_landau": (_c_i, [_c_i64p, _c_i, _c_p, _c_p, _c_d, _c_d, ctypes.POINTER(MgkLandauParams), _c_dp, _c_p,
                                   _c_p, _c_p, _c_p, _c_i, _c_i, _c_d, _c_p, _c_sz, _c_p, ctypes.POINTER(MgkLsInfo)]),
    "mgk_ls_solve_landau2d": (_c_i, [_c_i64p, _c_p, _c_p, _c_d, _c_d, ctypes.POINTER(MgkLandauParams), _c_i, _c_dp,
                                     _c_p, _c_p, _c_p, _c_p, _c_i, _c_i, _c_p, _c_i, _c_p, _c_i, _c_i, _c_d, _c_p,
                                     _c_sz, _c_p, ctypes.POINTER(MgkLsInfo)]),
    "mgk_ls_landau2d_ncg_workspace_bytes": (_c_i, [_c_i64p, ctypes.POINTER(_c_sz)]),
    "mgk_ls_solve_landau2d_ncg": (_c_i, [_c_i64p, _c_p, _c_p, _c_d, _c_d, ctypes.POINTER(MgkLandauParams), _c_i, _c_dp,
                                         _c_p, _c_p, _c_p, _c_i, _c_i, _c_d, _c_p, _c_sz, _c_p,
                                         ctypes.POINTER(MgkLsInfo), ctypes.POINTER(_c_i)]),
    "mgk_ls_solve_secant": (_c_i, [_c_i, _c_i64p, _c_i, _c_p, _c_p, _c_d, _c_d, ctypes.POINTER(MgkSecantParams), _c_dp,
                                   _c_p, _c_p, _c_p, _c_p, _c_i, _c_i, _c_d, _c_p, _c_sz, _c_p,
                                   ctypes.POINTER(MgkLsInfo)]),
    "mgk_heat_workspace_bytes": (_c_i, [_c_i, _c_i64p, ctypes.POINTER(_c_sz)]),
    "mgk_heat_step": (_c_i, [_c_i, _c_i64p, _c_p, _c_dp, _c_d, _c_d, _c_i, _c_d, _c_d, _c_p, _c_sz, _c_p]),
    "mgk_mt_fill": (_c_i, [ctypes.POINTER(MgkMtState), _c_i, _c_d, _c_d, ctypes.c_longlong, _c_dp]),
    "mgk_kmc_fs_flip": (_c_i, [ctypes.POINTER(MgkKmcFsSubstep), _c_i, ctypes.POINTER(ctypes.c_longlong),
                               ctypes.POINTER(_c_i), _c_d, ctypes.POINTER(MgkMtState), _c_dp, _c_dp,
                               ctypes.POINTER(_c_i), ctypes.POINTER(MgkFsInfo), _c_p]),
    "mgk_kmc_law_flip": (_c_i, [ctypes.POINTER(MgkKmcLoop), ctypes.POINTER(MgkKmcLawSubstep), _c_i,
                                ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(_c_i), _c_d,
                                ctypes.POINTER(MgkMtState), _c_dp, ctypes.POINTER(_c_i), ctypes.POINTER(MgkLsInfo),
                                _c_p]),
    "mgk_kmc_mixed_bc": (_c_i, [ctypes.POINTER(MgkKmcLoop), _c_i, ctypes.POINTER(_c_i), _c_dp, _c_i, _c_i, _c_d,
                                _c_d, _c_d, _c_i, _c_dp, ctypes.POINTER(_c_i), _c_dp, ctypes.POINTER(_c_i),
                                ctypes.POINTER(MgkLsInfo), ctypes.POINTER(_c_i), _c_p]),
    "mgk_kmc_run": (_c_i, [ctypes.POINTER(MgkKmcLoop), ctypes.POINTER(MgkMtState), ctypes.POINTER(MgkMtState),
                           _c_dp, _c_dp, _c_dp, _c_i, _c_i, ctypes.POINTER(_c_i), ctypes.POINTER(MgkKmcEvent),
                           ctypes.POINTER(_c_i), ctypes.POINTER(MgkLsInfo), ctypes.POINTER(_c_i), _c_p]),
    "mgk_kmc_run_batch": (_c_i, [_c_i, ctypes.POINTER(MgkKmcLoop), ctypes.POINTER(MgkMtState),
                                 ctypes.POINTER(MgkMtState), _c_dp, _c_dp, _c_dp, ctypes.POINTER(_c_i), _c_i,
                                 ctypes.POINTER(_c_i), ctypes.POINTER(MgkKmcEvent), _c_i, ctypes.POINTER(_c_i),
                                 ctypes.POINTER(_c_i), ctypes.POINTER(MgkLsInfo), ctypes.POINTER(_c_i), _c_p]),
    "mgk_kmc_cascade": (_c_i, [ctypes.POINTER(MgkKmcLoop), ctypes.POINTER(MgkMtState), _c_p, _c_d, _c_dp, _c_i,
                               _c_i, _c_i, _c_d, ctypes.POINTER(MgkCascadeFlip), _c_i, ctypes.POINTER(_c_i),
                               ctypes.POINTER(_c_i), _c_dp, _c_i, ctypes.POINTER(_c_i), ctypes.POINTER(_c_i),
                               ctypes.POINTER(MgkLsInfo), ctypes.POINTER(_c_i), _c_p]),
    "mgk_kmc_workspace_bytes": (_c_i, [ctypes.c_longlong, ctypes.POINTER(_c_sz)]),
    "mgk_kmc_barrier_rates": (_c_i, [_c_i, ctypes.c_longlong, _c_i, _c_i, _c_i, _c_p, _c_p, _c_p, _c_p, _c_p,
                                     _c_p, _c_p, ctypes.POINTER(MgkKmcParams), _c_p, _c_p, _c_p, _c_dp]),
    "mgk_kmc_select": (_c_i, [_c_i, ctypes.c_longlong, _c_i, _c_i, _c_i, _c_d, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
                              _c_p, ctypes.POINTER(MgkKmcParams), _c_p, _c_p, ctypes.POINTER(MgkKmcSel)]),
    "mgk_kmc_rates_from_q": (_c_i, [ctypes.c_longlong, _c_i, _c_p, _c_p, ctypes.POINTER(MgkKmcParams), _c_p,
                                    _c_p, _c_p, _c_p, _c_p, ctypes.POINTER(ctypes.c_longlong), _c_dp]),
    "mgk_kmc_select_array": (_c_i, [ctypes.c_longlong, _c_p, _c_d, _c_p, _c_p, ctypes.POINTER(MgkKmcSel)]),
    "mgk_find_below": (_c_i, [ctypes.c_longlong, _c_p, _c_d, _c_p, _c_p, ctypes.POINTER(ctypes.c_longlong)]),
    "mgk_kmc_flip": (_c_i, [_c_i, ctypes.POINTER(_c_i), _c_i, _c_i, ctypes.c_longlong, _c_i, _c_d, _c_d, _c_d,
                            _c_d, _c_d, _c_p, _c_p, _c_i, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i, _c_p]),
    "mgk_kmc_set_voxel": (_c_i, [ctypes.c_longlong, _c_i, _c_i, ctypes.c_longlong, _c_p, _c_p, _c_p, _c_p,
                                 _c_p]),
    "mgk_kmc_patch_workspace_bytes": (_c_i, [ctypes.POINTER(_c_sz)]),
    "mgk_kmc_patch": (_c_i, [_c_i, _c_i64p, _c_i, _c_p, _c_p, _c_d, ctypes.c_longlong, ctypes.POINTER(_c_i), _c_i,
                             _c_p, _c_dp, _c_dp, _c_p, _c_sz, _c_p, _c_dp]),
    "mgk_kmc_init_modes": (_c_i, [ctypes.c_longlong, _c_i, ctypes.c_uint64, _c_d, _c_d, _c_i, _c_d, _c_i, _c_d, _c_d,
                                  _c_p, _c_p, _c_p]),
    "mgk_fs_workspace_bytes": (_c_i, [_c_i64p, _c_i, ctypes.POINTER(_c_sz)]),
    "mgk_fs_solver_step": (_c_i, [_c_i64p, _c_i, _c_p, _c_p, _c_p, _c_p, ctypes.POINTER(MgkFsMaterial), _c_dp,
                                  _c_dp, ctypes.POINTER(_c_i), _c_i, _c_i, _c_d, _c_d, _c_d, _c_d, _c_i, _c_d, _c_i,
                                  _c_p, _c_p, _c_p, _c_sz, _c_p, ctypes.POINTER(MgkFsInfo)]),
    "mgk_fs_solver_step_al": (_c_i, [_c_i64p, _c_i, _c_p, _c_p, _c_p, _c_p, ctypes.POINTER(MgkFsMaterial), _c_dp,
                                     _c_dp, ctypes.POINTER(_c_i), _c_i, _c_i, _c_d, _c_d, _c_d, _c_i, _c_d, _c_i, _c_dp,
                                     _c_p, _c_p, _c_p, _c_sz, _c_p, ctypes.POINTER(MgkFsInfo)]),
    "mgk_fs_solver_step_dbfft": (_c_i, [_c_i64p, _c_i, _c_p, _c_p, _c_p, _c_p, ctypes.POINTER(MgkFsMaterial), _c_dp,
                                        _c_dp, ctypes.POINTER(_c_i), _c_i, _c_i, _c_d, _c_d, _c_d, _c_d, _c_i, _c_d, _c_i,
                                        _c_d, _c_p, _c_p, _c_p, _c_sz, _c_p, ctypes.POINTER(MgkFsInfo)]),
    "mgk_fs_solver_step_refined": (_c_i, [_c_i64p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, ctypes.POINTER(MgkFsMaterial),
                                          _c_dp, _c_dp, ctypes.POINTER(_c_i), _c_i, _c_i, _c_d, _c_d, _c_d, _c_d, _c_i,
                                          _c_d, _c_i, _c_p, _c_p, _c_p, _c_sz, _c_p, _c_sz, _c_p,
                                          ctypes.POINTER(MgkFsInfo)]),
    "mgk_fs_constitutive": (_c_i, [ctypes.c_longlong, _c_i, _c_p, _c_p, _c_p, _c_p, ctypes.POINTER(MgkFsMaterial),
                                   _c_p, _c_p, _c_p, _c_p]),
    "mgk_fs_constitutive_2d": (_c_i, [ctypes.c_longlong, _c_i, _c_p, _c_p, _c_p, _c_p, ctypes.POINTER(MgkFsMaterial),
                                      _c_p, _c_p, _c_p, _c_p, _c_p]),
    "mgk_fs_cauchy": (_c_i, [ctypes.c_longlong, _c_i, _c_p, _c_p, _c_p, _c_p]),
}


def load():
    """Load the library (once) and bind every exported symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not built: run `python -m mgkmc_amd.build` "
                           "(there is no CPU fallback for the HIP path)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    return load().mgk_last_error().decode(errors="replace")


def check(rc, what):
    """Map a C-ABI status to Python exceptions (MGK_NOT_CONVERGED is silent,
    as the reference's LS loop stops silently at max_iter)."""
    if rc in (MGK_OK, MGK_NOT_CONVERGED):
        return rc
    if rc == MGK_ERR_NONFINITE:
        return rc
    raise RuntimeError(f"{what} failed with status {rc}: {last_error()}")


def shape_arg(shape):
    arr = (ctypes.c_int64 * len(shape))(*[int(s) for s in shape])
    return arr


def dptr(t):
    """Device pointer of a torch tensor (or None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def host_dbl(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_c_dp)


# kernel ids of the live timing counters (mgk_host.h MgkKid)
KID = {"ls_tau_xfwd": 0, "ls_fft_mid": 1, "ls_pencil_gamma": 2, "ls_xinv_update": 3,
       "kmc_barrier_rates": 4, "kmc_select": 5, "fs_constitutive": 6, "fs_xfwd": 7, "fs_fft_mid": 8,
       "fs_pencil_proj": 9, "fs_xinv": 10, "fs_vector": 11, "ls_tau_xfwd_start": 12,
       "ls_xinv_update_start": 13}


def profile_read():
    """{kernel name: (total_ms, total_algorithmic_bytes, launches)}"""
    lib = load()
    out = {}
    for name, kid in KID.items():
        ms, by, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        lib.mgk_profile_read(kid, ctypes.byref(ms), ctypes.byref(by), ctypes.byref(n))
        if n.value:
            out[name] = (ms.value, by.value, n.value)
    return out
