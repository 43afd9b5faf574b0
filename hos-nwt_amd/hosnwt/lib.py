"""ctypes binding of libhosnwt.so (the C-ABI declared in include/hosnwt.h).

The library is built in-tree by ``hos-nwt_amd/build.sh`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# HOSNWT_LIB selects another build of the same library (e.g. a tuning variant).
LIB_PATH = os.environ.get("HOSNWT_LIB") or os.path.join(os.path.dirname(_HERE), "lib", "libhosnwt.so")

_dp = C.POINTER(C.c_double)


class HosParams(C.Structure):
    _fields_ = [
        ("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int), ("mHOS", C.c_int),
        ("p1", C.c_int), ("p2", C.c_int), ("icase", C.c_int), ("igeom", C.c_int),
        ("i_wmk", C.c_int), ("iramp", C.c_int), ("iabsnb", C.c_int), ("precision", C.c_int),
        ("xlen", C.c_double), ("ylen", C.c_double), ("d_hinge", C.c_double),
        ("Tramp", C.c_double), ("T_stop", C.c_double), ("T_mono", C.c_double),
        ("nu", _dp), ("n_freq", C.c_int), ("freq", _dp),
        ("pos_ini_re", _dp), ("pos_ini_im", _dp), ("coeffilt", C.c_double * 3),
        ("dposdy_ini_re", _dp), ("dposdy_ini_im", _dp), ("nu_x", _dp), ("nu_y", _dp),
        ("n_dip", C.c_int), ("dip_freq", _dp), ("dip_mux_re", _dp), ("dip_mux_im", _dp),
        ("dip_muz_re", _dp), ("dip_muz_im", _dp), ("dip_z", _dp), ("x_dip", C.c_double), ("n_side", C.c_int),
    ]


class HosStats(C.Structure):
    _fields_ = [
        ("n_rhs", C.c_longlong), ("n_big", C.c_longlong), ("n_base", C.c_longlong),
        ("big_bytes", C.c_longlong), ("big_n1", C.c_int), ("big_n2", C.c_int),
    ]


class HosOrderPolicy(C.Structure):
    _fields_ = [("m_min", C.c_int), ("judge_every", C.c_int), ("rate_tol", C.c_double), ("safety", C.c_double)]


class HosLearnedStats(C.Structure):
    _fields_ = [("attempts", C.c_longlong * 9), ("judges", C.c_longlong), ("reverts", C.c_longlong),
                ("max_judged_rate", C.c_double), ("max_pred_rate", C.c_double)]


FEAT_REC = 578  # HOS_FEAT_REC

EXPORTS = [
    "hos_create", "hos_destroy", "hos_last_error", "hos_initialize", "hos_set_state",
    "hos_get_state", "hos_rk_step", "hos_commit", "hos_end_ramp", "hos_advance", "hos_rhs",
    "hos_get_physical", "hos_get_slope_max", "hos_get_wmk", "hos_get_stats", "hos_synchronize",
    "hos_space_2_fourier", "hos_fourier_2_space", "hos_time_big_transform", "hos_time_big_batch", "hos_pass_log_begin", "hos_pass_log_end", "hos_check_slope",
    "hos_set_state2", "hos_get_state2", "hos_rhs2", "hos_set_state3", "hos_get_state3", "hos_rhs3",
    "hos_get_wmk_motion", "hos_get_modes", "hos_probes", "hos_advance_ensemble",
    "hos_set_order", "hos_get_order", "hos_record_features", "hos_take_features",
    "hos_mlp_info", "hos_mlp_create", "hos_mlp_destroy", "hos_mlp_set_params", "hos_mlp_get_params",
    "hos_mlp_forward", "hos_mlp_loss_grad", "hos_mlp_set_data", "hos_mlp_train",
    "hos_advance_learned", "hos_learned_last", "hos_learned_reset", "hos_mlp_forward_noalloc",
    "hos_advance_graph", "hos_advance_ensemble_graph", "hos_dense_product", "hos_get_schedule",
    "hos_swense_modes",
]

_lib = None


def load() -> C.CDLL:
    """Load libhosnwt.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not built: run hos-nwt_amd/build.sh")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    lib.hos_last_error.restype = C.c_char_p
    lib.hos_create.argtypes = [C.POINTER(HosParams), C.POINTER(vp)]
    lib.hos_destroy.argtypes = [vp]
    lib.hos_initialize.argtypes = [vp]
    lib.hos_set_state.argtypes = [vp] + [_dp] * 6
    lib.hos_get_state.argtypes = [vp] + [_dp] * 6
    lib.hos_rk_step.argtypes = [vp, C.c_double, C.c_double, _dp]
    lib.hos_commit.argtypes = [vp]
    lib.hos_end_ramp.argtypes = [vp]
    lib.hos_advance.argtypes = [vp, _dp, C.c_double, _dp, C.c_double, C.c_double, C.c_double,
                                C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.hos_advance_ensemble.argtypes = [C.c_int, C.POINTER(vp), _dp, C.c_double, _dp, C.c_double, _dp, _dp,
                                         C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.hos_rhs.argtypes = [vp, C.c_double, C.c_double] + [_dp] * 11
    lib.hos_rhs2.argtypes = [vp, C.c_double, C.c_double] + [_dp] * 17
    lib.hos_set_state2.argtypes = [vp] + [_dp] * 3
    lib.hos_get_state2.argtypes = [vp] + [_dp] * 3
    lib.hos_rhs3.argtypes = [vp, C.c_double, C.c_double] + [_dp] * 19
    lib.hos_set_state3.argtypes = [vp, _dp]
    lib.hos_get_state3.argtypes = [vp, _dp]
    lib.hos_get_physical.argtypes = [vp] + [_dp] * 4
    lib.hos_get_slope_max.argtypes = [vp, _dp, _dp]
    lib.hos_check_slope.argtypes = [vp, C.c_double, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.hos_get_wmk.argtypes = [vp, _dp, _dp]
    lib.hos_get_wmk_motion.argtypes = [vp, _dp, _dp, _dp, _dp]
    lib.hos_get_modes.argtypes = [vp, _dp, _dp]
    lib.hos_swense_modes.argtypes = [vp, C.c_double, C.c_int] + [_dp] * 8
    lib.hos_probes.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    lib.hos_get_stats.argtypes = [vp, C.POINTER(HosStats)]
    lib.hos_synchronize.argtypes = [vp]
    lib.hos_time_big_transform.argtypes = [vp, C.c_int, _dp]
    lib.hos_time_big_batch.argtypes = [vp, C.c_int, C.c_int, _dp]
    lib.hos_pass_log_begin.argtypes = [vp, C.c_int]
    lib.hos_pass_log_end.argtypes = [vp, _dp]
    ip, fp = C.POINTER(C.c_int), C.POINTER(C.c_float)
    lib.hos_set_order.argtypes = [vp, C.c_int]
    lib.hos_get_order.argtypes = [vp, ip]
    lib.hos_record_features.argtypes = [vp, C.c_int]
    lib.hos_take_features.argtypes = [vp, C.c_int, _dp, ip]
    lib.hos_mlp_info.argtypes = [ip, ip, ip, C.POINTER(C.c_longlong)]
    lib.hos_mlp_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.hos_mlp_destroy.argtypes = [vp]
    lib.hos_mlp_set_params.argtypes = [vp, fp]
    lib.hos_mlp_get_params.argtypes = [vp, fp]
    lib.hos_mlp_forward.argtypes = [vp, C.c_int, fp, fp]
    lib.hos_mlp_loss_grad.argtypes = [vp, C.c_int, fp, fp, _dp, fp]
    lib.hos_mlp_set_data.argtypes = [vp, C.c_int, fp, fp]
    lib.hos_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_ulonglong, fp, _dp]
    lib.hos_advance_learned.argtypes = [C.c_int, C.POINTER(vp), _dp, C.c_double, _dp, C.c_double, _dp, _dp, vp,
                                        C.POINTER(HosOrderPolicy), ip, ip, C.POINTER(HosLearnedStats)]
    lib.hos_learned_last.argtypes = [vp, _dp, fp]
    lib.hos_learned_reset.argtypes = [vp]
    lib.hos_mlp_forward_noalloc.argtypes = [vp, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.hos_advance_graph.argtypes = [vp, _dp, C.c_double, _dp, C.c_double, C.c_double, C.c_double,
                                      C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.hos_advance_ensemble_graph.argtypes = [C.c_int, C.POINTER(vp), _dp, C.c_double, _dp, C.c_double, _dp, _dp,
                                               C.POINTER(C.c_int), C.POINTER(C.c_int)]
    for f in ("hos_space_2_fourier", "hos_fourier_2_space"):
        getattr(lib, f).argtypes = [C.c_int] * 5 + [_dp, _dp]
    lib.hos_dense_product.argtypes = [C.c_int] * 9 + [_dp, _dp, _dp]
    lib.hos_get_schedule.argtypes = [vp, C.POINTER(C.c_int), _dp]
    for f in EXPORTS:
        if f != "hos_last_error":
            getattr(lib, f).restype = C.c_int
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError("hosnwt: " + load().hos_last_error().decode())


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(_dp)


def fptr(a: np.ndarray | None):
    """float32 pointer of a contiguous float32 array."""
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def fortran(a) -> np.ndarray:
    """Column-major float64 copy (the reference's array layout)."""
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def dense_product(a: np.ndarray, b: np.ndarray, a_kmajor: bool = False, b_kmajor: bool = True,
                  precision: int = 0, tile: int = 0) -> np.ndarray:
    """C = A B on the GPU's matrix cores, A: (batch, m, k) or (m, k), B: (k, n),
    shared by the batch, or (batch, k, n).  a_kmajor / b_kmajor choose the
    layout the operand is handed to the kernel in (k-contiguous or not)."""
    lib = load()
    a3 = np.asarray(a, dtype=np.float64)
    squeeze = a3.ndim == 2
    if squeeze:
        a3 = a3[None]
    batch, m, k = a3.shape
    b3 = np.asarray(b, dtype=np.float64)
    shared = b3.ndim == 2
    if shared:
        b3 = b3[None]
    if b3.shape[1] != k or (not shared and b3.shape[0] != batch):
        raise ValueError("dense_product: operand shapes do not match")
    n = b3.shape[2]
    # C order of (batch, m, k) is k-contiguous: column-major k x m per block
    ab = np.ascontiguousarray(a3 if a_kmajor else a3.transpose(0, 2, 1))
    bb = np.ascontiguousarray(b3.transpose(0, 2, 1) if b_kmajor else b3)
    cb = np.zeros((batch, n, m), dtype=np.float64)
    check(lib.hos_dense_product(precision, int(a_kmajor), int(b_kmajor), int(shared), m, n, k, batch, tile,
                                ptr(ab), ptr(bb), ptr(cb)))
    c = cb.transpose(0, 2, 1)
    return c[0] if squeeze else c


def r2r(grid_in: np.ndarray, kind_x: str, kind_y: str, forward: bool, precision: int = 0) -> np.ndarray:
    """SPACE_2_FOURIER / FOURIER_2_SPACE on the GPU for an (n1, n2) array."""
    lib = load()
    a = fortran(grid_in)
    if a.ndim == 1:
        a = a.reshape(-1, 1, order="F")
    n1, n2 = a.shape
    out = np.zeros((n1, n2), dtype=np.float64, order="F")
    kx = 0 if kind_x == "cos" else 1
    ky = 0 if kind_y == "cos" else 1
    f = lib.hos_space_2_fourier if forward else lib.hos_fourier_2_space
    check(f(n1, n2, kx, ky, precision, ptr(a), ptr(out)))
    return out
