"""ctypes binding of ``libfss_hip.so`` (the C ABI declared in ``include/fss.h``).

The library is built in-tree by :func:`fast_shepherd_score_amd.build.build_hip`.  Loading fails
loudly when it is missing: there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfss_hip.so")

# Mode / similarity ids (must match include/fss.h).
MODE_SHAPE = 0
MODE_SHAPE_ESP = 1
MODE_PHARM = 2
MODE_SHAPE_PHARM = 3
MODE_ESP_COMBO = 4
SIM_TANIMOTO = 0
SIM_TVERSKY = 1
N_PTYPES = 9
SEED_PHARM = 1
SEED_REF = 2
SEED_SURF = 4
COULOMB_SCALING = 1e4 / (4 * 55.263 * 3.141592653589793)   # score/constants.py:16

E_INVALID = 9001
E_TOO_LARGE = 9002
E_CLOSED = 9004
PENDING = 1
LIB_SHAPE = 1
LIB_PHARM = 2
LIB_SHAEP = 4
LIB_SURFACE = 8
SCREEN_MAX_TOPK = 1024
NOALLOC_SCRATCH_BYTES = 8 << 20
QPOOLS = 4                    # query-row pools: shape rows, pharmacophores, ShaEP surface, atoms
QPOOL_ROW_BYTES = (16, 36, 16, 20)

_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


class FssPairBatch(ctypes.Structure):
    _fields_ = [
        ("n_pairs", ctypes.c_int32),
        ("ref_pts", ctypes.c_void_p), ("fit_pts", ctypes.c_void_p),
        ("ref_rng", ctypes.c_void_p), ("fit_rng", ctypes.c_void_p),
        ("ref_anc", ctypes.c_void_p), ("ref_vec", ctypes.c_void_p),
        ("ref_typ", ctypes.c_void_p), ("ref_prng", ctypes.c_void_p),
        ("fit_anc", ctypes.c_void_p), ("fit_vec", ctypes.c_void_p),
        ("fit_typ", ctypes.c_void_p), ("fit_prng", ctypes.c_void_p),
        ("max_ref", ctypes.c_int32), ("max_fit", ctypes.c_int32),
        ("max_ref_ph", ctypes.c_int32), ("max_fit_ph", ctypes.c_int32),
        ("ref_srf", ctypes.c_void_p), ("fit_srf", ctypes.c_void_p),
        ("ref_srng", ctypes.c_void_p), ("fit_srng", ctypes.c_void_p),
        ("ref_atm", ctypes.c_void_p), ("fit_atm", ctypes.c_void_p),
        ("ref_rad", ctypes.c_void_p), ("fit_rad", ctypes.c_void_p),
        ("ref_arng", ctypes.c_void_p), ("fit_arng", ctypes.c_void_p),
        ("max_ref_srf", ctypes.c_int32), ("max_fit_srf", ctypes.c_int32),
        ("max_ref_atm", ctypes.c_int32), ("max_fit_atm", ctypes.c_int32),
    ]


class FssScoreParams(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int32),
        ("similarity", ctypes.c_int32),
        ("alpha", ctypes.c_double),
        ("lam", ctypes.c_double),
        ("color_weight", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("ph_alpha", ctypes.c_double * N_PTYPES),
        ("ph_k", ctypes.c_double * N_PTYPES),
        ("ph_cat", ctypes.c_int32 * N_PTYPES),
        ("esp_weight", ctypes.c_double),
        ("probe_radius", ctypes.c_double),
        ("coulomb", ctypes.c_double),
    ]


class FssAlignParams(ctypes.Structure):
    _fields_ = [
        ("n_restarts", ctypes.c_int32),
        ("max_steps", ctypes.c_int32),
        ("lr", ctypes.c_float), ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float), ("eps", ctypes.c_float),
        ("early_stop", ctypes.c_int32),
        ("es_tol", ctypes.c_float),
        ("es_patience", ctypes.c_int32),
        ("lanes_per_pose", ctypes.c_int32),
        ("restarts_per_wg", ctypes.c_int32),
        ("pair_order", ctypes.c_void_p),
        ("skip_finish", ctypes.c_int32),
        ("variant", ctypes.c_int32),
        ("groups_per_wg", ctypes.c_int32),
        ("pairs_per_wg", ctypes.c_int32),
        ("tile_rows", ctypes.c_int32),
    ]


class FssAlignPlanInfo(ctypes.Structure):
    _fields_ = [
        ("variant", ctypes.c_int32),
        ("tile_ref", ctypes.c_int32), ("tile_fit", ctypes.c_int32),
        ("block", ctypes.c_int32),
        ("grid", ctypes.c_int64),
        ("lds_bytes", ctypes.c_int64),
    ]


class FssAlignOut(ctypes.Structure):
    _fields_ = [
        ("q", ctypes.c_void_p), ("t", ctypes.c_void_p),
        ("pose_score", ctypes.c_void_p), ("best_score", ctypes.c_void_p),
        ("best_idx", ctypes.c_void_p), ("best_tf", ctypes.c_void_p),
        ("n_steps", ctypes.c_void_p),
    ]


class FssLibSpec(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32), ("channels", ctypes.c_int32),
        ("max_conformers", ctypes.c_int64), ("max_points", ctypes.c_int64),
        ("max_pharm", ctypes.c_int64), ("max_surface", ctypes.c_int64),
        ("max_atoms", ctypes.c_int64),
    ]


class FssLibRows(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in
                ("pts", "counts", "anc", "vec", "typ", "pcounts", "srf", "scounts", "atm", "rad", "acounts")]


class FssScreenParams(ctypes.Structure):
    _fields_ = [
        ("n_restarts", ctypes.c_int32), ("max_steps", ctypes.c_int32),
        ("lr", ctypes.c_float), ("beta1", ctypes.c_float),
        ("beta2", ctypes.c_float), ("eps", ctypes.c_float),
        ("early_stop", ctypes.c_int32), ("es_tol", ctypes.c_float),
        ("es_patience", ctypes.c_int32), ("top_k", ctypes.c_int32),
        ("launch_pairs", ctypes.c_int32), ("max_active", ctypes.c_int32),
        ("max_query_points", ctypes.c_int32), ("max_query_pharm", ctypes.c_int32),
        ("lanes_per_pose", ctypes.c_int32), ("restarts_per_wg", ctypes.c_int32),
        ("variant", ctypes.c_int32), ("seed_flags", ctypes.c_int32),
        ("quat_table", ctypes.c_void_p),
        ("max_query_surface", ctypes.c_int32), ("max_query_atoms", ctypes.c_int32),
        ("shape_from_surface", ctypes.c_int32),
        ("query_budget_bytes", ctypes.c_int64),
        ("pairs_per_wg", ctypes.c_int32),
        ("groups_per_wg", ctypes.c_int32),
    ]


class FssQuery(ctypes.Structure):
    _fields_ = [
        ("n_points", ctypes.c_int32), ("pts", ctypes.c_void_p),
        ("n_pharm", ctypes.c_int32), ("anc", ctypes.c_void_p),
        ("vec", ctypes.c_void_p), ("typ", ctypes.c_void_p),
        ("n_surface", ctypes.c_int32), ("srf", ctypes.c_void_p),
        ("n_atoms", ctypes.c_int32), ("atm", ctypes.c_void_p), ("rad", ctypes.c_void_p),
        ("all_scores", ctypes.c_int32),
    ]


class FssHits(ctypes.Structure):
    _fields_ = [
        ("capacity", ctypes.c_int32), ("n", ctypes.c_int32),
        ("score", ctypes.c_void_p), ("index", ctypes.c_void_p),
        ("transform", ctypes.c_void_p),
        ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double),
        ("all_scores", ctypes.c_void_p), ("n_all", ctypes.c_int64),
    ]


# name -> (restype, argtypes); every symbol include/fss.h declares.
_P = ctypes.c_void_p
SIGNATURES = {
    "fss_version": (ctypes.c_char_p, []),
    "fss_last_error": (ctypes.c_char_p, []),
    "fss_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "fss_self_overlap": (ctypes.c_int, [_P, _P, _P, _P]),
    "fss_overlap_grad": (ctypes.c_int, [_P, _P, ctypes.c_int32, _P, _P, _P, ctypes.c_int32,
                                        _P, _P, _P, _P, _P]),
    "fss_pose_loss_grad": (ctypes.c_int, [_P, _P, ctypes.c_int32, _P, ctypes.c_int32, _P,
                                          _P, _P, _P, _P, _P]),
    "fss_adam_step": (ctypes.c_int, [ctypes.c_int32, _P, _P, _P, _P, _P, ctypes.c_int32,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, _P, ctypes.c_int32, _P]),
    "fss_adam_qt": (ctypes.c_int, [ctypes.c_int32, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int32, _P]),
    "fss_early_stop_update": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _P, _P, _P, _P, _P,
                                             ctypes.c_int32, ctypes.c_float, ctypes.c_int32, _P]),
    "fss_pose_score": (ctypes.c_int, [_P, _P, ctypes.c_int32, _P, ctypes.c_int32, _P, _P, _P,
                                      _P, _P]),
    "fss_align": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "fss_align_noalloc": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "fss_align_plan": (ctypes.c_int, [_P, _P, _P, _P]),
    "fss_align_finish": (ctypes.c_int, [_P, _P, ctypes.c_int32, _P, _P, _P]),
    "fss_seed_poses": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, _P, _P, _P, _P]),
    "fss_score_fwd": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "fss_score_bwd": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "fss_score_fwd_f64": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "fss_score_bwd_f64": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fss_lib_create": (ctypes.c_int, [_P, _P]),
    "fss_lib_upload": (ctypes.c_int, [_P, ctypes.c_int32, _P, _P, _P, _P, _P, _P, _P]),
    "fss_lib_append": (ctypes.c_int, [_P, ctypes.c_int32, _P, _P]),
    "fss_lib_size": (ctypes.c_int, [_P, _P]),
    "fss_lib_destroy": (ctypes.c_int, [_P]),
    "fss_screen_create": (ctypes.c_int, [_P, _P, _P, _P]),
    "fss_screen_submit": (ctypes.c_int, [_P, _P, _P]),
    "fss_screen_poll": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32, _P]),
    "fss_screen_stats": (ctypes.c_int, [_P, _P, _P, _P]),
    "fss_screen_destroy": (ctypes.c_int, [_P]),
    "fss_plan_launch": (ctypes.c_int, [ctypes.c_int32, _P, ctypes.c_int64, _P]),
    "fss_plan_query_pools": (ctypes.c_int, [_P, _P, _P]),
    "fss_rowpool_create": (ctypes.c_int, [ctypes.c_int64, _P]),
    "fss_rowpool_alloc": (ctypes.c_int, [_P, ctypes.c_int64, _P]),
    "fss_rowpool_free": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int64]),
    "fss_rowpool_stats": (ctypes.c_int, [_P, _P, _P]),
    "fss_rowpool_destroy": (ctypes.c_int, [_P]),
    "fss_admit_fifo": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, _P, ctypes.c_int32, _P, _P]),
    "fss_screen_admission": (ctypes.c_int, [_P, _P, _P, _P]),
    "fss_screen_terms": (ctypes.c_int, [_P, _P, _P, _P]),
}

_lock = threading.Lock()
_lib = None


class FssError(RuntimeError):
    """A nonzero return code from libfss_hip.so, with the library's error text."""

    def __init__(self, fn: str, code: int, text: str):
        super().__init__(f"{fn} failed with code {code}: {text}")
        self.code = code


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libfss_hip.so once (raises if it was never built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise ImportError(
                f"{path} is missing: build it first with "
                "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc --offload-arch=gfx950)")
        lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def call(name: str, *args) -> None:
    """Call ``name`` and raise :class:`FssError` on a nonzero return code."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        text = lib.fss_last_error().decode(errors="replace")
        raise FssError(name, rc, text)


def ptr(t) -> int:
    """Device address of a tensor (0 for None)."""
    return 0 if t is None else int(t.data_ptr())


def byref(s) -> ctypes.c_void_p:
    return ctypes.cast(ctypes.pointer(s), ctypes.c_void_p)
