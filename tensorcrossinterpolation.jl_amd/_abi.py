"""ctypes mirror of include/tci.h and the loader of libtci_hip.so.

The library is built in-tree (`make -C tensorcrossinterpolation.jl_amd/csrc`,
or `__graft_entry__.build()`); importing this module never builds it.  There
is no CPU fallback: if the library or a gfx950 device is missing, the entry
points raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# the in-tree build; TCI_HIP_LIB names another build of the same library
LIB_PATH = os.environ.get("TCI_HIP_LIB") or os.path.join(PKG_DIR, "libtci_hip.so")

TCI_UNLIMITED = (1 << 63) - 1

# tci_status
STATUS = {
    0: "TCI_OK", 1: "TCI_ERR_NAN_L", 2: "TCI_ERR_NAN_U", 3: "TCI_ERR_ARG",
    4: "TCI_ERR_ZERO_MAXSAMPLE", 5: "TCI_ERR_NAN_TENSOR", 6: "TCI_ERR_NONSQUARE",
    7: "TCI_ERR_DIM", 8: "TCI_ERR_GENERIC", 9: "TCI_ERR_SINGULAR", 10: "TCI_ERR_CAPACITY",
    11: "TCI_ERR_OOM", 12: "TCI_ERR_DEVICE", 13: "TCI_ERR_PENDING",
}
TCI_OK, TCI_ERR_PENDING = 0, 13

# tci_functor_kind
F_LORENTZ_INT = 1
F_LORENTZ_WEIGHTED = 2
F_QUANTICS_GAUSS_COS = 3
F_GAUSS_SUM = 4
F_SUM = 5
F_TABLE = 6
F_TT = 7
F_QUANTICS_EXP = 8
FACTOR_LU, FACTOR_CI, FACTOR_SVD = 0, 1, 2
FLAG_NO_FINAL_SWEEP = 1        # TCI_FLAG_NO_FINAL_SWEEP


class tci_functor(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nparams", C.c_int32), ("params", C.POINTER(C.c_double)),
                ("niparams", C.c_int32), ("iparams", C.POINTER(C.c_int32))]


class tci_options(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("maxbonddim", C.c_int64), ("maxiter", C.c_int32),
                ("sweepstrategy", C.c_int32), ("normalizeerror", C.c_int32),
                ("ncheckhistory", C.c_int32), ("maxnglobalpivot", C.c_int32),
                ("nsearchglobalpivot", C.c_int32), ("tolmarginglobalsearch", C.c_double),
                ("strictlynested", C.c_int32), ("checkconvglobalpivot", C.c_int32),
                ("seed", C.c_uint64), ("chi_cap", C.c_int32), ("flags", C.c_int32),
                ("cache_log2", C.c_int32), ("pivotsearch", C.c_int32)]


class tci_state(C.Structure):
    _fields_ = [("nI", C.POINTER(C.c_int32)), ("Iset", C.POINTER(C.c_uint8)),
                ("nJ", C.POINTER(C.c_int32)), ("Jset", C.POINTER(C.c_uint8)),
                ("nI_hist", C.POINTER(C.c_int32)), ("Iset_hist", C.POINTER(C.c_uint8)),
                ("nJ_hist", C.POINTER(C.c_int32)), ("Jset_hist", C.POINTER(C.c_uint8)),
                ("maxsamplevalue", C.c_double)]


class tci_problem(C.Structure):
    _fields_ = [("L", C.c_int32), ("localdims", C.POINTER(C.c_int32)), ("f", tci_functor),
                ("npivots", C.c_int32), ("initialpivots", C.POINTER(C.c_int32)),
                ("opts", tci_options), ("state", C.POINTER(tci_state))]


def _pack_sets(sets, L, iside):
    """L sets of multi-indices → (counts int32[L], bytes back to back)."""
    counts = np.array([len(s) for s in sets], np.int32)
    chunks = []
    for b, s in enumerate(sets):
        ln = b if iside else L - 1 - b
        if len(s) and ln:
            chunks.append(np.asarray(s, np.uint8).reshape(len(s), ln).ravel())
    flat = np.concatenate(chunks) if chunks else np.zeros(1, np.uint8)
    return np.ascontiguousarray(counts), np.ascontiguousarray(flat)


class StateHolder:
    """Owns the buffers of a tci_state: the Iset/Jset of every site, the last
    history entry (or none) and maxsamplevalue (0: computed at the start)."""

    def __init__(self, L, Iset, Jset, Iset_hist=None, Jset_hist=None, maxsamplevalue=0.0):
        self.nI, self.I = _pack_sets(Iset, L, True)
        self.nJ, self.J = _pack_sets(Jset, L, False)
        self.st = tci_state(i32ptr(self.nI), u8ptr(self.I), i32ptr(self.nJ), u8ptr(self.J),
                            None, None, None, None, float(maxsamplevalue))
        if Iset_hist is not None:
            self.nIh, self.Ih = _pack_sets(Iset_hist, L, True)
            self.nJh, self.Jh = _pack_sets(Jset_hist, L, False)
            self.st.nI_hist, self.st.Iset_hist = i32ptr(self.nIh), u8ptr(self.Ih)
            self.st.nJ_hist, self.st.Jset_hist = i32ptr(self.nJh), u8ptr(self.Jh)


class tci_result(C.Structure):
    _fields_ = [("status", C.c_int32), ("niter", C.c_int32), ("L", C.c_int32),
                ("npivoterrors", C.c_int32), ("maxsamplevalue", C.c_double),
                ("workstats", C.c_double * 6)]


class tci_rrlu_tune(C.Structure):
    _fields_ = [("side", C.c_int32), ("tile", C.c_int32), ("grid_team", C.c_int32), ("grid_max_jobs", C.c_int32),
                ("ms_stream", C.c_float), ("ms_grid", C.c_float), ("ms_tile_batch", C.c_float),
                ("ms_stream_batch", C.c_float)]


class tci_lu_job(C.Structure):
    _fields_ = [("A", C.c_void_p), ("rowperm", C.c_void_p), ("colperm", C.c_void_p),
                ("m", C.c_int64), ("n", C.c_int64), ("maxrank", C.c_int64),
                ("reltol", C.c_double), ("abstol", C.c_double), ("leftorth", C.c_int32),
                ("status", C.c_int32), ("npivot", C.c_int64), ("error", C.c_double)]


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def i64ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def i32ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def u8ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


SWEEP = {"backandforth": 0, "forward": 1, "backward": 2}
PIVOTSEARCH = {"full": 0, "rook": 1}


def make_options(tolerance=1e-8, maxbonddim=TCI_UNLIMITED, maxiter=20, sweepstrategy="backandforth",
                 normalizeerror=True, ncheckhistory=3, maxnglobalpivot=5, nsearchglobalpivot=5,
                 tolmarginglobalsearch=10.0, strictlynested=False, checkconvglobalpivot=True,
                 seed=0, chi_cap=0, flags=0, cache_log2=0, pivotsearch="full"):
    if isinstance(sweepstrategy, str):
        sweepstrategy = SWEEP[sweepstrategy]
    if isinstance(pivotsearch, str):
        pivotsearch = PIVOTSEARCH[pivotsearch.lstrip(":")]
    return tci_options(float(tolerance), int(maxbonddim), int(maxiter), int(sweepstrategy),
                       int(bool(normalizeerror)), int(ncheckhistory), int(maxnglobalpivot),
                       int(nsearchglobalpivot), float(tolmarginglobalsearch), int(bool(strictlynested)),
                       int(bool(checkconvglobalpivot)), int(seed) & ((1 << 64) - 1), int(chi_cap), int(flags),
                       int(cache_log2), int(pivotsearch))


class ProblemHolder:
    """Owns the numpy buffers a tci_problem points into."""

    def __init__(self, localdims, kind, params=(), iparams=(), initialpivots=None, state=None, **opts):
        """state: a StateHolder to start from instead of initialpivots."""
        self.state = state
        self.localdims = np.ascontiguousarray(localdims, np.int32)
        self.L = len(self.localdims)
        self.params = np.ascontiguousarray(np.asarray(params, np.float64).ravel())
        self.iparams = np.ascontiguousarray(np.asarray(iparams, np.int32).ravel())
        if initialpivots is None:
            initialpivots = [[1] * self.L]
        self.pivots = np.ascontiguousarray(np.asarray(initialpivots, np.int32).reshape(-1, self.L))
        f = tci_functor(int(kind), len(self.params), dptr(self.params) if len(self.params) else None,
                        len(self.iparams), i32ptr(self.iparams) if len(self.iparams) else None)
        self.st = tci_problem(self.L, i32ptr(self.localdims), f, self.pivots.shape[0],
                              i32ptr(self.pivots), make_options(**opts),
                              C.pointer(state.st) if state is not None else None)


_LIB = None


def lib():
    """Load libtci_hip.so (built in-tree).  Raises if it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise OSError(f"libtci_hip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    L.tci_last_error.restype = C.c_char_p
    L.tci_version.restype = C.c_char_p
    L.tci_ctx_create.argtypes = [P(vp), C.c_int32, C.c_uint64]
    L.tci_ctx_destroy.argtypes = [vp]
    L.tci_ctx_synchronize.argtypes = [vp]
    L.tci_ctx_set_streams.argtypes = [vp, C.c_int32]
    L.tci_ctx_stream.argtypes = [vp, C.POINTER(vp)]
    L.tci_rrlu_f64.argtypes = [vp, P(C.c_double), C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                               C.c_double, C.c_double, C.c_int32, P(C.c_int64), P(C.c_int64),
                               P(C.c_double), C.c_int64, P(C.c_double), C.c_int64, P(C.c_int64),
                               P(C.c_double)]
    L.tci_luci_f64.argtypes = L.tci_rrlu_f64.argtypes
    L.tci_rrlu_batched_f64.argtypes = [vp, vp, C.c_int64]
    L.tci_pi_eval_f64.argtypes = [vp, P(tci_functor), C.c_int32, P(C.c_int32), P(C.c_uint8),
                                  C.c_int64, C.c_int32, P(C.c_uint8), C.c_int64, C.c_int32,
                                  P(C.c_double), P(C.c_double)]
    L.tci_options_default.argtypes = [P(tci_options)]
    L.tci_options_default.restype = None
    L.tci_solve_submit.argtypes = [vp, C.c_int32, P(tci_problem), P(vp)]
    L.tci_batch_query.argtypes = [vp]
    L.tci_batch_wait.argtypes = [vp]
    L.tci_batch_result.argtypes = [vp, C.c_int32, P(tci_result)]
    L.tci_batch_history.argtypes = [vp, C.c_int32, P(C.c_int64), P(C.c_double), P(C.c_int64)]
    L.tci_batch_linkdims.argtypes = [vp, C.c_int32, P(C.c_int64)]
    L.tci_batch_indexset.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, P(C.c_int32),
                                     P(C.c_uint8), C.c_int64]
    L.tci_batch_sitetensor.argtypes = [vp, C.c_int32, C.c_int32, P(C.c_int64), P(C.c_double),
                                       C.c_int64]
    L.tci_batch_pivoterrors.argtypes = [vp, C.c_int32, P(C.c_double), C.c_int64]
    L.tci_batch_bonderrors.argtypes = [vp, C.c_int32, P(C.c_double)]
    L.tci_batch_free.argtypes = [vp]
    L.tci_batch_elapsed_ms.argtypes = [vp, P(C.c_float)]
    L.tci_batch_profile.argtypes = [vp, C.c_int32, P(C.c_double)]
    L.tci_capacity_plan.argtypes = [P(tci_problem), P(C.c_int64), P(C.c_int64)]
    L.tci_ctx_set_rrlu_algo.argtypes = [vp, C.c_int32]
    L.tci_ctx_autotune_rrlu.argtypes = [vp, P(tci_rrlu_tune), C.c_int32, P(C.c_int32)]
    L.tci_ctx_set_rrlu_tuning.argtypes = [vp, P(tci_rrlu_tune), C.c_int32]
    L.tci_ctx_autotune_solve.argtypes = [vp, P(tci_problem), C.c_int32, P(C.c_float), P(C.c_int32)]
    L.tci_ctx_set_solve_tuning.argtypes = [vp, P(C.c_int32), P(C.c_int32), P(C.c_int32), C.c_int32]
    L.tci_batch_cachestats.argtypes = [vp, C.c_int32, P(C.c_int64), P(C.c_int64), P(C.c_int64)]
    L.tci_batch_cachedata.argtypes = [vp, C.c_int32, C.c_int64, P(C.c_uint8), P(C.c_double), P(C.c_int64)]
    L.tci_rrlu_batched_noalloc_f64.argtypes = [vp, vp, C.c_int64]
    L.tci_tt_evaluate_f64.argtypes = [vp, C.c_int32, P(C.c_int64), P(C.c_int32), P(C.c_double), C.c_int64,
                                      P(C.c_uint8), P(C.c_double)]
    L.tci_batch_evaluate.argtypes = [vp, C.c_int32, C.c_int64, P(C.c_uint8), P(C.c_double), P(C.c_double),
                                     P(C.c_double), P(C.c_double)]
    L.tci_gemm_f64.argtypes = [vp, C.c_int64, C.c_int64, C.c_int64, P(C.c_double), C.c_int64, P(C.c_double),
                               C.c_int64, P(C.c_double), C.c_int64]
    L.tci_tt_compress_f64.argtypes = [vp, C.c_int32, P(C.c_int64), P(C.c_int32), P(C.c_double), C.c_int32,
                                      C.c_double, C.c_int64, C.c_int32]
    L.tci_tt_sweep_f64.argtypes = [vp, C.c_int32, P(C.c_int64), P(C.c_int32), P(C.c_double), C.c_int32, C.c_int32,
                                   C.c_int64, C.c_double, C.c_double, P(C.c_int32), P(C.c_double)]
    L.tci_server_start.argtypes = [vp, P(tci_problem), C.c_int32, C.c_int32, C.c_int32, P(vp)]
    L.tci_server_submit.argtypes = [vp, P(tci_problem), P(vp)]
    L.tci_server_stop.argtypes = [vp]
    L.tci_server_poll.argtypes = [vp, P(vp), C.c_int32, P(C.c_int32)]
    L.tci_server_submit_many.argtypes = [vp, C.c_int32, P(tci_problem), P(vp), P(C.c_int32)]
    L.tci_tt_floatingzone_f64.argtypes = [vp, P(tci_functor), C.c_int32, P(C.c_int64), P(C.c_int32), P(C.c_double),
                                          C.c_int64, P(C.c_uint8), P(C.c_double), C.c_int64, C.c_double]
    _LIB = L
    return L


EXPORTED_SYMBOLS = [
    "tci_last_error", "tci_version", "tci_ctx_create", "tci_ctx_destroy", "tci_ctx_synchronize",
    "tci_ctx_set_streams", "tci_ctx_stream",
    "tci_rrlu_f64", "tci_luci_f64", "tci_rrlu_batched_f64", "tci_pi_eval_f64", "tci_options_default",
    "tci_solve_submit", "tci_batch_query", "tci_batch_wait", "tci_batch_result",
    "tci_batch_history", "tci_batch_linkdims", "tci_batch_indexset", "tci_batch_sitetensor",
    "tci_batch_pivoterrors", "tci_batch_bonderrors", "tci_batch_free", "tci_batch_elapsed_ms",
    "tci_batch_profile", "tci_capacity_plan", "tci_tt_evaluate_f64", "tci_batch_evaluate",
    "tci_ctx_set_rrlu_algo", "tci_ctx_autotune_rrlu", "tci_ctx_set_rrlu_tuning", "tci_ctx_autotune_solve", "tci_ctx_set_solve_tuning", "tci_rrlu_batched_noalloc_f64", "tci_batch_cachestats", "tci_batch_cachedata",
    "tci_gemm_f64", "tci_tt_compress_f64", "tci_tt_sweep_f64", "tci_tt_floatingzone_f64",
    "tci_server_start", "tci_server_submit", "tci_server_submit_many", "tci_server_poll",
    "tci_server_stop",
]
