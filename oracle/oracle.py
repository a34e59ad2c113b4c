"""ctypes wrapper of the CPU oracle (libtci_oracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP product:
only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  The product package never imports it.

The restated reference functions are cited in tci_oracle.c; this module only
marshals arguments and mirrors the reference's Julia return values
(1-based permutations, column-major matrices).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

INT64_MAX = (1 << 63) - 1

STATUS_NAMES = {
    0: "ok", 1: "lu.L contains NaNs", 2: "lu.U contains NaNs", 3: "ArgumentError",
    4: "maxsamplevalue is zero!", 5: "NaN in tensor", 6: "Pivot matrix is not square",
    7: "DimensionMismatch", 8: "ErrorException", 9: "SingularException",
}


class OracleError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_CB = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_int32), C.c_int32, C.c_void_p)


class _Functor(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nparams", C.c_int32), ("params", C.POINTER(C.c_double)),
                ("niparams", C.c_int32), ("iparams", C.POINTER(C.c_int32)),
                ("callback", _CB), ("user", C.c_void_p)]


class _Opts(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("maxbonddim", C.c_int64), ("maxiter", C.c_int32),
                ("sweepstrategy", C.c_int32), ("normalizeerror", C.c_int32),
                ("ncheckhistory", C.c_int32), ("maxnglobalpivot", C.c_int32),
                ("nsearchglobalpivot", C.c_int32), ("tolmarginglobalsearch", C.c_double),
                ("strictlynested", C.c_int32), ("checkconvglobalpivot", C.c_int32),
                ("seed", C.c_uint64), ("pivotsearch", C.c_int32)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libtci_oracle.so")
        if not os.path.exists(path):
            raise OSError(f"oracle library not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        P = C.POINTER
        L.orc_last_error.restype = C.c_char_p
        L.orc_rrlu.argtypes = [P(C.c_double), C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_double,
                               C.c_int32, P(C.c_int64), P(C.c_int64), P(C.c_double), P(C.c_double),
                               C.c_int64, P(C.c_int64), P(C.c_double)]
        L.orc_submatrixargmax.argtypes = [P(C.c_double), C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                          P(C.c_int64), P(C.c_int64)]
        for nm in ("orc_luci_left", "orc_luci_right"):
            getattr(L, nm).argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int32, P(C.c_int64),
                                       P(C.c_int64), P(C.c_double), P(C.c_double), C.c_int64,
                                       P(C.c_double)]
        L.orc_batcheval.argtypes = [P(_Functor), C.c_int32, P(C.c_int32), P(C.c_int32), C.c_int64,
                                    C.c_int32, P(C.c_int32), C.c_int64, C.c_int32, P(C.c_double)]
        L.orc_tci_new.restype = C.c_void_p
        L.orc_tci_new.argtypes = [C.c_int32, P(C.c_int32)]
        L.orc_tci_free.argtypes = [C.c_void_p]
        L.orc_tci_init.argtypes = [C.c_void_p, P(_Functor), C.c_int32, P(C.c_int32)]
        L.orc_addglobalpivots.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32)]
        L.orc_updatepivots.argtypes = [C.c_void_p, C.c_int32, P(_Functor), C.c_int32, C.c_double,
                                       C.c_double, C.c_int64]
        L.orc_sweep1site.argtypes = [C.c_void_p, P(_Functor), C.c_int32, C.c_double, C.c_double,
                                     C.c_int64, C.c_int32]
        L.orc_sweep2site.argtypes = [C.c_void_p, P(_Functor), C.c_int32, C.c_double, C.c_int64,
                                     C.c_int32, C.c_int32, C.c_int32]
        L.orc_fillsitetensors.argtypes = [C.c_void_p, P(_Functor)]
        L.orc_optimize.argtypes = [C.c_void_p, P(_Functor), P(_Opts)]
        L.orc_tci_len.argtypes = [C.c_void_p]
        L.orc_tci_setsize.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.orc_tci_getset.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(C.c_int32)]
        L.orc_tci_setset.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, P(C.c_int32)]
        L.orc_tci_sitetensor_dims.argtypes = [C.c_void_p, C.c_int32, P(C.c_int64)]
        L.orc_tci_sitetensor.argtypes = [C.c_void_p, C.c_int32, P(C.c_double)]
        L.orc_tci_maxsample.restype = C.c_double
        L.orc_tci_maxsample.argtypes = [C.c_void_p]
        L.orc_tci_set_maxsample.argtypes = [C.c_void_p, C.c_double]
        L.orc_tci_npivoterrors.argtypes = [C.c_void_p]
        L.orc_tci_pivoterrors.argtypes = [C.c_void_p, P(C.c_double)]
        L.orc_tci_bonderrors.argtypes = [C.c_void_p, P(C.c_double)]
        L.orc_tci_niter.argtypes = [C.c_void_p]
        L.orc_tci_history.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_double), P(C.c_int64)]
        L.orc_tci_evaluate.restype = C.c_double
        L.orc_tci_evaluate.argtypes = [C.c_void_p, P(C.c_int32)]
        L.orc_tci_workstats.argtypes = [C.c_void_p, P(C.c_double)]
        L.orc_convergencecriterion.argtypes = [C.c_int32, P(C.c_int64), P(C.c_double), P(C.c_int64),
                                               C.c_double, C.c_int64, C.c_int32, C.c_int32]
        L.orc_eval_points.argtypes = [C.c_void_p, C.c_int32, C.c_int64, P(C.c_int32), P(C.c_double)]
        L.orc_tt_evaluate.argtypes = [C.c_int32, P(C.c_int64), P(C.c_double), C.c_int64, P(C.c_int32),
                                      P(C.c_double)]
        L.orc_tt_sweep.argtypes = [C.c_int32, P(C.c_int64), P(C.c_int32), P(C.c_double), C.c_int32, C.c_int32,
                                   C.c_int64, C.c_double, C.c_double, P(C.c_int32), P(C.c_double)]
        L.orc_tt_compress.argtypes = [C.c_int32, P(C.c_int64), P(C.c_int32), P(C.c_double), C.c_int32,
                                      C.c_double, C.c_int64, C.c_int32]
        L.orc_arrlu.argtypes = [P(C.c_double), C.c_int64, C.c_int64, P(C.c_int64), C.c_int64, P(C.c_int64),
                                C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_uint64,
                                P(C.c_int64), P(C.c_int64), P(C.c_double), P(C.c_double), P(C.c_int64),
                                P(C.c_double)]
        L.orc_floatingzone.argtypes = [C.c_void_p, C.c_int32, P(C.c_int64), P(C.c_double), C.c_int64, P(C.c_int32),
                                       P(C.c_double), C.c_int64, C.c_double]
        L.orc_globalpivotfinder.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_int32, C.c_int32,
                                            C.c_double, C.c_uint64, P(C.c_int32)]
        L.orc_exp.restype = C.c_double
        L.orc_exp.argtypes = [C.c_double]
        L.orc_cos.restype = C.c_double
        L.orc_cos.argtypes = [C.c_double]
        L.orc_rng_seed.argtypes = [P(C.c_uint64), C.c_uint64]
        L.orc_rng_next.restype = C.c_uint64
        L.orc_rng_next.argtypes = [P(C.c_uint64)]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def _i32(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _check(code):
    if code != 0:
        raise OracleError(code, lib().orc_last_error().decode())


# --------------------------------------------------------------------------
# rrLU
# --------------------------------------------------------------------------
@dataclass
class RRLU:
    rowpermutation: np.ndarray   # 1-based, length m
    colpermutation: np.ndarray   # 1-based, length n
    L: np.ndarray                # m×p
    U: np.ndarray                # p×n
    leftorthogonal: bool
    npivot: int
    error: float


def submatrixargmax(A, k=0):
    A = np.asfortranarray(A, dtype=np.float64)
    r, c = C.c_int64(), C.c_int64()
    lib().orc_submatrixargmax(_dp(A), A.shape[0], A.shape[0], A.shape[1], k, C.byref(r), C.byref(c))
    return r.value, c.value


def rrlu(A, maxrank=INT64_MAX, reltol=1e-14, abstol=0.0, leftorthogonal=True) -> RRLU:
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    cap = max(min(m, n), 1)
    rp = np.zeros(max(m, 1), np.int64)
    cp = np.zeros(max(n, 1), np.int64)
    Lb = np.zeros((max(m, 1), cap), order="F")
    Ub = np.zeros((cap, max(n, 1)), order="F")
    p = C.c_int64()
    err = C.c_double()
    st = lib().orc_rrlu(_dp(A), m, n, int(maxrank), float(reltol), float(abstol), int(bool(leftorthogonal)),
                        _ip(rp), _ip(cp), _dp(Lb), _dp(Ub), cap, C.byref(p), C.byref(err))
    _check(st)
    pv = p.value
    return RRLU(rp[:m].copy(), cp[:n].copy(), np.array(Lb[:m, :pv]), np.array(Ub[:pv, :n]),
                bool(leftorthogonal), pv, err.value)


def arrlu(A, I0=(), J0=(), maxrank=INT64_MAX, reltol=1e-14, abstol=0.0, leftorthogonal=True, numrookiter=5,
          seed=0) -> RRLU:
    """arrlu (matrixlu.jl:227-293) of an explicit matrix, rook pivot search
    from the 1-based starting rows I0 and columns J0, randomsubset drawn
    from the project PRNG seeded with `seed`."""
    I0_in, J0_in = I0, J0
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    cap = max(min(m, n), 1)
    nI, nJ = len(I0), len(J0)
    I0 = np.zeros(max(m, 1), np.int64)
    J0 = np.zeros(max(n, 1), np.int64)
    I0[:nI] = np.asarray(I0_in, np.int64).reshape(-1)
    J0[:nJ] = np.asarray(J0_in, np.int64).reshape(-1)
    rp = np.zeros(max(m, 1), np.int64)
    cp = np.zeros(max(n, 1), np.int64)
    Lb = np.zeros(max(m * cap, 1))
    Ub = np.zeros(max(cap * n, 1))
    p = C.c_int64()
    err = C.c_double()
    _check(lib().orc_arrlu(_dp(A), m, n, _ip(I0), nI, _ip(J0), nJ, int(min(maxrank, INT64_MAX)), float(reltol), float(abstol),
                           int(bool(leftorthogonal)), int(numrookiter), int(seed), _ip(rp), _ip(cp), _dp(Lb), _dp(Ub),
                           C.byref(p), C.byref(err)))
    k = p.value
    return RRLU(rp[:m].copy(), cp[:n].copy(), Lb[:m * k].reshape((m, k), order="F").copy(),
                Ub[:k * n].reshape((k, n), order="F").copy(), bool(leftorthogonal), k, err.value)


def luci_factors(A, **kw):
    """MatrixLUCI left/right (matrixluci.jl:70-84)."""
    lu = rrlu(A, **kw)
    A = np.asarray(A)
    m, n = A.shape
    p = lu.npivot
    cap = max(p, 1)
    Lb = np.zeros((m, cap), order="F")
    Ub = np.zeros((cap, n), order="F")
    Lb[:, :p] = lu.L
    Ub[:p, :] = lu.U
    left = np.zeros((m, cap), order="F")
    right = np.zeros((cap, n), order="F")
    rp = np.ascontiguousarray(lu.rowpermutation, np.int64)
    cpm = np.ascontiguousarray(lu.colpermutation, np.int64)
    lib().orc_luci_left(m, n, p, int(lu.leftorthogonal), _ip(rp), _ip(cpm), _dp(Lb), _dp(Ub), cap, _dp(left))
    lib().orc_luci_right(m, n, p, int(lu.leftorthogonal), _ip(rp), _ip(cpm), _dp(Lb), _dp(Ub), cap, _dp(right))
    return lu, np.array(left[:, :p]), np.array(right[:p, :])


# --------------------------------------------------------------------------
# functors and TCI
# --------------------------------------------------------------------------
class Functor:
    """kind + params (see include/tci_functors.h) or a Python callback."""

    def __init__(self, kind=0, params=(), iparams=(), callback=None):
        self.kind = int(kind)
        self.params = np.ascontiguousarray(np.asarray(params, dtype=np.float64).ravel())
        self.iparams = np.ascontiguousarray(np.asarray(iparams, dtype=np.int32).ravel())
        if callback is not None:
            def _cb(x, L, _u, _f=callback):
                return float(_f([x[i] for i in range(L)]))
            self._cb = _CB(_cb)
        else:
            self._cb = _CB()
        self.st = _Functor(self.kind, len(self.params),
                           _dp(self.params) if len(self.params) else None,
                           len(self.iparams), _i32(self.iparams) if len(self.iparams) else None,
                           self._cb, None)


def batcheval(f: Functor, localdims, left, right):
    localdims = np.ascontiguousarray(localdims, np.int32)
    L = len(localdims)
    left = np.ascontiguousarray(np.asarray(left, np.int32).reshape(len(left), -1))
    right = np.ascontiguousarray(np.asarray(right, np.int32).reshape(len(right), -1))
    nl, nr = left.shape[1], right.shape[1]
    M = L - nl - nr
    shape = (left.shape[0],) + tuple(int(d) for d in localdims[nl:nl + M]) + (right.shape[0],)
    out = np.zeros(int(np.prod(shape)) or 1)
    lib().orc_batcheval(C.byref(f.st), L, _i32(localdims), _i32(left), left.shape[0], nl,
                        _i32(right), right.shape[0], nr, _dp(out))
    return out[:int(np.prod(shape))].reshape(shape, order="F")


def eval_points(f: Functor, points):
    """f at each row of points (npts×L, 1-based), host FP64."""
    x = np.ascontiguousarray(np.asarray(points, np.int32))
    out = np.zeros(max(x.shape[0], 1))
    lib().orc_eval_points(C.byref(f.st), x.shape[1], x.shape[0], _i32(x), _dp(out))
    return out[:x.shape[0]]


def tt_evaluate(cores, points):
    """evaluate(tt, x) (abstracttensortrain.jl:124-132) at each row of points."""
    L = len(cores)
    dims = np.ascontiguousarray(np.array([c.shape for c in cores], np.int64).ravel())
    flat = np.ascontiguousarray(np.concatenate([np.asarray(c, np.float64).ravel(order="F") for c in cores]))
    x = np.ascontiguousarray(np.asarray(points, np.int32))
    out = np.zeros(max(x.shape[0], 1))
    lib().orc_tt_evaluate(L, _ip(dims), _dp(flat), x.shape[0], _i32(x), _dp(out))
    return out[:x.shape[0]]


def floatingzone(f: Functor, cores, points, nsweeps=INT64_MAX, earlystoptol=float("inf")):
    """_floatingzone (globalsearch.jl:43-95) from each row of points.
    Returns (pivots npts×L, errors)."""
    L = len(cores)
    dims = np.ascontiguousarray(np.array([c.shape for c in cores], np.int64).ravel())
    flat = np.ascontiguousarray(np.concatenate([np.asarray(c, np.float64).ravel(order="F") for c in cores]))
    x = np.array(np.asarray(points, np.int32).reshape(-1, L), order="C")
    errs = np.zeros(max(x.shape[0], 1))
    lib().orc_floatingzone(C.byref(f.st), L, _ip(dims), _dp(flat), x.shape[0], _i32(x), _dp(errs), int(nsweeps),
                           float(earlystoptol))
    return x, errs[:x.shape[0]]


def _tt_pack(cores):
    links = np.array([cores[0].shape[0]] + [c.shape[2] for c in cores], np.int64)
    ld = np.array([c.shape[1] for c in cores], np.int32)
    flat = np.ascontiguousarray(np.concatenate([np.asarray(c, np.float64).ravel(order="F") for c in cores]))
    return links, ld, flat


def _tt_unpack(links, ld, flat):
    out, off = [], 0
    for b in range(len(ld)):
        sz = int(links[b] * ld[b] * links[b + 1])
        out.append(flat[off:off + sz].reshape((int(links[b]), int(ld[b]), int(links[b + 1])), order="F").copy())
        off += sz
    return out


_METHODS = {"LU": 0, "CI": 1}


def tt_sweep(cores, method="CI", forward=True, maxrank=INT64_MAX, reltol=1e-14, abstol=0.0):
    """One factorize-and-carry sweep over a TT (tensortrain.jl:147-183,
    conversion.jl:73-156).  Returns (cores, (rowindices, colindices) per bond,
    pivoterrors)."""
    links, ld, flat = _tt_pack(cores)
    L = len(cores)
    R = int(max(links[1:L])) if L > 1 else 0
    piv = np.zeros(max(2 * int(links[1:L].sum()), 1), np.int32)
    pe = np.zeros(R + 1)
    _check(lib().orc_tt_sweep(L, _ip(links), _i32(ld), _dp(flat), _METHODS[method], int(bool(forward)),
                              int(maxrank), float(reltol), float(abstol), _i32(piv), _dp(pe)))
    pivots, q = [], 0
    for b in range(L - 1):
        p = int(links[b + 1])
        pivots.append((piv[q:q + p].astype(np.int64).copy(), piv[q + p:q + 2 * p].astype(np.int64).copy()))
        q += 2 * p
    return _tt_unpack(links, ld, flat), pivots, pe


def tt_compress(cores, method="LU", tolerance=1e-12, maxbonddim=INT64_MAX, normalizeerror=True):
    """compress!(tt, method; tolerance, maxbonddim, normalizeerror) (tensortrain.jl:147-183)."""
    links, ld, flat = _tt_pack(cores)
    _check(lib().orc_tt_compress(len(cores), _ip(links), _i32(ld), _dp(flat), _METHODS[method],
                                 float(tolerance), int(maxbonddim), int(bool(normalizeerror))))
    return _tt_unpack(links, ld, flat)


def make_opts(tolerance=1e-8, maxbonddim=INT64_MAX, maxiter=20, sweepstrategy="backandforth",
              normalizeerror=True, ncheckhistory=3, maxnglobalpivot=5, nsearchglobalpivot=5,
              tolmarginglobalsearch=10.0, strictlynested=False, checkconvglobalpivot=True, seed=0,
              pivotsearch="full"):
    ss = {"backandforth": 0, "forward": 1, "backward": 2}[sweepstrategy]
    ps = {"full": 0, "rook": 1}[pivotsearch]
    return _Opts(float(tolerance), int(maxbonddim), int(maxiter), ss, int(normalizeerror),
                 int(ncheckhistory), int(maxnglobalpivot), int(nsearchglobalpivot),
                 float(tolmarginglobalsearch), int(strictlynested), int(checkconvglobalpivot), int(seed), ps)


class TCI:
    """Oracle TensorCI2 (tensorci2.jl:6-53)."""

    def __init__(self, localdims):
        self.localdims = np.ascontiguousarray(localdims, np.int32)
        self.L = len(self.localdims)
        self.h = lib().orc_tci_new(self.L, _i32(self.localdims))
        if not self.h:
            raise OracleError(8, lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tci_free(self.h)
            self.h = None

    def init(self, f: Functor, pivots):
        piv = np.ascontiguousarray(np.asarray(pivots, np.int32).reshape(-1, self.L))
        _check(lib().orc_tci_init(self.h, C.byref(f.st), piv.shape[0], _i32(piv)))

    def addglobalpivots(self, pivots):
        piv = np.ascontiguousarray(np.asarray(pivots, np.int32).reshape(-1, self.L))
        _check(lib().orc_addglobalpivots(self.h, piv.shape[0], _i32(piv)))

    def updatepivots(self, b0, f, leftorth=True, reltol=1e-14, abstol=0.0, maxbonddim=INT64_MAX):
        _check(lib().orc_updatepivots(self.h, b0, C.byref(f.st), int(leftorth), reltol, abstol, int(maxbonddim)))

    def sweep1site(self, f, forward=True, reltol=1e-14, abstol=0.0, maxbonddim=INT64_MAX, updatetensors=True):
        _check(lib().orc_sweep1site(self.h, C.byref(f.st), int(forward), reltol, abstol, int(maxbonddim),
                                    int(updatetensors)))

    def sweep2site(self, f, niter=2, abstol=1e-8, maxbonddim=INT64_MAX, sweepstrategy=0,
                   strictlynested=False, fillsitetensors=True):
        _check(lib().orc_sweep2site(self.h, C.byref(f.st), niter, abstol, int(maxbonddim), sweepstrategy,
                                    int(strictlynested), int(fillsitetensors)))

    def globalpivotfinder(self, f, abstol, nsearch=5, maxnglobalpivot=5, tolmargin=10.0, seed=0):
        """DefaultGlobalPivotFinder (globalpivotfinder.jl:143-195) on the current state."""
        out = np.zeros((max(nsearch, 1), self.L), np.int32)
        n = lib().orc_globalpivotfinder(self.h, C.byref(f.st), float(abstol), int(nsearch),
                                        int(maxnglobalpivot), float(tolmargin), int(seed), _i32(out))
        return [list(map(int, out[i])) for i in range(n)]

    def optimize(self, f, **kw):
        o = make_opts(**kw)
        _check(lib().orc_optimize(self.h, C.byref(f.st), C.byref(o)))
        return self.ranks, self.errors

    # ---- state ----
    def Iset(self, b0):
        n = lib().orc_tci_setsize(self.h, 0, b0)
        out = np.zeros((max(n, 1), max(b0, 1)), np.int32)
        lib().orc_tci_getset(self.h, 0, b0, _i32(out))
        return [list(map(int, out[i, :b0])) for i in range(n)]

    def Jset(self, b0):
        n = lib().orc_tci_setsize(self.h, 1, b0)
        ln = self.L - 1 - b0
        out = np.zeros((max(n, 1), max(ln, 1)), np.int32)
        lib().orc_tci_getset(self.h, 1, b0, _i32(out))
        return [list(map(int, out[i, :ln])) for i in range(n)]

    def setIset(self, b0, entries):
        a = np.ascontiguousarray(np.asarray(entries, np.int32).reshape(len(entries), -1)) if entries else np.zeros((0, 1), np.int32)
        lib().orc_tci_setset(self.h, 0, b0, len(entries), _i32(a))

    def setJset(self, b0, entries):
        a = np.ascontiguousarray(np.asarray(entries, np.int32).reshape(len(entries), -1)) if entries else np.zeros((0, 1), np.int32)
        lib().orc_tci_setset(self.h, 1, b0, len(entries), _i32(a))

    def linkdims(self):
        return [lib().orc_tci_setsize(self.h, 0, b) for b in range(1, self.L)]

    def rank(self):
        return max(self.linkdims())

    def sitetensor(self, b0):
        d = np.zeros(3, np.int64)
        lib().orc_tci_sitetensor_dims(self.h, b0, _ip(d))
        out = np.zeros(int(np.prod(d)) or 1)
        lib().orc_tci_sitetensor(self.h, b0, _dp(out))
        return out[:int(np.prod(d))].reshape(tuple(d), order="F")

    def sitetensors(self):
        return [self.sitetensor(b) for b in range(self.L)]

    @property
    def maxsamplevalue(self):
        return lib().orc_tci_maxsample(self.h)

    @property
    def pivoterrors(self):
        n = lib().orc_tci_npivoterrors(self.h)
        out = np.zeros(max(n, 1))
        lib().orc_tci_pivoterrors(self.h, _dp(out))
        return out[:n].copy()

    @property
    def bonderrors(self):
        out = np.zeros(self.L - 1)
        lib().orc_tci_bonderrors(self.h, _dp(out))
        return out

    @property
    def niter(self):
        return lib().orc_tci_niter(self.h)

    @property
    def ranks(self):
        n = self.niter
        r = np.zeros(max(n, 1), np.int64)
        lib().orc_tci_history(self.h, _ip(r), None, None)
        return r[:n].copy()

    @property
    def errors(self):
        n = self.niter
        e = np.zeros(max(n, 1))
        lib().orc_tci_history(self.h, None, _dp(e), None)
        return e[:n].copy()

    @property
    def nglobalpivots(self):
        n = self.niter
        g = np.zeros(max(n, 1), np.int64)
        lib().orc_tci_history(self.h, None, None, _ip(g))
        return g[:n].copy()

    def evaluate(self, x):
        x = np.ascontiguousarray(x, np.int32)
        return lib().orc_tci_evaluate(self.h, _i32(x))

    def workstats(self):
        out = np.zeros(6)
        lib().orc_tci_workstats(self.h, _dp(out))
        return dict(pi_elements=out[0], rrlu_argmax_bytes=out[1], rrlu_fused_bytes=out[2],
                    schur_flops=out[3], n_rrlu=int(out[4]), n_eval=int(out[5]))


def crossinterpolate2(f: Functor, localdims, initialpivots=None, **kw):
    t = TCI(localdims)
    if initialpivots is None:
        initialpivots = [[1] * len(localdims)]
    t.init(f, initialpivots)
    ranks, errors = t.optimize(f, **kw)
    return t, ranks, errors


def convergencecriterion(ranks, errors, nglobalpivots, tol, maxbonddim, ncheckhistory, checkconvglobalpivot=True):
    r = np.ascontiguousarray(ranks, np.int64)
    e = np.ascontiguousarray(errors, np.float64)
    g = np.ascontiguousarray(nglobalpivots, np.int64)
    return bool(lib().orc_convergencecriterion(len(r), _ip(r), _dp(e), _ip(g), float(tol), int(maxbonddim),
                                               int(ncheckhistory), int(checkconvglobalpivot)))
