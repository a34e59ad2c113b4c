"""ctypes wrapper of the TEST-ONLY CPU build of the device solve engine."""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_L = None


def lib():
    global _L
    if _L is None:
        L = C.CDLL(os.path.join(_HERE, "libengine_host.so"))
        vp = C.c_void_p
        L.eh_solve.restype = vp
        L.eh_solve.argtypes = [vp, C.POINTER(C.c_int)]
        L.eh_solve_mt.restype = vp
        L.eh_solve_mt.argtypes = [vp, C.POINTER(C.c_int)]
        L.eh_last_error.restype = C.c_char_p
        for nm in ("eh_free",):
            getattr(L, nm).argtypes = [vp]
        L.eh_niter.argtypes = [vp]
        L.eh_errsite.argtypes = [vp]
        L.eh_history.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.eh_setn.argtypes = [vp, C.c_int, C.c_int]
        L.eh_set.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int32)]
        L.eh_Tdims.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
        L.eh_T.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
        L.eh_maxsample.restype = C.c_double
        L.eh_maxsample.argtypes = [vp]
        L.eh_npivoterrors.argtypes = [vp]
        L.eh_pivoterrors.argtypes = [vp, C.POINTER(C.c_double)]
        L.eh_bonderrors.argtypes = [vp, C.POINTER(C.c_double)]
        L.eh_workstats.argtypes = [vp, C.POINTER(C.c_double)]
        L.eh_rrlu_mt.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int64, C.c_double,
                                 C.c_double, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_double)]
        L.eh_rrlu.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int64, C.c_double,
                              C.c_double, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                              C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _L = L
    return _L


class EngineSolve:
    def __init__(self, holder, threads=1):
        st = C.c_int()
        self.L = holder.L
        fn = lib().eh_solve_mt if threads > 1 else lib().eh_solve
        self.h = fn(C.byref(holder.st), C.byref(st))
        self.status = st.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().eh_free(self.h)

    @property
    def niter(self):
        return lib().eh_niter(self.h)

    def history(self):
        n = self.niter
        r = np.zeros(max(n, 1), np.int64); e = np.zeros(max(n, 1)); g = np.zeros(max(n, 1), np.int64)
        lib().eh_history(self.h, r.ctypes.data_as(C.POINTER(C.c_int64)), e.ctypes.data_as(C.POINTER(C.c_double)),
                         g.ctypes.data_as(C.POINTER(C.c_int64)))
        return r[:n], e[:n], g[:n]

    def set(self, which, b):
        n = lib().eh_setn(self.h, which, b)
        ln = b if which % 2 == 0 else self.L - 1 - b
        out = np.zeros(max(n * ln, 1), np.int32)
        lib().eh_set(self.h, which, b, out.ctypes.data_as(C.POINTER(C.c_int32)))
        return [list(map(int, out[i * ln:(i + 1) * ln])) for i in range(n)]

    def sitetensor(self, b):
        d = np.zeros(3, np.int64)
        lib().eh_Tdims(self.h, b, d.ctypes.data_as(C.POINTER(C.c_int64)))
        out = np.zeros(max(int(np.prod(d)), 1))
        lib().eh_T(self.h, b, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out[:int(np.prod(d))].reshape(tuple(d), order="F")

    @property
    def maxsample(self):
        return lib().eh_maxsample(self.h)

    def pivoterrors(self):
        n = lib().eh_npivoterrors(self.h)
        out = np.zeros(max(n, 1))
        lib().eh_pivoterrors(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out[:n]

    def bonderrors(self):
        out = np.zeros(self.L - 1)
        lib().eh_bonderrors(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out


def rrlu(mode, A, maxrank=(1 << 63) - 1, reltol=1e-14, abstol=0.0, leftorthogonal=True):
    """mode 0: register tile factor, 1: lean tile factor, 2: streaming rrlu_team."""
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = A.shape
    rp = np.zeros(max(m, 1), np.int32)
    cp = np.zeros(max(n, 1), np.int32)
    pv = np.zeros(max(m, n, 1))
    err = C.c_double()
    p = lib().eh_rrlu(mode, A.ctypes.data_as(C.POINTER(C.c_double)), m, n, int(maxrank), reltol, abstol,
                      int(leftorthogonal), rp.ctypes.data_as(C.POINTER(C.c_int32)),
                      cp.ctypes.data_as(C.POINTER(C.c_int32)), pv.ctypes.data_as(C.POINTER(C.c_double)),
                      C.byref(err))
    return dict(p=p, rowperm=rp[:m] + 1, colperm=cp[:n] + 1, pivots=pv[:max(p, 0)], error=err.value, A=A)


def rrlu_mt(mode, A, maxrank=(1 << 63) - 1, reltol=1e-14, abstol=0.0, leftorthogonal=True):
    """Register-tile rrLU on 256 host threads shaped like the device team."""
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = A.shape
    rp = np.zeros(max(m, 1), np.int32)
    cp = np.zeros(max(n, 1), np.int32)
    err = C.c_double()
    p = lib().eh_rrlu_mt(mode, A.ctypes.data_as(C.POINTER(C.c_double)), m, n, int(maxrank), reltol, abstol,
                         int(leftorthogonal), rp.ctypes.data_as(C.POINTER(C.c_int32)),
                         cp.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(err))
    return dict(p=p, rowperm=rp[:m] + 1, colperm=cp[:n] + 1, error=err.value)
