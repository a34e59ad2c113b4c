"""tci_amd — MI355X (gfx950) TCI2 sweep behind the reference's API.

Python mirror of TensorCrossInterpolation.jl's hot-path interface
(`crossinterpolate2`, `TensorCI2`, `optimize!` keywords, `rrlu`, `rrLU`
accessors, batch evaluation), calling libtci_hip.so through its C ABI
(include/tci.h).  All arithmetic runs in the gfx950 kernels; this module
only marshals arguments and results.  There is no CPU fallback: without the
built library or a gfx950 device every call raises.

Load it with importlib (the directory name is not a Python identifier), as
tests/conftest.py, bench.py and __graft_entry__.py do, under the name
`tci_amd`.
"""
from __future__ import annotations

import ctypes as C
import warnings
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from ._abi import (F_GAUSS_SUM, F_LORENTZ_INT, F_LORENTZ_WEIGHTED, F_QUANTICS_EXP,
                   F_QUANTICS_GAUSS_COS, F_SUM, F_TABLE, F_TT, TCI_UNLIMITED, ProblemHolder)

__all__ = [
    "TCIError", "ArgumentError", "Context", "default_context", "rrLU", "rrlu", "rrlu_batched", "MatrixLUCI",
    "left", "right", "rowindices", "colindices", "npivots", "pivoterrors", "lastpivoterror",
    "DeviceFunctor", "Lorentz", "WeightedLorentz", "QuanticsGaussCos", "GaussSum", "SumFunctor",
    "TableFunctor", "TTFunctor", "QuanticsExp", "batchevaluate", "TensorCI2", "crossinterpolate2",
    "crossinterpolate2_batch", "TCI_UNLIMITED", "capacity_plan", "tt_evaluate", "sample_points",
    "tensorci2_from_sets", "optimize", "CachedFunction", "TensorTrain", "compress", "tt_sweep",
    "tensorci2_from_tt", "floatingzone", "estimatetrueerror",
]


# --------------------------------------------------------------------------
# errors (tci_status → the reference's exception)
# --------------------------------------------------------------------------
class TCIError(RuntimeError):
    """ErrorException of the reference, or a device error; .status = tci_status."""

    def __init__(self, status, msg):
        super().__init__(f"{_abi.STATUS.get(status, status)}: {msg}")
        self.status = status


class ArgumentError(TCIError, ValueError):
    """ArgumentError of the reference (tensorci2.jl:736-754)."""


_MESSAGES = {
    1: "lu.L contains NaNs", 2: "lu.U contains NaNs", 4: "maxsamplevalue is zero!",
    5: "NaN in tensor", 6: "Pivot matrix is not square!", 9: "singular pivot matrix",
    10: "capacity exceeded (raise chi_cap)",
}


def _raise(status, msg=None):
    if msg is None:
        msg = _abi.lib().tci_last_error().decode() or _MESSAGES.get(status, "")
    if status == 3:
        raise ArgumentError(status, msg)
    raise TCIError(status, msg)


def _check(status):
    if status != 0:
        _raise(status)


# --------------------------------------------------------------------------
# context
# --------------------------------------------------------------------------
class Context:
    """One device + one HIP stream + an HBM budget (tci_ctx)."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        self._h = C.c_void_p()
        _check(_abi.lib().tci_ctx_create(C.byref(self._h), int(device), int(hbm_budget_bytes)))

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        _check(_abi.lib().tci_ctx_synchronize(self._h))

    _TUNE_FIELDS = ("side", "tile", "grid_team", "grid_max_jobs", "ms_stream", "ms_grid", "ms_tile_batch",
                    "ms_stream_batch")

    def autotune_rrlu(self):
        """Measure the standalone rrLU's kernel choices per size bucket on
        this device (tci_ctx_autotune_rrlu); AUTO follows them from now on.
        Returns the table as a list of dicts."""
        tab = (_abi.tci_rrlu_tune * 32)()
        n = C.c_int32()
        _check(_abi.lib().tci_ctx_autotune_rrlu(self._h, tab, 32, C.byref(n)))
        return [{k: getattr(tab[q], k) for k in self._TUNE_FIELDS} for q in range(n.value)]

    def set_rrlu_tuning(self, table):
        """Install a table returned by autotune_rrlu (an empty one returns
        AUTO to its fixed rules)."""
        tab = (_abi.tci_rrlu_tune * max(len(table), 1))()
        for q, row in enumerate(table):
            for k in self._TUNE_FIELDS:
                setattr(tab[q], k, row.get(k, 0))
        _check(_abi.lib().tci_ctx_set_rrlu_tuning(self._h, tab, len(table)))

    def autotune_solve(self, problem, batch=256):
        """Time `batch` copies of problem = (f, localdims, initialpivots,
        kwargs) with 64, 256, 512 and 1024 threads per solve and keep the
        fastest for batches of at least `batch` solves of at least this
        problem's capacity (tci_ctx_autotune_solve).  Returns (team, [ms at
        64, 256, 512, 1024])."""
        h = _holder(*problem)
        ms = (C.c_float * 4)()
        team = C.c_int32()
        _check(_abi.lib().tci_ctx_autotune_solve(self._h, C.byref(h.st), int(batch), ms, C.byref(team)))
        return team.value, list(ms)

    def set_solve_tuning(self, entries):
        """entries: [(min_mrcap, min_count, team)] (tci_ctx_set_solve_tuning);
        [] gives every batch the default team."""
        n = len(entries)
        keys = (C.c_int32 * max(n, 1))(*[int(e[0]) for e in entries])
        counts = (C.c_int32 * max(n, 1))(*[int(e[1]) for e in entries])
        teams = (C.c_int32 * max(n, 1))(*[int(e[2]) for e in entries])
        _check(_abi.lib().tci_ctx_set_solve_tuning(self._h, keys, counts, teams, n))

    def close(self):
        if self._h:
            _abi.lib().tci_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_DEFAULT = None


def default_context():
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Context(0)
    return _DEFAULT


# --------------------------------------------------------------------------
# rrLU (matrixlu.jl)
# --------------------------------------------------------------------------
@dataclass
class rrLU:
    """Fields of rrLU{Float64} (matrixlu.jl:71-91)."""
    rowpermutation: np.ndarray
    colpermutation: np.ndarray
    L: np.ndarray
    U: np.ndarray
    leftorthogonal: bool
    npivot: int
    error: float

    @property
    def shape(self):
        return (self.L.shape[0], self.U.shape[1])


def rrlu(A, maxrank=TCI_UNLIMITED, reltol=1e-14, abstol=0.0, leftorthogonal=True, ctx=None) -> rrLU:
    """rrlu(A; maxrank, reltol, abstol, leftorthogonal) (matrixlu.jl:217-225) on the GPU."""
    ctx = ctx or default_context()
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    cap = max(min(m, n, int(maxrank)), 0)
    rp = np.zeros(max(m, 1), np.int64)
    cp = np.zeros(max(n, 1), np.int64)
    Lb = np.zeros((max(m, 1), max(cap, 1)), order="F")
    Ub = np.zeros((max(cap, 1), max(n, 1)), order="F")
    p = C.c_int64()
    err = C.c_double()
    st = _abi.lib().tci_rrlu_f64(ctx.handle, _abi.dptr(A) if A.size else None, max(m, 1), m, n,
                                 int(maxrank), float(reltol), float(abstol), int(bool(leftorthogonal)),
                                 _abi.i64ptr(rp), _abi.i64ptr(cp), _abi.dptr(Lb), max(m, 1),
                                 _abi.dptr(Ub), max(cap, 1), C.byref(p), C.byref(err))
    _check(st)
    pv = p.value
    return rrLU(rp[:m].copy(), cp[:n].copy(), np.array(Lb[:m, :pv]), np.array(Ub[:pv, :n]),
                bool(leftorthogonal), pv, err.value)


class MatrixLUCI:
    """MatrixLUCI(A; maxrank, reltol, abstol, leftorthogonal) (matrixluci.jl:1-7)
    with left(luci)/right(luci) (matrixluci.jl:40-84) computed on the device
    (tci_luci_f64): with leftorthogonal, left = colstimespivotinv and
    right = rowmatrix; otherwise left = colmatrix and right =
    pivotinvtimesrows."""

    def __init__(self, A, maxrank=TCI_UNLIMITED, reltol=1e-14, abstol=0.0, leftorthogonal=True, ctx=None):
        ctx = ctx or default_context()
        A = np.asfortranarray(A, dtype=np.float64)
        m, n = A.shape
        cap = max(min(m, n, int(maxrank)), 0)
        rp = np.zeros(max(m, 1), np.int64)
        cp = np.zeros(max(n, 1), np.int64)
        lb = np.zeros((max(m, 1), max(cap, 1)), order="F")
        rb = np.zeros((max(cap, 1), max(n, 1)), order="F")
        p = C.c_int64()
        err = C.c_double()
        _check(_abi.lib().tci_luci_f64(ctx.handle, _abi.dptr(A) if A.size else None, max(m, 1), m, n,
                                       int(maxrank), float(reltol), float(abstol), int(bool(leftorthogonal)),
                                       _abi.i64ptr(rp), _abi.i64ptr(cp), _abi.dptr(lb), max(m, 1),
                                       _abi.dptr(rb), max(cap, 1), C.byref(p), C.byref(err)))
        pv = p.value
        self.shape = (m, n)
        self.rowpermutation, self.colpermutation = rp[:m].copy(), cp[:n].copy()
        self.leftorthogonal = bool(leftorthogonal)
        self.npivot, self.error = pv, err.value
        self._left, self._right = np.array(lb[:m, :pv]), np.array(rb[:pv, :n])

    def left(self):
        return self._left

    def right(self):
        return self._right

    def npivots(self):
        return self.npivot

    def rowindices(self):
        return self.rowpermutation[:self.npivot]

    def colindices(self):
        return self.colpermutation[:self.npivot]

    def lastpivoterror(self):
        return self.error


def left(lu, permute=True):
    """left(lu; permute) (matrixlu.jl:374-382), or left(luci) (matrixluci.jl:70-76)."""
    if isinstance(lu, MatrixLUCI):
        return lu.left()
    if not permute:
        return lu.L
    out = np.empty_like(lu.L)
    out[lu.rowpermutation - 1, :] = lu.L
    return out


def right(lu, permute=True):
    """right(lu; permute) (matrixlu.jl:384-392), or right(luci) (matrixluci.jl:78-84)."""
    if isinstance(lu, MatrixLUCI):
        return lu.right()
    if not permute:
        return lu.U
    out = np.empty_like(lu.U)
    out[:, lu.colpermutation - 1] = lu.U
    return out


def rowindices(lu: rrLU):
    return lu.rowpermutation[:lu.npivot]


def colindices(lu: rrLU):
    return lu.colpermutation[:lu.npivot]


def npivots(lu: rrLU):
    return lu.npivot


def pivoterrors(lu: rrLU):
    """[|diag|..., error] (matrixlu.jl:394-416)."""
    p = lu.npivot
    d = np.diag(lu.U[:p, :p]) if lu.leftorthogonal else np.diag(lu.L[:p, :p])
    return np.concatenate([np.abs(d), [lu.error]])


def lastpivoterror(lu: rrLU):
    return lu.error


def rrlu_batched(mats, maxrank=TCI_UNLIMITED, reltol=1e-14, abstol=0.0, leftorthogonal=True,
                 ctx=None, device_buffers=None):
    """Factor a list of matrices in one launch (tci_rrlu_batched_f64).

    Returns (rowperms, colperms, npivots, errors, statuses) per matrix and
    leaves the packed factors on the device.  `device_buffers` may pass an
    already-resident torch tensor holding the concatenated column-major
    matrices (as returned by `upload_matrices`) to skip the upload.
    """
    import torch
    ctx = ctx or default_context()
    if device_buffers is None:
        device_buffers = upload_matrices(mats)
    A, offs, shapes = device_buffers
    cnt = len(shapes)
    tot_perm = sum(m + n for m, n in shapes)
    perm = torch.zeros(max(tot_perm, 1), dtype=torch.int32, device=A.device)
    jobs = (_abi.tci_lu_job * cnt)()
    po = 0
    for i, (m, n) in enumerate(shapes):
        j = jobs[i]
        j.A = A.data_ptr() + 8 * offs[i]
        j.rowperm = perm.data_ptr() + 4 * po
        j.colperm = perm.data_ptr() + 4 * (po + m)
        j.m, j.n, j.maxrank = m, n, int(maxrank)
        j.reltol, j.abstol, j.leftorth = float(reltol), float(abstol), int(bool(leftorthogonal))
        po += m + n
    jb = torch.frombuffer(bytearray(jobs), dtype=torch.uint8).to(A.device)
    torch.cuda.synchronize(A.device)
    _check(_abi.lib().tci_rrlu_batched_f64(ctx.handle, C.c_void_p(jb.data_ptr()), cnt))
    ctx.synchronize()
    C.memmove(jobs, bytes(jb.cpu().numpy().tobytes()), C.sizeof(jobs))
    permh = perm.cpu().numpy()
    out = []
    po = 0
    for i, (m, n) in enumerate(shapes):
        j = jobs[i]
        out.append(dict(rowpermutation=permh[po:po + m].astype(np.int64),
                        colpermutation=permh[po + m:po + m + n].astype(np.int64),
                        npivot=int(j.npivot), error=float(j.error), status=int(j.status)))
        po += m + n
    return out


def upload_matrices(mats):
    """Concatenate column-major FP64 matrices into one device tensor."""
    import torch
    shapes = [tuple(int(x) for x in np.shape(a)) for a in mats]
    offs, o = [], 0
    for m, n in shapes:
        offs.append(o)
        o += m * n
    host = np.empty(max(o, 1))
    for a, off, (m, n) in zip(mats, offs, shapes):
        host[off:off + m * n] = np.asarray(a, np.float64).ravel(order="F")
    return torch.from_numpy(host).to("cuda"), offs, shapes


# --------------------------------------------------------------------------
# device functors (include/tci_functors.h)
# --------------------------------------------------------------------------
@dataclass
class DeviceFunctor:
    """A function f(σ) evaluated on the device: kind tag + parameter block."""
    kind: int
    params: np.ndarray = field(default_factory=lambda: np.zeros(0))
    iparams: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))

    def struct(self):
        self._p = np.ascontiguousarray(self.params, np.float64)
        self._ip = np.ascontiguousarray(self.iparams, np.int32)
        return _abi.tci_functor(self.kind, len(self._p), _abi.dptr(self._p) if len(self._p) else None,
                                len(self._ip), _abi.i32ptr(self._ip) if len(self._ip) else None)


def Lorentz(a=1.0):
    """f(v) = a / (1 + Σ v_i²) (README.md:24-28)."""
    return DeviceFunctor(F_LORENTZ_INT, np.array([a], np.float64))


def WeightedLorentz(c, D=16.0):
    """f(v) = 1 / (1 + Σ c_i (v_i/D)²)."""
    return DeviceFunctor(F_LORENTZ_WEIGHTED, np.concatenate([[float(D)], np.asarray(c, np.float64)]))


def QuanticsGaussCos(a, c, w, A=0.1, F=2.0 ** 20):
    """x = Σ(σ_i−1)2^{−i}; f = Σ a_k exp(−(x−c_k)²/(2w_k²)) + A cos(F x)."""
    K = len(a)
    trip = np.stack([np.asarray(a, float), np.asarray(c, float), np.asarray(w, float)], 1).ravel()
    return DeviceFunctor(F_QUANTICS_GAUSS_COS, np.concatenate([[K], trip, [A, F]]))


def GaussSum(a, w, c, d):
    """x_i = (σ_i−1)/(d−1); f = Σ_k a_k exp(−Σ_i (x_i−c_{k,i})²/(2w_k²)); c is K×L."""
    a = np.asarray(a, float)
    K = len(a)
    return DeviceFunctor(F_GAUSS_SUM, np.concatenate([[K, float(d)], a, np.asarray(w, float),
                                                       np.asarray(c, float).ravel()]))


def SumFunctor(s=1.0):
    """f(v) = s·Σ v_i (test_batcheval.jl:13-35)."""
    return DeviceFunctor(F_SUM, np.array([s], np.float64))


def TableFunctor(table):
    """f(v) = table[v_1, …, v_L] for a small dense table (column-major)."""
    t = np.asarray(table, np.float64)
    return DeviceFunctor(F_TABLE, t.ravel(order="F"), np.array(t.shape, np.int32))


def TTFunctor(cores):
    """f(v) = T_1[1,v_1,:]·…·T_L[:,v_L,1] for given cores (χ_{i-1}, d_i, χ_i)."""
    L = len(cores)
    chis = [cores[0].shape[0]] + [c.shape[2] for c in cores]
    ds = [c.shape[1] for c in cores]
    params = np.concatenate([np.asarray(c, np.float64).ravel(order="F") for c in cores])
    return DeviceFunctor(F_TT, params, np.array(chis + ds, np.int32))


def QuanticsExp(alpha1, beta1, alpha2=0.0, beta2=0.0):
    """x = Σ(σ_i−1)2^{−i}; f = α₁ exp(β₁x) + α₂ exp(β₂x)."""
    return DeviceFunctor(F_QUANTICS_EXP, np.array([alpha1, beta1, alpha2, beta2], np.float64))


class CachedFunction:
    """CachedFunction{Float64,K}(f, localdims) (cachedfunction.jl:8-60) for a
    device functor: passed to crossinterpolate2 in place of f, every
    evaluation of the solve goes through a memo table on the device
    (2^log2size entries), keyed by Σ coeff_i·(σ_i - 1) (177-184) in 256 bits.
    keybits is K's width; a domain whose largest key does not fit below
    typemax(K) is refused as the reference refuses it (19-21).  The reference's
    table lives as long as the wrapper; here each solve starts an empty table,
    and after a solve `cachedata()` holds that solve's table."""

    def __init__(self, f: DeviceFunctor, localdims, keybits=128, log2size=20):
        self.f = f
        self.localdims = [int(d) for d in localdims]
        self.keybits = int(keybits)
        self.log2size = int(log2size)
        if self.keybits > 256:
            raise ArgumentError(3, "keys are at most 256 bits on the device")
        self.coeffs = [1]
        for d in self.localdims[:-1]:
            self.coeffs.append(self.coeffs[-1] * d)
        if sum(c * (d - 1) for c, d in zip(self.coeffs, self.localdims)) >= (1 << self.keybits) - 1:
            raise TCIError(8, "Overflow in CachedFunction. Use ValueType = a bigger type with fixed size, "
                              "e.g., BitIntegers.UInt256")
        self._data = {}
        self.stats = None

    def encodecachekey(self, index):
        """_key / encodecachekey (cachedfunction.jl:177-186)."""
        if len(index) != len(self.coeffs):
            raise TCIError(8, "Invalid length of indexset")
        return sum(c * (int(x) - 1) for c, x in zip(self.coeffs, index))

    def decodecachekey(self, key):
        """decodecachekey (cachedfunction.jl:188-196)."""
        out = []
        for d in self.localdims:
            key, r = divmod(key, d)
            out.append(r + 1)
        return out

    def cachedata(self):
        """cachedata(cf) (cachedfunction.jl:198-203) of the last solve's table."""
        return dict(self._data)

    def iscached(self, index):
        return tuple(int(x) for x in index) in self._data


def batchevaluate(f: DeviceFunctor, localdims, leftindexset, rightindexset, M, ctx=None):
    """_batchevaluate_dispatch(Float64, f, localdims, Iset, Jset, Val(M)) (batcheval.jl:32-61).

    Returns (array of shape (|Iset|, d…, |Jset|), max|array|)."""
    ctx = ctx or default_context()
    localdims = np.ascontiguousarray(localdims, np.int32)
    L = len(localdims)
    nI, nJ = len(leftindexset), len(rightindexset)
    nl = len(leftindexset[0]) if nI else 0
    nr = len(rightindexset[0]) if nJ else 0
    if nI and nJ and nl + M + nr != L:
        raise ArgumentError(3, "Invalid number of central indices")
    shape = (nI,) + tuple(int(d) for d in localdims[nl:nl + M]) + (nJ,)
    if nI * nJ == 0:
        return np.zeros(tuple(0 for _ in shape)), 0.0
    left = np.ascontiguousarray(np.asarray(leftindexset, np.uint8).reshape(nI, nl))
    right = np.ascontiguousarray(np.asarray(rightindexset, np.uint8).reshape(nJ, nr))
    out = np.zeros(int(np.prod(shape)))
    mx = C.c_double(0.0)
    fs = f.struct()
    _check(_abi.lib().tci_pi_eval_f64(ctx.handle, C.byref(fs), L, _abi.i32ptr(localdims),
                                      _abi.u8ptr(left), nI, nl, _abi.u8ptr(right), nJ, nr,
                                      _abi.dptr(out), C.byref(mx)))
    return out.reshape(shape, order="F"), mx.value


# --------------------------------------------------------------------------
# tensor-train evaluation (abstracttensortrain.jl:124-132)
# --------------------------------------------------------------------------
def _points(points, L):
    x = np.asarray(points)
    if x.ndim == 1:
        x = x.reshape(1, -1)
    if x.shape[1] != L:
        raise ArgumentError(3, f"To evaluate a tt of length {L}, you have to provide {L} indices, "
                               f"but there were {x.shape[1]}.")
    return np.ascontiguousarray(x, np.uint8) if x.size else np.zeros((0, L), np.uint8)


def tt_evaluate(sitetensors, points, ctx=None):
    """evaluate(tt, indexset) (abstracttensortrain.jl:124-132) at every row of
    `points` (1-based, shape npts×L), on the device.  Returns npts values."""
    ctx = ctx or default_context()
    L = len(sitetensors)
    x = _points(points, L)
    links = np.array([sitetensors[0].shape[0]] + [t.shape[2] for t in sitetensors], np.int64)
    ld = np.array([t.shape[1] for t in sitetensors], np.int32)
    cores = np.ascontiguousarray(np.concatenate([np.asarray(t, np.float64).ravel(order="F")
                                                 for t in sitetensors]))
    out = np.zeros(max(x.shape[0], 1))
    _check(_abi.lib().tci_tt_evaluate_f64(ctx.handle, L, _abi.i64ptr(links), _abi.i32ptr(ld),
                                          _abi.dptr(cores), x.shape[0], _abi.u8ptr(x), _abi.dptr(out)))
    return out[:x.shape[0]]


class TensorTrain:
    """TensorTrain{Float64,3} (tensortrain.jl:1-64): site tensors of shape
    (χ_b, d_b, χ_{b+1}); evaluate/compress run on the device."""

    def __init__(self, sitetensors):
        ts = [np.asarray(t, np.float64) for t in (sitetensors.sitetensors if hasattr(sitetensors, "sitetensors")
                                                  else sitetensors)]
        for b, t in enumerate(ts):
            if t.ndim != 3:
                raise ArgumentError(3, f"site tensor {b + 1} must have 3 dimensions")
            if b + 1 < len(ts) and t.shape[2] != ts[b + 1].shape[0]:
                raise ArgumentError(3, f"The tensors at {b + 1} and {b + 2} must share a common link index")
        self.sitetensors = [np.array(t, order="F") for t in ts]

    def __len__(self):
        return len(self.sitetensors)

    def linkdims(self):
        """[size(T, 3) for T in sitetensors[1:end-1]] (abstracttensortrain.jl:29-31)."""
        return [t.shape[2] for t in self.sitetensors[:-1]]

    def rank(self):
        return max(self.linkdims()) if len(self) > 1 else 0

    def sitedims(self):
        return [[t.shape[1]] for t in self.sitetensors]

    def evaluate(self, points, ctx=None):
        return tt_evaluate(self.sitetensors, points, ctx)

    def __call__(self, indexset):
        return float(self.evaluate(np.asarray(indexset).reshape(1, -1))[0])

    def _packed(self):
        links = np.array([self.sitetensors[0].shape[0]] + [t.shape[2] for t in self.sitetensors], np.int64)
        ld = np.array([t.shape[1] for t in self.sitetensors], np.int32)
        flat = np.ascontiguousarray(np.concatenate([t.ravel(order="F") for t in self.sitetensors]))
        return links, ld, flat

    def _unpack(self, links, ld, flat):
        out, off = [], 0
        for b in range(len(ld)):
            sz = int(links[b] * ld[b] * links[b + 1])
            out.append(np.array(flat[off:off + sz].reshape((int(links[b]), int(ld[b]), int(links[b + 1])),
                                                           order="F"), order="F"))
            off += sz
        self.sitetensors = out


_FACTOR = {"LU": _abi.FACTOR_LU, "CI": _abi.FACTOR_CI, "SVD": _abi.FACTOR_SVD}


def _method(method):
    key = method.lstrip(":").upper() if isinstance(method, str) else method
    if key not in _FACTOR:
        raise TCIError(8, "Not implemented yet.")
    return _FACTOR[key]


def compress(tt, method="LU", tolerance=1e-12, maxbonddim=TCI_UNLIMITED, normalizeerror=True, ctx=None):
    """compress!(tt, method; tolerance, maxbonddim, normalizeerror)
    (tensortrain.jl:147-183) on the device, in place: a left-to-right sweep
    without truncation, then a right-to-left one truncating each bond with
    rrLU (:LU) or MatrixLUCI (:CI).  :SVD is refused (ArgumentError)."""
    code = _method(method)
    ctx = ctx or default_context()
    links, ld, flat = tt._packed()
    _check(_abi.lib().tci_tt_compress_f64(ctx.handle, len(tt), _abi.i64ptr(links), _abi.i32ptr(ld),
                                          _abi.dptr(flat), code, float(tolerance), int(maxbonddim),
                                          int(bool(normalizeerror))))
    tt._unpack(links, ld, flat)


def _kron_left_at(Iset, d, q):
    """Entry q (0-based) of kronecker(Iset, localdim) (tensorci2.jl:315-320),
    whose Iset index runs fastest, without building the product."""
    s, i = divmod(q, len(Iset))
    return list(Iset[i]) + [s + 1]


def _kron_right_at(d, Jset, q):
    """Entry q (0-based) of kronecker(localdim, Jset) (tensorci2.jl:322-327),
    whose local index runs fastest, without building the product."""
    j, s = divmod(q, d)
    return [s + 1] + list(Jset[j])


def tt_sweep(tt, method="CI", forward=True, maxrank=TCI_UNLIMITED, reltol=1e-14, abstol=0.0, ctx=None):
    """One sweep of factorizations over tt on the device (tci_tt_sweep_f64),
    in place: per bond, the rrLU of the grouped site (leftorthogonal =
    forward), the site replaced by left (forward) or right (backward) of
    rrLU (:LU) or MatrixLUCI (:CI) and the other factor multiplied into the
    neighbour (tensortrain.jl:155-182, conversion.jl:105-141).  Returns the
    per-bond rowindices, colindices and the running max of pivoterrors."""
    code = _method(method)
    ctx = ctx or default_context()
    L = len(tt)
    links, ld, flat = tt._packed()
    R = int(max(links[1:L])) if L > 1 else 0
    piv = np.zeros(max(2 * int(links[1:L].sum()), 1), np.int32)
    pe = np.zeros(R + 1)
    _check(_abi.lib().tci_tt_sweep_f64(ctx.handle, L, _abi.i64ptr(links), _abi.i32ptr(ld), _abi.dptr(flat),
                                       code, int(bool(forward)), int(maxrank), float(reltol),
                                       float(abstol), _abi.i32ptr(piv), _abi.dptr(pe)))
    tt._unpack(links, ld, flat)
    rows, cols, q = [], [], 0
    for b in range(L - 1):
        p = int(links[b + 1])
        rows.append([int(v) for v in piv[q:q + p]])
        cols.append([int(v) for v in piv[q + p:q + 2 * p]])
        q += 2 * p
    return rows, cols, pe


def _sweep1sitegetindices(tt, forward, spectator=None, maxbonddim=TCI_UNLIMITED, tolerance=0.0, ctx=None):
    """sweep1sitegetindices!(tt, forwardsweep, spectatorindices; maxbonddim,
    tolerance) (conversion.jl:73-156): the factorizations and products on the
    device (MatrixLUCI with abstol = tolerance, leftorthogonal = forward),
    the index bookkeeping here.  Mutates tt and spectator."""
    L = len(tt)
    ld = [s[0] for s in tt.sitedims()]
    rows, cols, pe = tt_sweep(tt, "CI", forward, maxbonddim, 1e-14, tolerance, ctx)
    indexset = [[[]]]
    for i in range(L - 1):
        if forward:
            b = i   # site b, bond b
            indexset.append([_kron_left_at(indexset[-1], ld[b], r - 1) for r in rows[b]])
            if spectator:
                spectator[b] = [spectator[b][c - 1] for c in cols[b]]
        else:
            b = L - 2 - i   # site b+1, bond b
            indexset.append([_kron_right_at(ld[b + 1], indexset[-1], c - 1) for c in cols[b]])
            if spectator:
                spectator[b + 1] = [spectator[b + 1][r - 1] for r in rows[b]]
    return (indexset if forward else indexset[::-1]), pe


def tensorci2_from_tt(tt, tolerance=1e-12, maxbonddim=TCI_UNLIMITED, maxiter=3, ctx=None):
    """TensorCI2{Float64}(tt::TensorTrain; tolerance, maxbonddim, maxiter)
    (conversion.jl:158-176): pivot sets from alternating MatrixLUCI sweeps
    over tt (which is modified in place, as in the reference), the site
    tensors of the swept train, maxsamplevalue the largest |entry|."""
    if len(tt) < 2:
        raise TCIError(8, "localdims should have at least 2 elements!")
    Iset, _ = _sweep1sitegetindices(tt, True, maxbonddim=maxbonddim, tolerance=tolerance, ctx=ctx)
    Jset, pivoterrors = _sweep1sitegetindices(tt, False, maxbonddim=maxbonddim, tolerance=tolerance, ctx=ctx)
    for it in range(3, maxiter + 1):
        if it % 2 == 1:
            Isetnew, pivoterrors = _sweep1sitegetindices(tt, True, Jset, ctx=ctx)
            if Isetnew == Iset:
                break
        else:
            Jsetnew, pivoterrors = _sweep1sitegetindices(tt, False, Iset, ctx=ctx)
            if Jsetnew == Jset:
                break
    t = TensorCI2.__new__(TensorCI2)
    t.localdims = [s[0] for s in tt.sitedims()]
    L = len(t.localdims)
    t.Iset, t.Jset = Iset, Jset
    t.sitetensors = tt.sitetensors
    t.pivoterrors = np.asarray(pivoterrors)
    t.bonderrors = np.zeros(L - 1)
    t.maxsamplevalue = float(max(np.max(np.abs(s)) for s in t.sitetensors))
    t.Iset_history, t.Jset_history = [], []
    t.status, t.niter = 0, 0
    t.ranks, t.errors, t.nglobalpivots = np.zeros(0, np.int64), np.zeros(0), np.zeros(0, np.int64)
    t.workstats = {}
    return t


def floatingzone(tt, f, initialpoints, nsweeps=TCI_UNLIMITED, earlystoptol=float("inf"), ctx=None):
    """_floatingzone(ttcache, f; earlystoptol, nsweeps, initp)
    (globalsearch.jl:43-95) from every row of initialpoints, one workgroup
    per point (tci_tt_floatingzone_f64).  Returns (pivots npts×L, errors)."""
    if not isinstance(f, DeviceFunctor):
        raise TypeError("f must be a DeviceFunctor")
    if nsweeps <= 0:
        raise TCIError(8, "nsweeps should be positive!")
    ctx = ctx or default_context()
    tt = tt if isinstance(tt, TensorTrain) else TensorTrain(tt)
    L = len(tt)
    x = np.array(_points(initialpoints, L), order="C")
    links, ld, flat = tt._packed()
    errs = np.zeros(max(x.shape[0], 1))
    fs = f.struct()
    _check(_abi.lib().tci_tt_floatingzone_f64(ctx.handle, C.byref(fs), L, _abi.i64ptr(links),
                                              _abi.i32ptr(ld), _abi.dptr(flat), x.shape[0], _abi.u8ptr(x),
                                              _abi.dptr(errs), int(nsweeps), float(earlystoptol)))
    return x.astype(np.int64), errs[:x.shape[0]]


def estimatetrueerror(tt, f, nsearch=100, initialpoints=None, seed=0, ctx=None):
    """estimatetrueerror(tt, f; nsearch, initialpoints) (globalsearch.jl:18-41):
    the floating-zone search from each start point on the device, then the
    (pivot, error) pairs in descending error order (ties keep their order,
    NaN first), repeats dropped.  Random start points come from a seeded
    numpy generator (the reference draws from Julia's default RNG)."""
    if nsearch <= 0 and initialpoints is None:
        raise TCIError(8, "No search is performed")
    if nsearch < 0:
        raise TCIError(8, "nsearch must be non-negative")
    tt = tt if isinstance(tt, TensorTrain) else TensorTrain(tt)
    if initialpoints is None:
        initialpoints = sample_points([s[0] for s in tt.sitedims()], nsearch, seed)
    piv, errs = floatingzone(tt, f, initialpoints, ctx=ctx)
    order = sorted(range(len(errs)), key=lambda q: (0 if errs[q] != errs[q] else 1, -errs[q] if errs[q] == errs[q] else 0))
    out, seen = [], set()
    for q in order:
        key = (tuple(int(v) for v in piv[q]), "nan" if errs[q] != errs[q] else float(errs[q]))
        if key not in seen:
            seen.add(key)
            out.append(([int(v) for v in piv[q]], float(errs[q])))
    return out


def sample_points(localdims, n, seed=0):
    """n seeded random points of the domain (1-based, n×L bytes)."""
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(1, int(d) + 1, n) for d in localdims], 1).astype(np.uint8)


# --------------------------------------------------------------------------
# TensorCI2 / crossinterpolate2 (tensorci2.jl)
# --------------------------------------------------------------------------
class TensorCI2:
    """Result of a device solve with the fields of TensorCI2{Float64} (tensorci2.jl:6-40).

    Iset[b]/Jset[b] are lists of 1-based multi-indices (b = 0..L-1 for sites
    1..L); sitetensors[b] has shape (χ_{b}, d_b, χ_{b+1})."""

    def __init__(self, batch, i, localdims):
        lib = _abi.lib()
        self.localdims = [int(d) for d in localdims]
        L = len(self.localdims)
        r = _abi.tci_result()
        _check(lib.tci_batch_result(batch, i, C.byref(r)))
        self.status = r.status
        self.niter = r.niter
        self.maxsamplevalue = r.maxsamplevalue
        self.workstats = dict(zip(["pi_elements", "rrlu_argmax_bytes", "rrlu_fused_bytes",
                                   "schur_flops", "n_rrlu", "n_eval"], list(r.workstats)))
        n = max(r.niter, 1)
        ranks = np.zeros(n, np.int64)
        errs = np.zeros(n)
        ng = np.zeros(n, np.int64)
        lib.tci_batch_history(batch, i, _abi.i64ptr(ranks), _abi.dptr(errs), _abi.i64ptr(ng))
        self.ranks, self.errors, self.nglobalpivots = ranks[:r.niter], errs[:r.niter], ng[:r.niter]
        pe = np.zeros(max(r.npivoterrors, 1))
        lib.tci_batch_pivoterrors(batch, i, _abi.dptr(pe), len(pe))
        self.pivoterrors = pe[:r.npivoterrors]
        be = np.zeros(max(L - 1, 1))
        lib.tci_batch_bonderrors(batch, i, _abi.dptr(be))
        self.bonderrors = be[:L - 1]
        self.Iset, self.Jset = [], []
        hI, hJ = [], []
        for which, dst in ((0, self.Iset), (1, self.Jset), (2, hI), (3, hJ)):
            for site in range(1, L + 1):
                ln = site - 1 if which % 2 == 0 else L - site
                cnt = C.c_int32()
                buf = np.zeros(1 << 16, np.uint8)
                lib.tci_batch_indexset(batch, i, which, site, C.byref(cnt), _abi.u8ptr(buf), len(buf))
                if cnt.value * ln > len(buf):
                    buf = np.zeros(cnt.value * ln, np.uint8)
                    lib.tci_batch_indexset(batch, i, which, site, C.byref(cnt), _abi.u8ptr(buf), len(buf))
                dst.append([list(map(int, buf[e * ln:(e + 1) * ln])) for e in range(cnt.value)])
        # Iset_history / Jset_history (tensorci2.jl:12-13): only the last
        # entry is ever read (sweep2site!, 874-877), so only it is kept
        has_hist = any(len(s) for s in hI)
        self.Iset_history = [hI] if has_hist else []
        self.Jset_history = [hJ] if has_hist else []
        self.sitetensors = []
        if self.status == 0:
            for site in range(1, L + 1):
                d3 = np.zeros(3, np.int64)
                lib.tci_batch_sitetensor(batch, i, site, _abi.i64ptr(d3), None, 0)
                out = np.zeros(max(int(np.prod(d3)), 1))
                _check(lib.tci_batch_sitetensor(batch, i, site, _abi.i64ptr(d3), _abi.dptr(out), len(out)))
                self.sitetensors.append(out[:int(np.prod(d3))].reshape(tuple(d3), order="F"))

    def __len__(self):
        return len(self.localdims)

    def linkdims(self):
        """[|Iset[b+1]|] (tensorci2.jl:83-85)."""
        return [len(self.Iset[b]) for b in range(1, len(self.localdims))]

    def rank(self):
        return max(self.linkdims())

    def pivoterror(self):
        """max(bonderrors) (tensorci2.jl:157-159)."""
        return float(np.max(self.bonderrors))

    def evaluate(self, points, ctx=None):
        """TensorTrain(tci) evaluated at each row of points (tensortrain.jl:62-64,
        abstracttensortrain.jl:124-132), on the device."""
        return tt_evaluate(self.sitetensors, points, ctx)

    def __call__(self, indexset):
        return float(self.evaluate(np.asarray(indexset).reshape(1, -1))[0])


def _resolve_tolerance(tolerance, pivottolerance):
    """optimize! tolerance/pivottolerance resolution (tensorci2.jl:734-746)."""
    if pivottolerance is not None:
        if tolerance is not None and tolerance != pivottolerance:
            raise ArgumentError(3, "Got different values for pivottolerance and tolerance in "
                                   "optimize!(TCI2). For TCI2, both of these options have the same "
                                   "meaning. Please assign only `tolerance`.")
        warnings.warn("The option `pivottolerance` of `optimize!(tci::TensorCI2, f)` is deprecated. "
                      "Please update your code to use `tolerance`, as `pivottolerance` will be "
                      "removed in the future.", DeprecationWarning, stacklevel=3)
        return pivottolerance
    return 1e-8 if tolerance is None else tolerance


def _holder(f, localdims, initialpivots, kw):
    kw = dict(kw)
    tol = _resolve_tolerance(kw.pop("tolerance", None), kw.pop("pivottolerance", None))
    ps = str(kw.pop("pivotsearch", "full")).lstrip(":")
    if ps not in _abi.PIVOTSEARCH:
        raise ArgumentError(3, f"Unknown pivot search strategy {ps}. Choose from :rook, :full.")
    kw["pivotsearch"] = ps
    kw.pop("checkbatchevaluatable", None)   # device functors are batch evaluators
    kw.pop("verbosity", None)
    kw.pop("loginterval", None)
    state = kw.pop("_state", None)
    if isinstance(f, CachedFunction):
        kw["cache_log2"] = f.log2size
        f = f.f
    if not isinstance(f, DeviceFunctor):
        raise TypeError("f must be a DeviceFunctor (a Julia/Python closure cannot run on the GPU)")
    return ProblemHolder(localdims, f.kind, f.params, f.iparams, initialpivots, state=state, tolerance=tol, **kw)


def _state_of(tci, maxsamplevalue):
    L = len(tci.localdims)
    hist = (tci.Iset_history[-1], tci.Jset_history[-1]) if tci.Iset_history else (None, None)
    return _abi.StateHolder(L, tci.Iset, tci.Jset, hist[0], hist[1], maxsamplevalue)


def tensorci2_from_sets(f, localdims, Iset, Jset, ctx=None):
    """TensorCI2{Float64}(f, localdims, Iset, Jset) (tensorci2.jl:58-72): a
    TCI with the given pivot sets, no history, and maxsamplevalue the
    maximum of |f| over every Iset[b] ⊕ σ ⊕ Jset[b], evaluated on the device.
    Iset[b]/Jset[b] are lists of 1-based multi-indices (b = 0..L-1)."""
    L = len(localdims)
    if len(Iset) != L or len(Jset) != L:
        raise ArgumentError(3, "Iset and Jset need one set per site")
    st = _abi.StateHolder(L, Iset, Jset, None, None, 0.0)
    (t,) = crossinterpolate2_batch([(f, localdims, None, dict(maxiter=0, flags=_abi.FLAG_NO_FINAL_SWEEP,
                                                             _state=st))], ctx=ctx)
    if t.status != 0:
        _raise(t.status, _MESSAGES.get(t.status, ""))
    return t


def optimize(tci, f, ctx=None, **kwargs):
    """optimize!(tci, f; kwargs...) (tensorci2.jl:700-850) continuing from an
    existing TensorCI2 (its Iset, Jset, last history entry and
    maxsamplevalue), on the device.  Updates tci in place, as the reference
    mutates it, and returns (ranks, errors) of this call."""
    kw = dict(kwargs)
    kw["_state"] = _state_of(tci, tci.maxsamplevalue)
    (t,) = crossinterpolate2_batch([(f, tci.localdims, None, kw)], ctx=ctx)
    if t.status != 0:
        _raise(t.status, _MESSAGES.get(t.status, ""))
    for k in ("Iset", "Jset", "Iset_history", "Jset_history", "sitetensors", "pivoterrors", "bonderrors",
              "maxsamplevalue", "status", "niter", "ranks", "errors", "nglobalpivots", "workstats"):
        setattr(tci, k, getattr(t, k))
    return t.ranks, t.errors


def _fetch_cache(lib, b, k, cf):
    """The memo table of solve k of batch b into cf (stats and cachedata)."""
    ne, nl, nh, cnt = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
    _check(lib.tci_batch_cachestats(b, k, C.byref(ne), C.byref(nl), C.byref(nh)))
    _check(lib.tci_batch_cachedata(b, k, 0, None, None, C.byref(cnt)))
    L = len(cf.localdims)
    pts = np.zeros((max(cnt.value, 1), L), np.uint8)
    vals = np.zeros(max(cnt.value, 1))
    _check(lib.tci_batch_cachedata(b, k, cnt.value, _abi.u8ptr(pts), _abi.dptr(vals), C.byref(cnt)))
    cf._data = {tuple(int(v) for v in pts[q]): float(vals[q]) for q in range(cnt.value)}
    cf.stats = dict(entries=ne.value, lookups=nl.value, hits=nh.value)


def capacity_plan(holder):
    """(chi_used, chi_limit) of a problem (tci_capacity_plan)."""
    used, limit = C.c_int64(), C.c_int64()
    _check(_abi.lib().tci_capacity_plan(C.byref(holder.st), C.byref(used), C.byref(limit)))
    return used.value, limit.value


def crossinterpolate2_batch(problems, ctx=None, return_batch=False, error_sample=None):
    """Submit many crossinterpolate2 problems as ONE launch.

    problems: list of (f, localdims, initialpivots_or_None, kwargs).
    Returns a list of TensorCI2 in order (and the device time of the
    launches in ms with return_batch).

    The reference's index sets grow without bound; the device's have a
    capacity fixed at submit.  A solve that ends with TCI_ERR_CAPACITY while
    its capacity is below min(maxbonddim, rank bound) is resubmitted (with
    the other such solves, as one launch) at twice the capacity, so the
    default options succeed wherever the reference does.  A caller that
    sets chi_cap explicitly gets the CAPACITY status instead.

    error_sample=(npts, seed): every successful solve's TT and f are
    evaluated on the device at npts seeded random points (sample_points)
    while the solve is still resident; t.sample_error holds
    (max|TT - f|, max|f|) over them."""
    ctx = ctx or default_context()
    lib = _abi.lib()
    problems = list(problems)
    out = [None] * len(problems)
    chi = [None] * len(problems)     # capacity override per problem (None: as given)
    pending = list(range(len(problems)))
    total_ms = 0.0
    while pending:
        holders = []
        for i in pending:
            f, ld, piv, kw = problems[i]
            kw = dict(kw)
            if chi[i] is not None:
                kw["chi_cap"] = chi[i]
            holders.append(_holder(f, ld, piv, kw))
        arr = (_abi.tci_problem * len(holders))(*[h.st for h in holders])
        b = C.c_void_p()
        _check(lib.tci_solve_submit(ctx.handle, len(holders), arr, C.byref(b)))
        retry = []
        try:
            _check(lib.tci_batch_wait(b))
            for k, i in enumerate(pending):
                t = TensorCI2(b, k, problems[i][1])
                user_cap = int(problems[i][3].get("chi_cap", 0) or 0)
                if t.status == 10 and user_cap == 0:
                    used, limit = capacity_plan(holders[k])
                    if used < limit:
                        chi[i] = int(min(2 * used, limit, 8192))
                        if chi[i] > used:
                            retry.append(i)
                            continue
                if error_sample and t.status == 0:
                    npts, seed = error_sample
                    x = sample_points(problems[i][1], int(npts), seed)
                    me, mf = C.c_double(), C.c_double()
                    _check(lib.tci_batch_evaluate(b, k, x.shape[0], _abi.u8ptr(x), None, None,
                                                  C.byref(me), C.byref(mf)))
                    t.sample_error = (me.value, mf.value)
                if isinstance(problems[i][0], CachedFunction) and t.status == 0:
                    _fetch_cache(lib, b, k, problems[i][0])
                out[i] = t
            ms = C.c_float()
            lib.tci_batch_elapsed_ms(b, C.byref(ms))
            total_ms += ms.value
        finally:
            lib.tci_batch_free(b)
        pending = retry
    if return_batch:
        return out, total_ms
    return out


def crossinterpolate2(ValueType, f, localdims, initialpivots=None, ctx=None, **kwargs):
    """crossinterpolate2(Float64, f, localdims, initialpivots; kwargs...) (tensorci2.jl:943-953).

    Returns (tci, ranks, errors) like the reference; raises the reference's
    errors (ArgumentError, ErrorException) as TCIError subclasses."""
    if ValueType not in (float, np.float64, "Float64"):
        raise TypeError("only Float64 solves run on the device")
    (t,) = crossinterpolate2_batch([(f, localdims, initialpivots, kwargs)], ctx=ctx)
    if t.status != 0:
        _raise(t.status, _MESSAGES.get(t.status, ""))
    return t, t.ranks, t.errors
