"""Solves that start from pivot sets instead of initial pivots, with the
device engine's own code built for the CPU, against the oracle: the
constructor TensorCI2{T}(f, localdims, Iset, Jset) (tensorci2.jl:58-72,
maxsamplevalue over every Iset[b] ⊕ σ ⊕ Jset[b]) and optimize! continuing an
existing TCI (its sets, the last history entry and maxsamplevalue).  No GPU."""
import numpy as np
import pytest

import eh
import oracle as O
from conftest import load_pkg
from problems import TABLE, c3_spec, oracle_opts, specs

_abi = load_pkg()._abi


def _eh_state(e, L, with_hist=True):
    I = [e.set(0, b) for b in range(L)]
    J = [e.set(1, b) for b in range(L)]
    hI = [e.set(2, b) for b in range(L)] if with_hist else None
    hJ = [e.set(3, b) for b in range(L)] if with_hist else None
    return _abi.StateHolder(L, I, J, hI, hJ, e.maxsample)


def _reconstructed_max(f, ld, I, J):
    # reconstractglobalpivotsfromijset (tensorci2.jl:171-190)
    pts = [i + [v] + j for b in range(len(ld)) for i in I[b] for j in J[b] for v in range(1, ld[b] + 1)]
    return float(np.max(np.abs(O.eval_points(f, np.array(pts, np.int32)))))


def test_constructor_from_sets_cpu():
    # test_tensorci2.jl:461-475: a 10×10 random matrix as f, maxbonddim 5
    rng = np.random.default_rng(1234)
    M = rng.random((10, 10))
    ld = [10, 10]
    f = O.Functor(TABLE, M.ravel(order="F"), [10, 10])
    t, _, _ = O.crossinterpolate2(f, ld, maxbonddim=5)
    I = [t.Iset(b) for b in range(2)]
    J = [t.Jset(b) for b in range(2)]
    st = _abi.StateHolder(2, I, J, None, None, 0.0)
    h = _abi.ProblemHolder(ld, TABLE, M.ravel(order="F"), [10, 10], None, state=st, maxiter=0,
                           flags=_abi.FLAG_NO_FINAL_SWEEP)
    e = eh.EngineSolve(h)
    assert e.status == 0
    assert [e.set(0, b) for b in range(2)] == I and [e.set(1, b) for b in range(2)] == J
    assert e.maxsample == _reconstructed_max(f, ld, I, J)
    assert e.maxsample == t.maxsamplevalue        # the reference test's first assertion


_REENTRY = [s for s in specs() if s[0] in {"lorentz5_tight", "gauss_sum", "random_tt", "lorentz8_nosearch"}] + \
    [c3_spec(0)]


@pytest.mark.parametrize("spec", _REENTRY, ids=lambda s: s[0])
def test_optimize_reentry_cpu(spec):
    """optimize! twice on one TCI (the oracle, in place) equals a device solve
    of the first call followed by a device solve started from its state."""
    name, kind, params, ip, ld, piv, opts = spec
    o1 = dict(oracle_opts(opts), maxiter=2)
    o2 = dict(oracle_opts(opts), maxiter=4)
    f = O.Functor(kind, params, ip)
    t, _, _ = O.crossinterpolate2(f, ld, piv, **o1)
    r2, e2 = t.optimize(f, **o2)
    h1 = _abi.ProblemHolder(ld, kind, params, ip, piv, **o1)
    e1 = eh.EngineSolve(h1)
    assert e1.status == 0
    st = _eh_state(e1, len(ld))
    h2 = _abi.ProblemHolder(ld, kind, params, ip, None, state=st, **o2)
    e = eh.EngineSolve(h2)
    assert e.status == 0
    r, er, ng = e.history()
    assert list(r) == list(r2)
    assert np.array_equal(er, e2)
    assert e.maxsample == t.maxsamplevalue
    for b in range(len(ld)):
        assert e.set(0, b) == t.Iset(b) and e.set(1, b) == t.Jset(b)
        assert np.array_equal(e.sitetensor(b), t.sitetensor(b))


def test_state_is_checked():
    ld = [3, 3]
    bad = _abi.StateHolder(2, [[[]], [[4]]], [[[1]], [[]]], None, None, 0.0)   # 4 > d = 3
    h = _abi.ProblemHolder(ld, TABLE, np.ones(9), [3, 3], None, state=bad)
    lib = load_pkg()._abi.lib()
    import ctypes as C
    used, lim = C.c_int64(), C.c_int64()
    assert lib.tci_capacity_plan(C.byref(h.st), C.byref(used), C.byref(lim)) == 3
