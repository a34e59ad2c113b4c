"""The memo table of f (CachedFunction, cachedfunction.jl:8-203) in the
device engine's own code built for the CPU: solves with and without it give
the same bits, the table answers repeated evaluations, and the key and its
overflow check follow the reference.  No GPU."""
import numpy as np
import pytest

import eh
import oracle as O
from conftest import load_pkg
from problems import c4_spec, oracle_opts, specs

pkg = load_pkg()
_abi = pkg._abi

_SPECS = [s for s in specs() if s[0] in {"lorentz8", "gauss_sum", "random_tt", "quantics_gc", "weighted_lorentz"}]


@pytest.mark.parametrize("spec", _SPECS, ids=lambda s: s[0])
def test_cached_solve_equals_plain(spec):
    name, kind, params, ip, ld, piv, opts = spec
    plain = eh.EngineSolve(_abi.ProblemHolder(ld, kind, params, ip, piv, **opts))
    cached = eh.EngineSolve(_abi.ProblemHolder(ld, kind, params, ip, piv, cache_log2=16, **opts))
    assert plain.status == cached.status == 0
    r1, e1, g1 = plain.history()
    r2, e2, g2 = cached.history()
    assert list(r1) == list(r2) and np.array_equal(e1, e2) and list(g1) == list(g2)
    assert plain.maxsample == cached.maxsample
    for b in range(len(ld)):
        assert plain.set(0, b) == cached.set(0, b) and plain.set(1, b) == cached.set(1, b)
        assert np.array_equal(plain.sitetensor(b), cached.sitetensor(b))


def test_key_and_overflow_follow_the_reference():
    f = pkg.SumFunctor(1.0)
    cf = pkg.CachedFunction(f, [3, 4, 5])
    # coeffs 1, 3, 12 (cachedfunction.jl:13-18)
    assert cf.coeffs == [1, 3, 12]
    assert cf.encodecachekey([3, 4, 5]) == 2 + 3 * 3 + 12 * 4
    for key in (0, 7, 59):
        assert cf.encodecachekey(cf.decodecachekey(key)) == key
    # 16^40 = 2^160 points overflow a 128-bit key, not a 256-bit one
    with pytest.raises(pkg.TCIError):
        pkg.CachedFunction(pkg.WeightedLorentz(np.ones(40)), [16] * 40, keybits=128)
    pkg.CachedFunction(pkg.WeightedLorentz(np.ones(40)), [16] * 40, keybits=256)
    # the device refuses a domain beyond 2^256 points, and a bad table size
    import ctypes as C
    used, lim = C.c_int64(), C.c_int64()
    h = _abi.ProblemHolder([255] * 33, 5, [1.0], cache_log2=16)
    assert pkg._abi.lib().tci_capacity_plan(C.byref(h.st), C.byref(used), C.byref(lim)) == 3
    h = _abi.ProblemHolder([16] * 40, 5, [1.0], cache_log2=3)
    assert pkg._abi.lib().tci_capacity_plan(C.byref(h.st), C.byref(used), C.byref(lim)) == 3
    h = _abi.ProblemHolder([16] * 40, 5, [1.0], cache_log2=16)
    assert pkg._abi.lib().tci_capacity_plan(C.byref(h.st), C.byref(used), C.byref(lim)) == 0
