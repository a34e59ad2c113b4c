"""The memo table of f on the GPU (CachedFunction, cachedfunction.jl:8-203):
a solve through the wrapper equals the plain solve and the oracle bit for
bit, the table answers repeated evaluations, and cachedata holds f's values
at the points it decodes from the keys."""
import numpy as np
import pytest

import oracle as O
from problems import c4_spec, oracle_opts, specs
from test_gpu_solves import _check, _tci_problem

pytestmark = pytest.mark.gpu

_SPECS = [s for s in specs() if s[0] in {"lorentz8", "gauss_sum", "quantics_gc", "random_tt"}] + [c4_spec(0)]


@pytest.mark.parametrize("spec", _SPECS, ids=lambda s: s[0])
def test_cached_solve(tci, ctx, spec):
    name, kind, params, ip, ld, piv, opts = spec
    print("cache", name, flush=True)
    f, _, _, kw = _tci_problem(tci, spec)
    cf = tci.CachedFunction(f, ld, keybits=256, log2size=18)
    (t,) = tci.crossinterpolate2_batch([(cf, ld, piv, dict(kw))])
    _check(t, spec)                                   # equals the oracle bit for bit
    st = cf.stats
    assert st["entries"] == 1 << 18
    assert 0 < st["hits"] < st["lookups"]
    data = cf.cachedata()
    assert 0 < len(data) <= st["lookups"]
    pts = np.array(list(data.keys())[:5000], np.int32)
    want = O.eval_points(O.Functor(kind, params, ip), pts)
    got = np.array([data[tuple(p)] for p in pts.tolist()])
    assert np.array_equal(got, want)
    assert cf.iscached(pts[0])
