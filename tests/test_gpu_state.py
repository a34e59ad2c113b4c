"""Solves started from pivot sets on the GPU: the constructor
TensorCI2{Float64}(f, localdims, Iset, Jset) mirroring
test_tensorci2.jl:461-475, and optimize! continuing an existing TCI, both
against the oracle bit for bit."""
import numpy as np
import pytest

import oracle as O
from problems import TABLE, c4_spec, oracle_opts, specs

pytestmark = pytest.mark.gpu


def test_initialize_with_local_pivots_list(tci, ctx):
    rng = np.random.default_rng(1234)
    M = rng.random((10, 10))
    f = tci.TableFunctor(M)
    t, ranks, errors = tci.crossinterpolate2(float, f, [10, 10], maxbonddim=5)
    t2 = tci.tensorci2_from_sets(f, [10, 10], t.Iset, t.Jset)
    assert t2.maxsamplevalue == t.maxsamplevalue
    assert t2.Iset == t.Iset and t2.Jset == t.Jset
    assert t2.Iset_history == []
    o, _, _ = O.crossinterpolate2(O.Functor(TABLE, M.ravel(order="F"), [10, 10]), [10, 10], maxbonddim=5)
    assert t.Iset == [o.Iset(b) for b in range(2)]


_REENTRY = [s for s in specs() if s[0] in {"lorentz5_tight", "gauss_sum", "random_tt", "weighted_lorentz"}] + \
    [c4_spec(0)]


@pytest.mark.parametrize("spec", _REENTRY, ids=lambda s: s[0])
def test_optimize_reentry(tci, ctx, spec):
    name, kind, params, ip, ld, piv, opts = spec
    print("reentry", name, flush=True)
    f = tci.DeviceFunctor(kind, np.asarray(params, float), np.asarray(ip, np.int32))
    o1 = dict(oracle_opts(opts), maxiter=2)
    o2 = dict(oracle_opts(opts), maxiter=4)
    t, _, _ = tci.crossinterpolate2(float, f, ld, piv, **o1)
    assert t.Iset_history              # a solve that swept has a history entry
    ranks, errors = tci.optimize(t, f, **o2)
    of = O.Functor(kind, params, ip)
    o, _, _ = O.crossinterpolate2(of, ld, piv, **o1)
    r2, e2 = o.optimize(of, **o2)
    assert list(ranks) == list(r2)
    assert np.array_equal(errors, e2)
    assert t.maxsamplevalue == o.maxsamplevalue
    for b in range(len(ld)):
        assert t.Iset[b] == o.Iset(b) and t.Jset[b] == o.Jset(b)
        assert np.array_equal(t.sitetensors[b], o.sitetensor(b))
