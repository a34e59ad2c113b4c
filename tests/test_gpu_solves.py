"""Whole crossinterpolate2 solves on the GPU against the oracle, bit for bit,
across every functor family (tests/problems.py), one launch for all of them
and each alone."""
import numpy as np
import pytest

import oracle as O
from problems import oracle_opts, specs

pytestmark = pytest.mark.gpu


def _tci_problem(tci, spec):
    name, kind, params, ip, ld, piv, opts = spec
    return (tci.DeviceFunctor(kind, np.asarray(params, float), np.asarray(ip, np.int32)), ld, piv, dict(opts))


def _check(t, spec):
    name, kind, params, ip, ld, piv, opts = spec
    try:
        o, _, _ = O.crossinterpolate2(O.Functor(kind, params, ip), ld, piv, **oracle_opts(opts))
        ost = 0
    except O.OracleError as ex:
        o, ost = None, ex.code
    assert t.status == ost, name
    if ost:
        return
    assert list(t.ranks) == list(o.ranks), name
    assert np.array_equal(t.errors, o.errors), name
    assert t.maxsamplevalue == o.maxsamplevalue, name
    for b in range(len(ld)):
        assert t.Iset[b] == o.Iset(b), (name, b)
        assert t.Jset[b] == o.Jset(b), (name, b)
        assert np.array_equal(t.sitetensors[b], o.sitetensor(b)), (name, b)
    assert np.array_equal(t.pivoterrors, o.pivoterrors), name
    assert np.array_equal(t.bonderrors, o.bonderrors), name


@pytest.mark.parametrize("spec", specs(), ids=lambda s: s[0])
def test_solve_alone(tci, ctx, spec):
    print("solve", spec[0], flush=True)
    (t,) = tci.crossinterpolate2_batch([_tci_problem(tci, spec)])
    _check(t, spec)


def test_all_in_one_launch(tci, ctx):
    sp = specs()
    ts = tci.crossinterpolate2_batch([_tci_problem(tci, s) for s in sp])
    for t, s in zip(ts, sp):
        _check(t, s)


def test_error_statuses(tci, ctx):
    with pytest.raises(tci.TCIError) as e:
        tci.crossinterpolate2(float, tci.Lorentz(0.0), [2] * 5)
    assert e.value.status == 4
    with pytest.raises(tci.ArgumentError):
        tci.crossinterpolate2(float, tci.Lorentz(1.0), [2] * 5, tolerance=0.0)
    with pytest.raises(tci.TCIError) as e:
        tci.crossinterpolate2(float, tci.Lorentz(1.0), [10] * 8, tolerance=1e-12, chi_cap=4, maxiter=3)
    assert e.value.status == 10
