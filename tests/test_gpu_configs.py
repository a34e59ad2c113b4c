"""The solve configurations at their full sizes on the GPU against the oracle,
bit for bit: the plain Lorentzian (L=8, d=10, tol 1e-8), the quantics
workload (L=30, d=2, tol 1e-10, maxbonddim 200) and the high-dimensional one
(L=40, d=16, tol 1e-10, maxbonddim 256), each in parity mode (no global
search) and with the optimize! defaults.  Also the capacity retry: a solve
whose rank outgrows the default capacity succeeds with default options."""
import json
import os

import numpy as np
import pytest

import oracle as O
from problems import TABLE, config_specs, oracle_opts
from test_gpu_solves import _check, _tci_problem

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "configs.json")


@pytest.mark.parametrize("spec", config_specs(), ids=lambda s: s[0])
def test_config_solve(tci, ctx, spec):
    print("config", spec[0], flush=True)
    (t,) = tci.crossinterpolate2_batch([_tci_problem(tci, spec)])
    _check(t, spec)
    # and the committed golden of the same configuration (ranks, link dims)
    gold = json.load(open(GOLD))[spec[0]]
    assert list(map(int, t.ranks)) == gold["ranks"]
    assert t.linkdims() == gold["linkdims"]


def test_configs_in_one_launch(tci, ctx):
    sp = config_specs()
    ts = tci.crossinterpolate2_batch([_tci_problem(tci, s) for s in sp])
    for t, s in zip(ts, sp):
        _check(t, s)


def test_capacity_retry_with_default_options(tci, ctx):
    # a random 10^4 table: its rank at the middle bond is 100 > the default
    # capacity of 64, so the first launch ends with CAPACITY and the solve is
    # resubmitted at capacity 100; the result equals the oracle's
    rng = np.random.default_rng(7)
    tab = rng.random((10, 10, 10, 10))
    f = tci.TableFunctor(tab)
    t, ranks, errors = tci.crossinterpolate2(float, f, [10] * 4, tolerance=1e-12)
    assert t.status == 0
    assert max(t.linkdims()) > 64
    o, oranks, oerrors = O.crossinterpolate2(O.Functor(TABLE, tab.ravel(order="F"), [10] * 4), [10] * 4,
                                             tolerance=1e-12)
    assert list(ranks) == list(oranks)
    assert np.array_equal(errors, oerrors)
    for b in range(4):
        assert t.Iset[b] == o.Iset(b) and t.Jset[b] == o.Jset(b)
        assert np.array_equal(t.sitetensors[b], o.sitetensor(b))
