"""Tensor-train evaluation on the device (abstracttensortrain.jl:124-132)
against the oracle's restatement, bit for bit, and the normalized error
max|TT - f| / max|f| of every full-size configuration on 10^5 seeded random
points, with f evaluated on the host in FP64 by the oracle."""
import json
import os

import numpy as np
import pytest

import oracle as O
from problems import TT, config_specs, oracle_opts, random_tt
from test_gpu_solves import _tci_problem

pytestmark = pytest.mark.gpu

NPTS = 100_000
HERE = os.path.dirname(os.path.abspath(__file__))
ERRS = json.load(open(os.path.join(HERE, "golden", "config_errors.json")))


def test_tt_evaluate_random_tt(tci, ctx):
    # a random TT with uneven bonds: device values equal the oracle's chain
    # bit for bit, and numpy's contraction to rounding
    _, _, cores = random_tt([1, 3, 7, 5, 2, 1], [2, 4, 3, 5, 2], 11)
    x = tci.sample_points([2, 4, 3, 5, 2], 5000, seed=3)
    got = tci.tt_evaluate(cores, x)
    want = O.tt_evaluate(cores, x)
    assert np.array_equal(got, want)
    ref = np.array([np.linalg.multi_dot([c[:, int(v) - 1, :] for c, v in zip(cores, p)]).item() for p in x[:200]])
    np.testing.assert_allclose(got[:200], ref, rtol=1e-13)
    # the TT functor (the same chain, in the functor code) agrees too
    flat, ip, _ = random_tt([1, 3, 7, 5, 2, 1], [2, 4, 3, 5, 2], 11)
    fv = O.eval_points(O.Functor(TT, flat, ip), x.astype(np.int32))
    assert np.array_equal(got, fv)


def test_tt_evaluate_argument_errors(tci, ctx):
    _, _, cores = random_tt([1, 2, 1], [3, 3], 1)
    with pytest.raises(tci.ArgumentError):
        tci.tt_evaluate(cores, np.array([[1, 2, 3]]))     # wrong length (the reference's message)
    with pytest.raises(tci.ArgumentError):
        tci.tt_evaluate(cores, np.array([[1, 4]]))        # index out of range
    assert tci.tt_evaluate(cores, np.zeros((0, 2))).shape == (0,)


@pytest.mark.parametrize("spec", config_specs(), ids=lambda s: s[0])
def test_config_error_on_random_points(tci, ctx, spec):
    name, kind, params, ip, ld, piv, opts = spec
    print("config error", name, flush=True)
    (t,) = tci.crossinterpolate2_batch([_tci_problem(tci, spec)], error_sample=(NPTS, 12345))
    assert t.status == 0
    x = tci.sample_points(ld, NPTS, seed=12345)
    tt = t.evaluate(x)
    # the device TT equals the oracle's TT (the same solve) at every point
    o, _, _ = O.crossinterpolate2(O.Functor(kind, params, ip), ld, piv, **oracle_opts(opts))
    assert np.array_equal(tt, O.tt_evaluate(o.sitetensors(), x.astype(np.int32)))
    # the error against f evaluated on the host in FP64
    fh = O.eval_points(O.Functor(kind, params, ip), x.astype(np.int32))
    err = np.max(np.abs(tt - fh)) / np.max(np.abs(fh))
    # the device's own f gives the same figures (its f bits equal the host's)
    me, mf = t.sample_error
    assert mf == np.max(np.abs(fh)) and me == np.max(np.abs(tt - fh))
    # recorded per configuration: the oracle (the reference algorithm) has
    # exactly this error; where it exceeds the tolerance the reference's
    # pivot-based error estimate missed features of f (see the golden)
    assert err == ERRS[name]["err_maxf"]
    assert me / t.maxsamplevalue == ERRS[name]["err_maxsample"]
    if ERRS[name]["meets_tol"]:
        assert me / t.maxsamplevalue <= opts["tolerance"]
