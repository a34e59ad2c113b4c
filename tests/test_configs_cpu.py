"""The full-size solve configurations on the CPU: the oracle reproduces the
committed goldens (tests/golden/configs.json: ranks, link dims, errors and a
digest of every pivot set), and the device engine's own code built for the
CPU (one-thread team, with and without the on-chip areas) equals the oracle
bit for bit on each of them.  No GPU."""
import json
import os

import numpy as np
import pytest

import oracle as O
from problems import config_specs, oracle_opts
from test_engine_host import _compare

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "configs.json")))


def _digest(t, L):
    import hashlib
    h = hashlib.sha256()
    for b in range(L):
        h.update(json.dumps([t.Iset(b), t.Jset(b)]).encode())
    return h.hexdigest()


@pytest.mark.parametrize("spec", config_specs(), ids=lambda s: s[0])
def test_oracle_reproduces_config_golden(spec):
    name, kind, params, ip, ld, piv, opts = spec
    t, r, e = O.crossinterpolate2(O.Functor(kind, params, ip), ld, piv, **oracle_opts(opts))
    g = GOLD[name]
    assert r.tolist() == g["ranks"]
    assert e.tolist() == g["errors"]
    assert t.linkdims() == g["linkdims"]
    assert t.nglobalpivots.tolist() == g["nglobalpivots"]
    assert t.maxsamplevalue == g["maxsamplevalue"]
    assert _digest(t, len(ld)) == g["sets_sha256"]
    # the invariants the reference's tests check on every TCI: square pivot
    # matrices, bond dims within the rank bound and the error bound met
    L = len(ld)
    for b in range(L - 1):
        assert len(t.Iset(b + 1)) == len(t.Jset(b))
        lp = int(np.prod(ld[:b + 1], dtype=float)) if b < 30 else 1 << 62
        assert len(t.Iset(b + 1)) <= min(lp, int(opts.get("maxbonddim", 1 << 62)))
    assert e[-1] <= opts["tolerance"]


@pytest.mark.parametrize("spec", config_specs(), ids=lambda s: s[0])
def test_engine_cpu_build_equals_oracle_on_config(spec):
    _compare(spec, 1)


@pytest.fixture
def host_lds(monkeypatch):
    monkeypatch.setenv("EH_LDS", "1")


@pytest.mark.parametrize("spec", [s for s in config_specs() if "parity" in s[0]], ids=lambda s: s[0])
def test_engine_cpu_build_with_onchip_areas_on_config(spec, host_lds):
    _compare(spec, 1)


def test_oracle_tt_evaluate_pinned():
    """The oracle's TT evaluation (abstracttensortrain.jl:124-132): equal to
    the TT functor's chain bit for bit, to numpy's matrix products to
    rounding, and exact on a rank-1 TT whose value is a product of numbers."""
    from problems import TT, random_tt
    flat, ip, cores = random_tt([1, 3, 7, 5, 2, 1], [2, 4, 3, 5, 2], 11)
    rng = np.random.default_rng(0)
    x = np.stack([rng.integers(1, d + 1, 3000) for d in [2, 4, 3, 5, 2]], 1).astype(np.int32)
    got = O.tt_evaluate(cores, x)
    assert np.array_equal(got, O.eval_points(O.Functor(TT, flat, ip), x))
    ref = np.array([np.linalg.multi_dot([c[:, v - 1, :] for c, v in zip(cores, p)]).item() for p in x])
    np.testing.assert_allclose(got, ref, rtol=1e-13)
    # rank 1 with power-of-two entries: every product is exact
    r1 = [np.array([[[2.0 ** (b + s)] for s in range(3)]]) for b in range(4)]
    y = np.array([[1, 2, 3, 1], [3, 3, 3, 3]], np.int32)
    want = [np.prod([2.0 ** (b + int(v) - 1) for b, v in enumerate(p)]) for p in y]
    assert O.tt_evaluate(r1, y).tolist() == want


@pytest.mark.parametrize("name", sorted(json.load(open(os.path.join(HERE, "golden", "config_errors.json")))))
def test_oracle_reproduces_config_errors(name):
    g = json.load(open(os.path.join(HERE, "golden", "config_errors.json")))[name]
    spec = [s for s in config_specs() if s[0] == name][0]
    _, kind, params, ip, ld, piv, opts = spec
    f = O.Functor(kind, params, ip)
    t, _, _ = O.crossinterpolate2(f, ld, piv, **oracle_opts(opts))
    rng = np.random.default_rng(g["seed"])
    x = np.stack([rng.integers(1, int(d) + 1, g["npts"]) for d in ld], 1).astype(np.int32)
    tt = O.tt_evaluate(t.sitetensors(), x)
    fv = O.eval_points(f, x)
    assert float(np.max(np.abs(tt - fv))) == g["max_abs_err"]
    assert float(np.max(np.abs(fv))) == g["max_abs_f"]
    # the TT interpolates f at every pivot of the final sets (the
    # interpolation property of the cross, to rounding)
    L = len(ld)
    for b in range(L - 1):
        for i in t.Iset(b + 1)[:4]:
            for j in t.Jset(b)[:4]:
                p = np.array([i + j], np.int32)
                assert abs(O.tt_evaluate(t.sitetensors(), p)[0] - O.eval_points(f, p)[0]) <= 1e-9 * t.maxsamplevalue
