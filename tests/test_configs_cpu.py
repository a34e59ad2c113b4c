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
