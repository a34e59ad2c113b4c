"""Served mode: requests admitted while others run (continuous batching over
several streams) give the same result, bit for bit, as each request solved
alone."""
import numpy as np
import pytest

from problems import served_mix, specs

pytestmark = pytest.mark.gpu


def _req(tci, spec):
    name, kind, params, ip, ld, piv, opts = spec
    return (tci.DeviceFunctor(kind, np.asarray(params, float), np.asarray(ip, np.int32)), ld, piv, dict(opts))


def _same(a, b, name):
    assert a.status == b.status, name
    if a.status:
        return
    assert list(a.ranks) == list(b.ranks), name
    assert np.array_equal(a.errors, b.errors), name
    assert a.maxsamplevalue == b.maxsamplevalue, name
    for s in range(len(a.localdims)):
        assert a.Iset[s] == b.Iset[s], (name, s)
        assert a.Jset[s] == b.Jset[s], (name, s)
        assert np.array_equal(a.sitetensors[s], b.sitetensors[s]), (name, s)


def test_served_equals_alone(tci, ctx):
    from importlib import import_module
    serving = import_module("tci_amd.serving")
    reqs = [_req(tci, s) for s in served_mix(24, seed=11)]
    reqs += [_req(tci, s) for s in specs() if s[0] in ("lorentz8", "gauss_sum", "random_tt", "weighted_lorentz")]
    arrivals = serving.poisson_arrivals(len(reqs), rate_per_s=400.0, seed=3)
    srv = serving.TCIServer(ctx, streams=3, max_batch=8)
    res, stats = srv.run(reqs, arrivals, keep_tci=True)
    assert stats.batches > 1                      # admitted over time, not as one batch
    for i, r in enumerate(reqs):
        (alone,) = tci.crossinterpolate2_batch([r], ctx=ctx)
        _same(res[i].tci, alone, f"request {i}")
        assert res[i].latency >= 0


def test_served_mix_equals_oracle(tci, ctx):
    """The served workload's solves (one launch, mixed sizes) equal the oracle."""
    from test_gpu_solves import _check
    mix = served_mix(12, seed=21)
    out = tci.crossinterpolate2_batch([_req(tci, s) for s in mix], ctx=ctx)
    for t, s in zip(out, mix):
        _check(t, s)
