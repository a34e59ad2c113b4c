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


def _nan_requests(tci):
    """A Lorentzian with a = NaN (maxsamplevalue is zero! at the start) and
    a 5×5 table with one NaN entry (the solve runs, its errors are NaN)."""
    tab = np.arange(1.0, 26.0).reshape(5, 5)
    tab[3, 2] = np.nan
    return [(tci.Lorentz(float("nan")), [10] * 5, None, dict(tolerance=1e-8)),
            (tci.TableFunctor(tab), [5, 5], None, dict(tolerance=1e-8))]


def test_served_nan_requests_do_not_disturb_the_batch(tci, ctx):
    """NaN requests inside served batches end with the oracle's status or
    NaN history, and every other request equals its solve alone."""
    from importlib import import_module
    import oracle as O
    serving = import_module("tci_amd.serving")
    good = [_req(tci, s) for s in served_mix(10, seed=31)]
    bad = _nan_requests(tci)
    reqs = good[:5] + bad[:1] + good[5:] + bad[1:]
    arrivals = np.zeros(len(reqs))                 # everything in one batch
    srv = serving.TCIServer(ctx, streams=2, max_batch=64)
    res, stats = srv.run(reqs, arrivals, keep_tci=True)
    assert stats.batches == 1
    assert res[5].status == 4                      # maxsamplevalue is zero!
    tab = np.asarray(bad[1][0].params).reshape((5, 5), order="F")
    t, _, _ = O.crossinterpolate2(O.Functor(6, tab.ravel(order="F"), [5, 5]), [5, 5], tolerance=1e-8)
    assert np.array_equal(res[-1].errors, t.errors, equal_nan=True)
    assert list(res[-1].ranks) == list(t.ranks)
    for i, r in enumerate(reqs):
        if i in (5, len(reqs) - 1):
            continue
        (alone,) = tci.crossinterpolate2_batch([r], ctx=ctx)
        _same(res[i].tci, alone, f"request {i}")


def test_served_small_hbm_budget(tci):
    """With an HBM budget that holds only a few solves' arenas at once the
    loop admits smaller batches and waits for running ones; every result
    still equals the solve alone."""
    from importlib import import_module
    serving = import_module("tci_amd.serving")
    reqs = [_req(tci, s) for s in served_mix(16, seed=41)]
    small = tci.Context(0, hbm_budget_bytes=96 << 20)
    try:
        srv = serving.TCIServer(small, streams=3, max_batch=16)
        res, stats = srv.run(reqs, np.zeros(len(reqs)), keep_tci=True)
        assert stats.batches > 1
        for i, r in enumerate(reqs):
            (alone,) = tci.crossinterpolate2_batch([r])
            _same(res[i].tci, alone, f"request {i}")
    finally:
        small.close()
