"""GPU parity: the HIP path (through the C ABI) against the oracle.

Bit-exact for pivot sequences, permutations, ranks, pivot sets and every
FP64 value the engine computes with a fixed operation order (the oracle
restates the same order); the checks to a tolerance (TT evaluation against
the functor, the error estimates) are in tests/test_gpu_tt_eval.py and
tests/test_gpu_tt.py.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _same_float(a, b):
    return a == b or (np.isnan(a) and np.isnan(b))


def test_context(tci, ctx):
    assert tci._abi.lib().tci_version().decode().startswith("tci-mi355x")


def test_rrlu_golden_pivots(tci, ctx):
    with open(os.path.join(GOLD, "rrlu_pivots.json")) as fh:
        cases = json.load(fh)
    for c in cases:
        print("case", c["name"], flush=True)
        lu = tci.rrlu(np.array(c["A"]), **c["kwargs"])
        assert lu.rowpermutation.tolist() == c["rowpermutation"], c["name"]
        assert lu.colpermutation.tolist() == c["colpermutation"], c["name"]
        assert lu.npivot == c["npivot"], c["name"]
        err = c["error"] if not isinstance(c["error"], str) else float(c["error"])
        assert _same_float(lu.error, err), c["name"]
        assert np.array_equal(lu.L, np.array(c["L"]).reshape(lu.L.shape)), c["name"]
        assert np.array_equal(lu.U, np.array(c["U"]).reshape(lu.U.shape)), c["name"]


@pytest.mark.parametrize("seed", range(4))
def test_rrlu_random_vs_oracle(tci, ctx, seed):
    rng = np.random.default_rng(seed)
    for t in range(5):
        m, n = (int(x) for x in rng.integers(1, 300, 2))
        kind = t % 3
        if kind == 0:
            A = rng.standard_normal((m, n))
        elif kind == 1:
            r = int(rng.integers(1, 12))
            A = rng.standard_normal((m, r)) @ rng.standard_normal((r, n)) + 1e-12 * rng.standard_normal((m, n))
        else:
            A = rng.integers(-1, 2, (m, n)).astype(float)
        lo = bool(rng.integers(0, 2))
        kw = dict(reltol=1e-10, leftorthogonal=lo)
        print("seed", seed, "case", t, m, n, kw, flush=True)
        a = tci.rrlu(A, **kw)
        b = O.rrlu(A, **kw)
        assert np.array_equal(a.rowpermutation, b.rowpermutation)
        assert np.array_equal(a.colpermutation, b.colpermutation)
        assert a.npivot == b.npivot
        assert _same_float(a.error, b.error)
        assert np.array_equal(a.L, b.L) and np.array_equal(a.U, b.U)


def _cmp_tci(t, o, L):
    assert list(t.ranks) == list(o.ranks)
    assert np.array_equal(t.errors, o.errors)
    assert t.maxsamplevalue == o.maxsamplevalue
    for b in range(L):
        assert t.Iset[b] == o.Iset(b), b
        assert t.Jset[b] == o.Jset(b), b
        assert np.array_equal(t.sitetensors[b], o.sitetensor(b)), b
    assert np.array_equal(t.pivoterrors, o.pivoterrors)
    assert np.array_equal(t.bonderrors, o.bonderrors)


@pytest.mark.parametrize("nsearch", [0, 5])
def test_crossinterpolate2_lorentz8_bitwise(tci, ctx, nsearch):
    """README.md:24-28 example: L=8, d=10, tol 1e-8, initial pivot ones."""
    t, ranks, errors = tci.crossinterpolate2(float, tci.Lorentz(1.0), [10] * 8, tolerance=1e-8,
                                             nsearchglobalpivot=nsearch)
    o, _, _ = O.crossinterpolate2(O.Functor(1, [1.0]), [10] * 8, tolerance=1e-8, nsearchglobalpivot=nsearch)
    _cmp_tci(t, o, 8)
    with open(os.path.join(GOLD, "tci_c1.json")) as fh:
        g = json.load(fh)[f"nsearch{nsearch}"]
    assert list(ranks) == g["ranks"]
    assert t.linkdims() == g["linkdims"]


def test_batch_alone_equals_mixed(tci, ctx):
    """Each solve's result is the same alone and inside a mixed batch."""
    rng = np.random.default_rng(7)
    probs = [(tci.Lorentz(1.0), [10] * 8, None, dict(tolerance=1e-8)),
             (tci.Lorentz(0.5), [10] * 5, None, dict(tolerance=1e-12, maxiter=200)),
             (tci.WeightedLorentz(rng.uniform(0.5, 2, 6), 16.0), [16] * 6, None, dict(tolerance=1e-10)),
             (tci.Lorentz(1.0), [10] * 5, [[1, 1, 1, 1, 1], [10, 8, 10, 4, 4]], dict(tolerance=1e-12, maxiter=50))]
    mixed = tci.crossinterpolate2_batch(probs)
    for p, m in zip(probs, mixed):
        (alone,) = tci.crossinterpolate2_batch([p])
        assert alone.status == 0 and m.status == 0
        assert list(alone.ranks) == list(m.ranks)
        assert np.array_equal(alone.errors, m.errors)
        for b in range(len(p[1])):
            assert alone.Iset[b] == m.Iset[b]
            assert np.array_equal(alone.sitetensors[b], m.sitetensors[b])


def test_pivoterrors_diag(tci, ctx):
    """test_tensorci2.jl:27-39: pivoterrors == [1.0, 1e-5, 0.0] exactly."""
    tab = np.diag([1.0, 1e-5, 0.0])
    t, _, _ = tci.crossinterpolate2(float, tci.TableFunctor(tab), [3, 3], [[1, 1]], tolerance=1e-8)
    assert t.pivoterrors.tolist() == [1.0, 1e-5, 0.0]


def test_rrlu_batched_c2_shapes_vs_oracle(tci, ctx):
    """Matrices of the batched-rrLU workload (rows/cols up to 1024, A = X·Y +
    1e-12·E, rank up to 128, reltol 1e-10), uneven sizes in one launch: the
    pivot sequences, pivot counts and errors equal the oracle's."""
    rng = np.random.default_rng(7)
    shapes = [(1024, 1024, 128), (1000, 37, 20), (33, 900, 64), (517, 611, 8), (256, 256, 100)]
    mats = []
    for m, n, r in shapes:
        X = rng.standard_normal((m, r))
        Y = rng.standard_normal((r, n))
        mats.append(X @ Y + 1e-12 * rng.standard_normal((m, n)))
    res = tci.rrlu_batched(mats, reltol=1e-10, ctx=ctx)
    for A, g in zip(mats, res):
        o = O.rrlu(A, reltol=1e-10)
        assert g["status"] == 0
        assert g["npivot"] == o.npivot, A.shape
        assert np.array_equal(g["rowpermutation"], o.rowpermutation), A.shape
        assert np.array_equal(g["colpermutation"], o.colpermutation), A.shape
        assert g["error"] == o.error, A.shape


def test_ctx_stream_carries_the_batched_rrlu(tci, ctx):
    """tci_ctx_stream hands out the stream the batched rrLU runs on: timing
    events recorded on it around the call bracket the launch, and the
    results still equal the oracle's."""
    import ctypes as C
    import torch
    lib = tci._abi.lib()
    sp = C.c_void_p()
    assert lib.tci_ctx_stream(ctx.handle, C.byref(sp)) == 0 and sp.value
    s = torch.cuda.ExternalStream(sp.value)
    rng = np.random.default_rng(11)
    mats = [rng.standard_normal((60, 8)) @ rng.standard_normal((8, 50)) for _ in range(3)]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    res = tci.rrlu_batched(mats, reltol=1e-10, ctx=ctx)
    e1.record(s)
    e1.synchronize()
    assert e0.elapsed_time(e1) > 0.0
    for A, g in zip(mats, res):
        o = O.rrlu(A, reltol=1e-10)
        assert g["npivot"] == o.npivot
        assert np.array_equal(g["rowpermutation"], o.rowpermutation)
        assert np.array_equal(g["colpermutation"], o.colpermutation)
