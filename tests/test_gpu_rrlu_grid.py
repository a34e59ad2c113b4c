"""The cooperative whole-chip rrLU (one workgroup per CU, logical row and
column swaps, one grid barrier per pivot) against the oracle bit for bit:
random, low-rank and tie-heavy matrices in both orientations, NaN verdicts,
shapes with one row or one column, sides up to 1 500, and one bond of the
high-dimensional workload's size — 4 352 × 4 352 with 256 pivots
(matrixlu.jl:141-181).  Selected with tci_ctx_set_rrlu_algo(GRID); AUTO
picks it for a single large matrix."""
import time

import numpy as np
import pytest

import oracle as O
from test_engine_host import _nan_cases, _rrlu_cases

pytestmark = pytest.mark.gpu

AUTO, GRID = 0, 3


@pytest.fixture
def grid_ctx(tci, ctx):
    c = tci.Context(0)
    assert tci._abi.lib().tci_ctx_set_rrlu_algo(c.handle, GRID) == 0
    yield c
    c.close()


def _same(a, b):
    return a == b or (np.isnan(a) and np.isnan(b))


def _oracle(A, **kw):
    try:
        return 0, O.rrlu(A, **kw)
    except O.OracleError as e:
        return e.code, None


def _check(tci, ctx, A, **kw):
    st, b = _oracle(A, **kw)
    try:
        a = tci.rrlu(A, ctx=ctx, **kw)
        ast = 0
    except tci.TCIError as e:
        a, ast = None, e.status
    assert ast == st, (A.shape, kw)
    if st:
        return
    assert a.npivot == b.npivot, (A.shape, kw)
    assert np.array_equal(a.rowpermutation, b.rowpermutation), (A.shape, kw)
    assert np.array_equal(a.colpermutation, b.colpermutation), (A.shape, kw)
    assert _same(a.error, b.error), (A.shape, kw)
    assert np.array_equal(a.L, b.L) and np.array_equal(a.U, b.U), (A.shape, kw)


@pytest.mark.parametrize("seed", range(3))
def test_grid_random_lowrank_ties(tci, grid_ctx, seed):
    for A, lo, rt in _rrlu_cases(500 + seed, 300):
        print("grid case", seed, A.shape, lo, rt, flush=True)
        _check(tci, grid_ctx, A, reltol=rt, leftorthogonal=lo)


@pytest.mark.parametrize("seed", range(2))
def test_grid_nan_verdicts(tci, grid_ctx, seed):
    for A, lo in _nan_cases(700 + seed):
        print("grid nan case", seed, A.shape, lo, flush=True)
        _check(tci, grid_ctx, A, leftorthogonal=lo)


def test_grid_thin_and_maxrank(tci, grid_ctx):
    rng = np.random.default_rng(11)
    for shape in ((1, 1), (1, 57), (57, 1), (3, 700), (700, 3)):
        A = rng.standard_normal(shape)
        for lo in (True, False):
            _check(tci, grid_ctx, A, leftorthogonal=lo)
    A = rng.standard_normal((90, 80))
    for mr in (1, 2, 17):
        _check(tci, grid_ctx, A, maxrank=mr)
        _check(tci, grid_ctx, A, maxrank=mr, leftorthogonal=False)
    _check(tci, grid_ctx, np.zeros((40, 30)))                # one pivot of 0, error 0 stop
    _check(tci, grid_ctx, A, abstol=0.5)


def test_grid_large_auto(tci, ctx):
    """A single 1 500 × 1 200 low-rank matrix: AUTO takes the cooperative path."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((1500, 90)) @ rng.standard_normal((90, 1200)) + 1e-11 * rng.standard_normal((1500, 1200))
    for lo in (True, False):
        print("auto large", lo, flush=True)
        _check(tci, ctx, A, reltol=1e-9, leftorthogonal=lo)


def test_grid_c4_bond(tci, ctx):
    """One bond of the L=40, d=16 workload: a 4 352 × 4 352 Pi with 256
    pivots (maxbonddim 256), bitwise against the oracle."""
    rng = np.random.default_rng(4)
    m = n = 4352
    A = rng.standard_normal((m, 300)) @ rng.standard_normal((300, n)) + 1e-12 * rng.standard_normal((m, n))
    A = np.asfortranarray(A)
    t0 = time.perf_counter()
    a = tci.rrlu(A, maxrank=256, reltol=1e-14, ctx=ctx)
    t1 = time.perf_counter()
    b = O.rrlu(A, maxrank=256, reltol=1e-14)
    t2 = time.perf_counter()
    print(f"c4 bond: device call {t1 - t0:.3f} s (host copies included), oracle {t2 - t1:.1f} s", flush=True)
    assert a.npivot == b.npivot == 256
    assert np.array_equal(a.rowpermutation, b.rowpermutation)
    assert np.array_equal(a.colpermutation, b.colpermutation)
    assert a.error == b.error
    assert np.array_equal(a.L, b.L) and np.array_equal(a.U, b.U)
