"""The register-tile rrLU (the solve engine's hot factorization, matrices up
to 128×128) through the standalone entry points, selected with
tci_ctx_set_rrlu_algo(TILE), against the oracle bit for bit: random,
low-rank and tie-heavy matrices in both orientations, NaN verdicts, tiny
magnitudes whose squares are subnormal, and the 80×118 low-rank matrix on
which an earlier tile design diverged.  The packed factors the batched
entry point leaves in A (rrlu!, matrixlu.jl:194-204), trailing block
included, equal the streaming kernel's."""
import numpy as np
import pytest

import oracle as O
from test_engine_host import _nan_cases, _oracle_status, _rrlu_cases

pytestmark = pytest.mark.gpu

AUTO, STREAM, TILE = 0, 1, 2


@pytest.fixture
def tile_ctx(tci, ctx):
    c = tci.Context(0)
    assert tci._abi.lib().tci_ctx_set_rrlu_algo(c.handle, TILE) == 0
    yield c
    c.close()


def _same(a, b):
    return a == b or (np.isnan(a) and np.isnan(b))


def _check(tci, ctx, A, **kw):
    st, b = _oracle_status_kw(A, **kw)
    try:
        a = tci.rrlu(A, ctx=ctx, **kw)
        ast = 0
    except tci.TCIError as e:
        a, ast = None, e.status
    assert ast == st, (A.shape, kw)
    if st:
        return
    assert a.npivot == b.npivot, (A.shape, kw)
    assert np.array_equal(a.rowpermutation, b.rowpermutation)
    assert np.array_equal(a.colpermutation, b.colpermutation)
    assert _same(a.error, b.error)
    assert np.array_equal(a.L, b.L) and np.array_equal(a.U, b.U)


def _oracle_status_kw(A, **kw):
    try:
        return 0, O.rrlu(A, **kw)
    except O.OracleError as e:
        return e.code, None


@pytest.mark.parametrize("seed", range(4))
def test_tile_random_lowrank_ties(tci, tile_ctx, seed):
    for A, lo, rt in _rrlu_cases(seed, 128):
        print("tile case", seed, A.shape, lo, rt, flush=True)
        _check(tci, tile_ctx, A, reltol=rt, leftorthogonal=lo)
        _check(tci, tile_ctx, A, reltol=rt, leftorthogonal=not lo)


def test_tile_maxrank_and_abstol(tci, tile_ctx):
    rng = np.random.default_rng(9)
    A = rng.standard_normal((100, 77))
    for kw in (dict(maxrank=5), dict(maxrank=77), dict(abstol=0.5), dict(abstol=0.5, leftorthogonal=False),
               dict(maxrank=3, leftorthogonal=False)):
        _check(tci, tile_ctx, A, **kw)


@pytest.mark.parametrize("seed", range(3))
def test_tile_nan_verdicts(tci, tile_ctx, seed):
    for A, lo in _nan_cases(300 + seed):
        _check(tci, tile_ctx, A, leftorthogonal=lo)


@pytest.mark.parametrize("scale", [1e-150, 1e-158, 1e-162, 1e-170, 1e-200])
def test_tile_tiny_magnitudes(tci, tile_ctx, scale):
    rng = np.random.default_rng(int(-np.log10(scale)))
    for t in range(4):
        m, n = (int(x) for x in rng.integers(2, 41, 2))
        A = rng.standard_normal((m, n)) * scale
        if t % 2:
            A[rng.random((m, n)) < 0.3] = 0.0
        _check(tci, tile_ctx, A, leftorthogonal=bool(t & 1))


def test_tile_earlier_divergence_matrix(tci, tile_ctx):
    # seed 2, case 1 of the random rrLU parity test: 80×118, rank < 12 plus
    # 1e-12 noise, leftorthogonal=false — the matrix on which an earlier
    # register-tile design chose a different pivot 9 on the GPU
    rng = np.random.default_rng(2)
    cases = []
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
        cases.append((A, lo))
    A, lo = cases[1]
    assert A.shape == (80, 118) and lo is False
    _check(tci, tile_ctx, A, reltol=1e-10, leftorthogonal=False)
    _check(tci, tile_ctx, A, reltol=0.0, leftorthogonal=False)
    _check(tci, tile_ctx, A, reltol=1e-10, leftorthogonal=True)


def test_batched_packed_factors_tile_equals_stream(tci, ctx):
    """The batched entry point on uneven matrices: the tile kernel's packed
    A (L, U and the trailing block) equals the streaming kernel's bit for
    bit, and so do permutations, pivots and errors."""
    import torch
    rng = np.random.default_rng(5)
    mats = []
    for t in range(40):
        m, n = (int(x) for x in rng.integers(1, 129, 2))
        r = int(rng.integers(1, 20))
        A = rng.standard_normal((m, r)) @ rng.standard_normal((r, n)) + 1e-9 * rng.standard_normal((m, n))
        if t % 7 == 3:
            A = rng.integers(-1, 2, (m, n)).astype(float)
        mats.append(A)
    out = {}
    for algo in (STREAM, TILE):
        c = tci.Context(0)
        assert tci._abi.lib().tci_ctx_set_rrlu_algo(c.handle, algo) == 0
        buf = tci.upload_matrices(mats)
        for lo in (True, False):
            A = buf[0].clone()
            res = tci.rrlu_batched(mats, reltol=1e-12, leftorthogonal=lo, ctx=c, device_buffers=(A, buf[1], buf[2]))
            torch.cuda.synchronize()
            out[(algo, lo)] = (A.cpu().numpy(), res)
        c.close()
    for lo in (True, False):
        As, rs = out[(STREAM, lo)]
        At, rt = out[(TILE, lo)]
        assert np.array_equal(As, At)
        for a, b, M in zip(rs, rt, mats):
            assert a["npivot"] == b["npivot"] and a["status"] == b["status"]
            assert np.array_equal(a["rowpermutation"], b["rowpermutation"])
            assert np.array_equal(a["colpermutation"], b["colpermutation"])
            assert _same(a["error"], b["error"])
        # and the oracle's pivots
        for a, M in zip(rt, mats):
            o = O.rrlu(M, reltol=1e-12, leftorthogonal=lo)
            assert a["npivot"] == o.npivot
            assert np.array_equal(a["rowpermutation"], o.rowpermutation)


def test_tile_refuses_large_jobs(tci, tile_ctx):
    with pytest.raises(tci.ArgumentError):
        tci.rrlu(np.ones((129, 4)), ctx=tile_ctx)
