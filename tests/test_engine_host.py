"""The device engine's own code, compiled for the CPU, against the oracle.

tests/engine_host builds tci_engine.h (the code the gfx950 solve kernel
runs) with a one-thread team, and the register-tile rrLU with a 256-thread
host team shaped like the device team.  Every result here must equal the
oracle bit for bit.  No GPU.
"""
import numpy as np
import pytest

import eh
import oracle as O
from problems import oracle_opts, specs

import importlib.util, os, sys  # noqa: E401
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from conftest import load_pkg  # noqa: E402

_abi = load_pkg()._abi


def _solve_both(spec, threads=1):
    name, kind, params, ip, ld, piv, opts = spec
    h = _abi.ProblemHolder(ld, kind, params, ip, piv, **opts)
    e = eh.EngineSolve(h, threads)
    try:
        t, r, er = O.crossinterpolate2(O.Functor(kind, params, ip), ld, piv, **oracle_opts(opts))
        ost = 0
    except O.OracleError as ex:
        t, ost = None, ex.code
    return e, t, ost


# the 256-thread emulation is slow (std::barrier per reduction): a subset
_MT = {"lorentz5_pivots", "quantics_exp_small_maxbond", "random_tt", "diag_table"}
_CASES = [pytest.param(s, 1, id=s[0] + "-1") for s in specs()] + \
         [pytest.param(s, 256, id=s[0] + "-256") for s in specs() if s[0] in _MT]


_CASES_LDS = [pytest.param(s, 1, id=s[0] + "-lds") for s in specs()] + \
             [pytest.param(s, 256, id=s[0] + "-256-lds") for s in specs()
              if s[0] in {"lorentz5_pivots", "random_tt"}]


@pytest.fixture
def host_lds(monkeypatch):
    monkeypatch.setenv("EH_LDS", "1")


@pytest.mark.parametrize("spec,threads", _CASES_LDS)
def test_engine_with_onchip_areas_equals_oracle(spec, threads, host_lds):
    """The engine with the on-chip areas the device has (Pi area, index
    tables, pivot sets held apart from the HBM arena) equals the oracle."""
    _compare(spec, threads)


@pytest.mark.parametrize("spec,threads", _CASES)
def test_engine_equals_oracle(spec, threads):
    _compare(spec, threads)


def _compare(spec, threads):
    """The engine (on a 1-thread team, or on 256 threads shaped like a 4-wave
    device team) equals the oracle bit for bit."""
    e, t, ost = _solve_both(spec, threads)
    assert e.status == ost
    if ost:
        return
    r, er, ng = e.history()
    assert list(r) == list(t.ranks)
    assert np.array_equal(er, t.errors)
    assert list(ng) == list(t.nglobalpivots)
    assert e.maxsample == t.maxsamplevalue
    L = len(spec[4])
    for b in range(L):
        assert e.set(0, b) == t.Iset(b)
        assert e.set(1, b) == t.Jset(b)
        assert np.array_equal(e.sitetensor(b), t.sitetensor(b))
    assert np.array_equal(e.pivoterrors(), t.pivoterrors)
    assert np.array_equal(e.bonderrors(), t.bonderrors)


def test_engine_errors():
    h = _abi.ProblemHolder([2] * 5, 1, [0.0], tolerance=1e-8)          # f == 0 at the pivot
    assert eh.EngineSolve(h).status == 4
    h = _abi.ProblemHolder([2] * 5, 1, [1.0], tolerance=0.0)           # tolerance 0, unlimited χ
    assert eh.EngineSolve(h).status == 3
    h = _abi.ProblemHolder([10] * 8, 1, [1.0], tolerance=1e-12, chi_cap=4, maxiter=3)
    assert eh.EngineSolve(h).status == 10                               # capacity → loud error


def _rrlu_cases(seed, maxside):
    rng = np.random.default_rng(seed)
    for t in range(6):
        m, n = (int(x) for x in rng.integers(1, maxside + 1, 2))
        kind = t % 3
        if kind == 0:
            A = rng.standard_normal((m, n))
        elif kind == 1:
            r = int(rng.integers(1, 12))
            A = rng.standard_normal((m, r)) @ rng.standard_normal((r, n)) + 1e-12 * rng.standard_normal((m, n))
        else:
            A = rng.integers(-1, 2, (m, n)).astype(float)
        yield A, bool(rng.integers(0, 2)), [0.0, 1e-10, 1e-3][t % 3]


@pytest.mark.parametrize("seed", range(3))
def test_tile_and_stream_rrlu_equal_oracle(seed):
    for A, lo, rt in _rrlu_cases(seed, 160):
        b = O.rrlu(A, reltol=rt, leftorthogonal=lo) if rt > 0 or True else None
        for mode in (0, 1, 2):
            a = eh.rrlu(mode, A, reltol=rt, leftorthogonal=lo)
            if a["p"] == -100:
                continue
            assert a["p"] == b.npivot, (mode, A.shape)
            assert np.array_equal(a["rowperm"], b.rowpermutation)
            assert np.array_equal(a["colperm"], b.colpermutation)
            assert a["error"] == b.error or (np.isnan(a["error"]) and np.isnan(b.error))


@pytest.mark.parametrize("seed", [1])
def test_tile_rrlu_multithreaded_equals_oracle(seed):
    """256 host threads shaped like a 4-wave device team (2×32 doubles per lane, 128×128)."""
    for A, lo, rt in _rrlu_cases(100 + seed, 128):
        b = O.rrlu(A, reltol=rt, leftorthogonal=lo)
        for mode in (0, 1):
            a = eh.rrlu_mt(mode, A, reltol=rt, leftorthogonal=lo)
            assert a["p"] == b.npivot, (mode, A.shape)
            assert np.array_equal(a["rowperm"], b.rowpermutation)
            assert np.array_equal(a["colperm"], b.colpermutation)


def _nan_cases(seed):
    rng = np.random.default_rng(seed)
    for t in range(8):
        m, n = (int(x) for x in rng.integers(1, 41, 2))
        A = rng.integers(-2, 3, (m, n)).astype(float)
        if t % 4 == 0:
            A[rng.integers(0, m), :] = np.nan                   # a NaN row
        elif t % 4 == 1:
            A[rng.random((m, n)) < 0.1] = np.nan                 # scattered NaNs
        elif t % 4 == 2:
            A[:] = np.nan                                        # nothing to pick
        else:
            A[:, rng.integers(0, n)] = np.nan                    # a NaN column
        yield A, bool(t & 1)


def _oracle_status(A, lo):
    try:
        b = O.rrlu(A, leftorthogonal=lo)
        return 0, b
    except O.OracleError as e:
        return e.code, None


@pytest.mark.parametrize("seed", range(3))
def test_tile_rrlu_nan_verdicts_equal_oracle(seed):
    """NaN entries: the same error (lu.L / lu.U contains NaNs) or the same
    pivots as the reference restatement, on both host teams."""
    for A, lo in _nan_cases(300 + seed):
        st, b = _oracle_status(A, lo)
        for a in (eh.rrlu(0, A, leftorthogonal=lo), eh.rrlu_mt(0, A, leftorthogonal=lo),
                  eh.rrlu(2, A, leftorthogonal=lo)):
            if st:
                assert a["p"] == -st, (A.shape, lo, st, a["p"])
            else:
                assert a["p"] == b.npivot
                assert np.array_equal(a["rowperm"], b.rowpermutation)
                assert np.array_equal(a["colperm"], b.colpermutation)


@pytest.mark.parametrize("scale", [1e-150, 1e-158, 1e-162, 1e-170, 1e-200])
def test_tile_rrlu_tiny_magnitudes_equal_oracle(scale):
    """Entries whose squares are subnormal or underflow to zero: ties in abs2
    among different magnitudes must resolve exactly as the reference's
    column-major first maximum does."""
    rng = np.random.default_rng(int(-np.log10(scale)))
    for t in range(4):
        m, n = (int(x) for x in rng.integers(2, 41, 2))
        A = rng.standard_normal((m, n)) * scale
        if t % 2:
            A[rng.random((m, n)) < 0.3] = 0.0
        lo = bool(t & 1)
        st, b = _oracle_status(A, lo)
        for a in (eh.rrlu(0, A, leftorthogonal=lo), eh.rrlu_mt(0, A, leftorthogonal=lo)):
            if st:
                assert a["p"] == -st
                continue
            assert a["p"] == b.npivot, (scale, t)
            assert np.array_equal(a["rowperm"], b.rowpermutation)
            assert np.array_equal(a["colperm"], b.colpermutation)
            assert a["error"] == b.error or (np.isnan(a["error"]) and np.isnan(b.error))
