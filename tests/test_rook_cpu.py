"""Rook pivot search (arrlu, matrixlu.jl:227-293, on the lazily evaluated Pi
of updatepivots!, tensorci2.jl:476-503, 552-595) on the CPU.

The oracle's arrlu is pinned by what the mathematics fixes: on matrices of
exact rank r it finds r pivots whose cross A[:,J]·A[I,J]⁻¹·A[I,:] is A; its
L·U equals A with the rows and columns permuted; its permutations are
permutations; the pivot rows and columns it returns are the full-pivot
choice of its last submatrix.  The reference's own rook assertions
(test_tensorci2.jl:55-100, 247-340) hold for the oracle's TCI.  Then the
device engine's code, built for the CPU, equals the oracle bit for bit with
pivotsearch = :rook.  The randomsubset draws come from the project PRNG,
so no bitwise comparison with the reference's Julia RNG is possible
(parity with the reference: unpinned beyond these properties)."""
import numpy as np
import pytest

import eh
import oracle as O
from conftest import load_pkg
from problems import LORENTZ, QEXP, oracle_opts, rook_specs
from rrlu_np import rrlu_np

_abi = load_pkg()._abi


@pytest.mark.parametrize("seed", range(4))
def test_oracle_arrlu_exact_rank(seed):
    rng = np.random.default_rng(100 + seed)
    for t in range(10):
        m, n = (int(x) for x in rng.integers(4, 70, 2))
        r = int(rng.integers(1, min(m, n, 8) + 1))
        A = rng.standard_normal((m, r)) @ rng.standard_normal((r, n))
        for lo in (True, False):
            I0 = [int(rng.integers(1, m + 1))]
            J0 = [int(rng.integers(1, n + 1))]
            res = O.arrlu(A, I0=I0, J0=J0, leftorthogonal=lo, seed=seed * 10 + t)
            p = res.npivot
            assert p == r, (m, n, r, p)
            assert sorted(res.rowpermutation) == list(range(1, m + 1))
            assert sorted(res.colpermutation) == list(range(1, n + 1))
            I, J = res.rowpermutation[:p] - 1, res.colpermutation[:p] - 1
            scale = np.abs(A).max()
            cross = A[:, J] @ np.linalg.solve(A[np.ix_(I, J)], A[I, :])
            assert np.abs(A - cross).max() <= 1e-11 * scale
            Ap = A[np.ix_(res.rowpermutation - 1, res.colpermutation - 1)]
            assert np.abs(Ap - res.L @ res.U).max() <= 1e-11 * scale
            d = np.diag(res.L if not lo else res.U)[:p]
            assert np.all(np.diag(res.U if not lo else res.L)[:p] == 1.0)
            assert np.all(d != 0)


def test_oracle_arrlu_last_submatrix_is_full_pivoting():
    """With numrookiter = 1, leftorthogonal and J0 at least half the
    columns, the random push adds every other column, so arrlu factors all
    of A once (columns in a random order): its row pivots, L, pivot values
    and error are those of the full-pivot rrLU of A, bit for bit (each
    entry's updates do not depend on the column order)."""
    rng = np.random.default_rng(7)
    A = rng.standard_normal((30, 25))
    for seed in range(3):
        res = O.arrlu(A, I0=[], J0=list(range(1, 14)), numrookiter=1, seed=seed, maxrank=4)
        ref = rrlu_np(A, maxrank=4)
        assert res.npivot == ref["npivot"] == 4
        assert np.array_equal(res.rowpermutation, ref["rowpermutation"])
        assert np.array_equal(res.L, ref["L"])
        assert res.error == ref["error"]
        assert np.array_equal(np.diag(res.U), np.diag(ref["U"][:, :4]))


def test_oracle_rook_reference_assertions():
    """test_tensorci2.jl:247-340 with pivotsearch = :rook (Float64)."""
    f = O.Functor(LORENTZ, [1.0])
    ld = [10] * 5
    t3, _, _ = O.crossinterpolate2(f, ld, tolerance=1e-12, maxiter=200, pivotsearch="rook")
    assert max(t3.bonderrors) <= 2e-12
    assert t3.rank() <= 200
    pts = np.array(np.meshgrid(*[np.arange(1, 4)] * 5, indexing="ij")).reshape(5, -1).T.astype(np.int32)
    v = np.array([t3.evaluate(list(x)) for x in pts])
    fv = O.eval_points(f, pts)
    assert np.allclose(v, fv, rtol=1e-10, atol=1e-12)
    piv = [[1, 1, 1, 1, 1], [10, 8, 10, 4, 4], [5, 4, 8, 9, 3], [7, 7, 10, 5, 9], [7, 7, 10, 5, 9]]
    t4, _, _ = O.crossinterpolate2(f, ld, piv, tolerance=1e-12, maxiter=200, pivotsearch="rook")
    assert max(t4.bonderrors) <= 2e-12


@pytest.mark.parametrize("strictlynested,nsearch", [(False, 0), (True, 0), (False, 10)])
def test_oracle_rook_trivial_mps(strictlynested, nsearch):
    """test_tensorci2.jl:55-100: exp(-x) on 8 quantics sites, maxbonddim 1,
    pivotsearch :rook — every bond stays at 1."""
    f = O.Functor(QEXP, [1.0, -1.0, 0.0, 0.0])
    t, _, _ = O.crossinterpolate2(f, [2] * 8, [[1] * 8, [1] + [2] * 7], tolerance=1e-4, maxbonddim=1, maxiter=2,
                                  normalizeerror=False, nsearchglobalpivot=nsearch, strictlynested=strictlynested,
                                  pivotsearch="rook")
    assert all(x == 1 for x in t.linkdims())


_CASES = [pytest.param(s, 1, id=s[0]) for s in rook_specs()] + \
         [pytest.param(s, 256, id=s[0] + "-256") for s in rook_specs() if s[0] in {"lorentz5_pivots_rook",
                                                                                    "random_tt_rook"}]


@pytest.mark.parametrize("spec,threads", _CASES)
def test_engine_rook_equals_oracle(spec, threads):
    name, kind, params, ip, ld, piv, opts = spec
    h = _abi.ProblemHolder(ld, kind, params, ip, piv, **opts)
    e = eh.EngineSolve(h, threads)
    try:
        t, _, _ = O.crossinterpolate2(O.Functor(kind, params, ip), ld, piv, **oracle_opts(opts))
        ost = 0
    except O.OracleError as ex:
        t, ost = None, ex.code
    assert e.status == ost
    if ost:
        return
    r, er, ng = e.history()
    assert list(r) == list(t.ranks)
    assert np.array_equal(er, t.errors)
    assert list(ng) == list(t.nglobalpivots)
    assert e.maxsample == t.maxsamplevalue
    for b in range(len(ld)):
        assert e.set(0, b) == t.Iset(b)
        assert e.set(1, b) == t.Jset(b)
        assert np.array_equal(e.sitetensor(b), t.sitetensor(b))
    assert np.array_equal(e.pivoterrors(), t.pivoterrors)
    assert np.array_equal(e.bonderrors(), t.bonderrors)
