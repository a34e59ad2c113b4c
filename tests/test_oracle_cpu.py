"""The oracle pinned against the reference's own known answers.

Every check here is one of: a fixture + expected value copied from the
reference's tests (cited file:line), a closed form, brute force, or the
independent numpy restatement agreeing bit for bit.  No GPU.
"""
import json
import os

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st
from hypothesis.extra import numpy as hnp

import oracle as O
from rrlu_np import rrlu_np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLD, "reference_fixtures.json")) as _fh:
    FX = json.load(_fh)


def A_(name):
    return np.array(FX[name]["A"])


def left_right(lu):
    """left(lu)·right(lu) with the row/column permutations undone."""
    Lp = np.empty_like(lu.L)
    Lp[lu.rowpermutation - 1, :] = lu.L
    Up = np.empty_like(lu.U)
    Up[:, lu.colpermutation - 1] = lu.U
    return Lp, Up


# ---------------------------------------------------------------- argmax
def test_argmax_first_max_column_major():
    """submatrixargmax(abs2, A, k): first maximum in column-major order
    (test_matrixlu.jl:21-28 pins the column-major convention)."""
    A = A_("argmax_10x8")
    for k in range(8):
        r, c = O.submatrixargmax(A, k)
        blk = (A[k:, k:] ** 2).ravel(order="F")
        q = int(np.argmax(blk))
        assert (r, c) == (k + q % (10 - k), k + q // (10 - k))
    assert O.submatrixargmax(A, 7) == (int(np.argmax(A[7:, 7:] ** 2) % 3) + 7, 7)


def test_argmax_ties_and_nan():
    A = np.array([[1.0, -2.0], [2.0, np.nan]])
    assert O.submatrixargmax(A, 0) == (1, 0)       # |2| twice: first in column-major order
    N = np.full((3, 3), np.nan)
    assert O.submatrixargmax(N, 1) == (1, 1)       # all NaN → (k, k)


# ------------------------------------------------------------------ rrLU
def test_rrlu_4x4_reconstructs():
    """test_matrixlu.jl:54-69."""
    A = A_("rrlu_4x4")
    lu = O.rrlu(A)
    L, U = lu.L, lu.U
    assert np.all(np.diag(L) == 1.0) and np.allclose(L, np.tril(L))
    assert np.allclose(U, np.triu(U))
    Lp, Up = left_right(lu)
    assert np.allclose(Lp @ Up, A, rtol=1e-12, atol=1e-14)


def test_rrlu_truncated():
    """test_matrixlu.jl:88-97: npivot == 1."""
    assert O.rrlu(A_("truncated_3x3")).npivot == 1


def test_rrlu_maxrank():
    """test_matrixlu.jl:99-121."""
    lu = O.rrlu(A_("approx_8x6"), maxrank=4)
    assert lu.npivot == 4 and lu.L.shape == (8, 4) and lu.U.shape == (4, 6)
    assert np.array_equal(lu.L, np.tril(lu.L)) and np.array_equal(lu.U, np.triu(lu.U))


def test_rrlu_reltol_lowrank():
    """test_matrixlu.jl:123-138 with seeded noise instead of Julia's rand."""
    A = A_("approx_8x6")
    rng = np.random.default_rng(1)
    A = np.hstack([A, A + 1e-3 * rng.random((8, 6))])
    lu = O.rrlu(A, reltol=1e-2)
    assert lu.npivot < 8
    Lp, Up = left_right(lu)
    assert np.max(np.abs(Lp @ Up - A)) < 1e-2


def test_rrlu_exact_rank3():
    """test_matrixlu.jl:141-165 / test_matrixluci.jl:48-74."""
    A = A_("lowrank_p") @ A_("lowrank_q")
    lu = O.rrlu(A)
    assert lu.npivot == 3
    Lp, Up = left_right(lu)
    assert np.allclose(Lp @ Up, A, rtol=1e-12, atol=1e-14)


def test_rrlu_pivoterrors_identity():
    """test_matrixlu.jl:167-175: pivoterrors == [1, 1, 0], lastpivoterror == 0."""
    lu = O.rrlu(A_("identity_2x2"))
    pe = list(np.abs(np.diag(lu.U[:lu.npivot, :lu.npivot]))) + [lu.error]
    assert pe == [1.0, 1.0, 0.0] and lu.error == 0.0


def test_rrlu_lastpivoterror():
    """test_matrixlu.jl:177-195."""
    A = A_("lastpivot_5x5")
    lu = O.rrlu(A, maxrank=2)
    assert lu.npivot + 1 == 3 and lu.error > 0
    assert O.rrlu(A, abstol=0.5).error < 0.5
    assert O.rrlu(A, abstol=0.0).error == 0.0


def test_rrlu_tiny_values():
    """test_matrixlu.jl:197-211."""
    A = A_("tiny_4x4")
    lu = O.rrlu(A, abstol=1e-3)
    assert lu.npivot == 1 and lu.error > 0
    Lp, Up = left_right(lu)
    assert np.max(np.abs(Lp @ Up - A)) < 1e-3


def test_rrlu_nan_raises():
    """lu.L contains NaNs (matrixlu.jl:164-166)."""
    A = np.array([[1.0, 2.0], [np.nan, 1.0]])
    with pytest.raises(O.OracleError) as e:
        O.rrlu(A)
    assert e.value.code in (1, 2)


def test_luci_closed_forms():
    """MatrixLUCI factors vs the closed forms of test_matrixluci.jl:23-37:
    colstimespivotinv = A[:,J]·A[I,J]⁻¹, pivotinvtimesrows = A[I,J]⁻¹·A[I,:]."""
    A = A_("luci_8x6")
    lu, left, right = O.luci_factors(A, maxrank=4)
    I = lu.rowpermutation[:4] - 1
    J = lu.colpermutation[:4] - 1
    P = A[np.ix_(I, J)]
    assert np.allclose(left, A[:, J] @ np.linalg.inv(P), rtol=1e-10, atol=1e-12)
    lu2, l2, r2 = O.luci_factors(A, maxrank=4, leftorthogonal=False)
    I2 = lu2.rowpermutation[:4] - 1
    J2 = lu2.colpermutation[:4] - 1
    P2 = A[np.ix_(I2, J2)]
    assert np.allclose(r2, np.linalg.inv(P2) @ A[I2, :], rtol=1e-10, atol=1e-12)
    assert np.allclose(left @ right, A[:, J] @ np.linalg.inv(P) @ A[I, :], rtol=1e-10, atol=1e-12)


def test_luci_exact_lowrank():
    A = A_("luci_p") @ A_("luci_q")
    lu, left, right = O.luci_factors(A)
    assert lu.npivot == 3
    assert np.allclose(left @ right, A, rtol=1e-12, atol=1e-14)


def _run_both(A, **kw):
    """Both restatements; an error is compared by its reference message."""
    try:
        a = O.rrlu(A, **kw)
    except O.OracleError as e:
        a = ("error", e.code)
    try:
        b = rrlu_np(A, **kw)
    except FloatingPointError as e:
        b = ("error", 1 if "lu.L" in str(e) else 2)
    if isinstance(a, tuple) or isinstance(b, tuple):
        assert a == b
        return None
    _same(a, b)
    return a


def _same(a, b):
    assert list(a.rowpermutation) == list(b["rowpermutation"])
    assert list(a.colpermutation) == list(b["colpermutation"])
    assert a.npivot == b["npivot"]
    assert np.array_equal(a.L, b["L"]) and np.array_equal(a.U, b["U"])
    assert a.error == b["error"] or (np.isnan(a.error) and np.isnan(b["error"]))


def test_golden_pivots_from_both_restatements():
    with open(os.path.join(GOLD, "rrlu_pivots.json")) as fh:
        cases = json.load(fh)
    for c in cases:
        A = np.array(c["A"])
        a = O.rrlu(A, **c["kwargs"])
        assert list(a.rowpermutation) == c["rowpermutation"], c["name"]
        assert a.npivot == c["npivot"], c["name"]
        kw = {k: v for k, v in c["kwargs"].items()}
        _same(a, rrlu_np(A, **kw))


@settings(max_examples=150, deadline=None)
@given(hnp.arrays(np.float64, st.tuples(st.integers(1, 12), st.integers(1, 12)),
                  elements=st.sampled_from([0.0, 1.0, -1.0, 2.0, 0.5, -3.0, 1e-300, 7.25])),
       st.booleans(), st.sampled_from([0.0, 1e-10, 0.3]))
def test_c_and_numpy_restatements_agree(A, lo, rt):
    """Tie-heavy small matrices: the two restatements agree bit for bit."""
    _run_both(A, reltol=rt, leftorthogonal=lo)


@settings(max_examples=60, deadline=None)
@given(hnp.arrays(np.float64, st.tuples(st.integers(1, 30), st.integers(1, 30)),
                  elements=st.floats(-1e3, 1e3, allow_nan=False)), st.booleans())
def test_c_and_numpy_agree_random(A, lo):
    _run_both(A, leftorthogonal=lo)


# ---------------------------------------------------------- evaluation
def test_batcheval_closed_forms():
    """test_batcheval.jl:13-35: f = sum, M = 1 and M = 2."""
    f = O.Functor(5, [1.0])
    ld = [2, 2, 2, 2, 2]
    left = [[1, 1]] * 4 + [[2, 1]]
    right = [[1, 2]] * 3
    r = O.batcheval(f, ld, left, right)
    ref = np.array([[[sum(l) + c + sum(rr) for rr in right] for c in (1, 2)] for l in left])
    assert np.array_equal(r, ref)
    r2 = O.batcheval(f, ld, [[1], [2]], right)
    ref2 = np.array([[[[sum(l) + c + cp + sum(rr) for rr in right] for cp in (1, 2)] for c in (1, 2)]
                     for l in ([1], [2])])
    assert np.array_equal(r2, ref2)


def test_project_exp_cos_accuracy():
    """The functors' own exp/cos: within 2 ulp of numpy's over the ranges the
    workloads use (exp of [-700, 700], cos of x·2^20 with x ∈ [0, 1))."""
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-700, 700, 4000), rng.uniform(-1, 1, 2000), [0.0, -0.0, 1.0]])
    for x in xs:
        y, r = O.lib().orc_exp(float(x)), np.exp(x)
        assert abs(y - r) <= 2 * np.spacing(r), x
    for x in rng.uniform(0, 1, 3000) * 2.0 ** 20:
        y, r = O.lib().orc_cos(float(x)), np.cos(x)
        assert abs(y - r) <= 4e-16 * max(1.0, abs(x) * 2 ** -20), x
    assert np.isnan(O.lib().orc_exp(float("nan"))) and O.lib().orc_exp(1000.0) == float("inf")


# -------------------------------------------------------------------- TCI2
def test_tci_pivoterrors_diag():
    """test_tensorci2.jl:27-39: tci.pivoterrors == [1.0, 1e-5, 0.0]."""
    tab = np.diag([1.0, 1e-5, 0.0])
    t, _, _ = O.crossinterpolate2(O.Functor(6, tab.ravel(order="F"), [3, 3]), [3, 3], [[1, 1]], tolerance=1e-8)
    assert list(t.pivoterrors) == [1.0, 1e-5, 0.0]


def test_tci_trivial_mps():
    """test_tensorci2.jl:215-232."""
    t = O.TCI([2] * 5)
    t.init(O.Functor(5, [0.5]), [[1] * 5])
    assert t.linkdims() == [1] * 4 and t.rank() == 1


def test_tci_lorentz_updatepivots_and_globalpivot():
    """test_tensorci2.jl:247-273: maxbonddim=2 sweeps give linkdims 2; a global
    pivot plus the 1-site sweeps of makecanonical! give 3."""
    f = O.Functor(1, [1.0])
    t = O.TCI([10] * 5)
    t.init(f, [[1] * 5])
    assert t.linkdims() == [1] * 4
    for b in range(4):
        t.updatepivots(b, f, True, reltol=1e-8, maxbonddim=2)
    assert t.linkdims() == [2] * 4
    t.addglobalpivots([[2, 9, 10, 5, 7]])
    t.sweep1site(f, True, reltol=0.0, abstol=0.0, updatetensors=False)
    t.sweep1site(f, False, reltol=1e-12, updatetensors=False)
    t.sweep1site(f, True, reltol=1e-12, updatetensors=True)
    assert t.linkdims() == [3] * 4
    assert len(t.Iset(0)) == 1 and len(t.Jset(4)) == 1


@pytest.mark.parametrize("pivots", [None, [[1, 1, 1, 1, 1], [10, 8, 10, 4, 4], [5, 4, 8, 9, 3],
                                           [7, 7, 10, 5, 9], [7, 7, 10, 5, 9]]])
def test_tci_lorentz_tight_tolerance(pivots):
    """test_tensorci2.jl:297-339: pivoterror <= 2e-12 and value ≈ f on 3^5 points."""
    f = O.Functor(1, [1.0])
    t, ranks, errors = O.crossinterpolate2(f, [10] * 5, pivots, tolerance=1e-12, maxiter=200)
    assert np.max(t.bonderrors) <= 2e-12
    assert t.rank() <= 200
    import itertools
    for v in itertools.product(range(1, 4), repeat=5):
        assert abs(t.evaluate(list(v)) - 1.0 / (1 + sum(x * x for x in v))) < 1e-12


def test_tci_forward_rank_matches():
    """test_tensorci2.jl:275-295: rank with sweepstrategy=:forward equals the
    rank after repeated updatepivots! (full pivot search)."""
    f = O.Functor(1, [1.0])
    t = O.TCI([10] * 5)
    t.init(f, [[1] * 5])
    for b in range(4):
        t.updatepivots(b, f, True, reltol=1e-8, maxbonddim=2)
    t.addglobalpivots([[2, 9, 10, 5, 7]])
    t.sweep1site(f, True, reltol=0.0, abstol=0.0, updatetensors=False)
    t.sweep1site(f, False, reltol=1e-12, updatetensors=False)
    t.sweep1site(f, True, reltol=1e-12, updatetensors=True)
    for _ in range(4, 21):
        for b in range(4):
            t.updatepivots(b, f, True, reltol=1e-8)
    t2, _, _ = O.crossinterpolate2(f, [10] * 5, tolerance=1e-8, maxiter=8, sweepstrategy="forward")
    assert t.rank() == t2.rank()


def test_tci_random_tt_reconstruction():
    """test_tensorci2.jl:477-502: TCI of a random TT reproduces it."""
    rng = np.random.default_rng(5)
    bd = [1, 2, 3, 2, 1]
    ld = [2, 3, 3, 2]
    cores = [rng.random((bd[i], ld[i], bd[i + 1])) for i in range(4)]
    f = O.Functor(7, np.concatenate([c.ravel(order="F") for c in cores]), bd + ld)
    t, _, _ = O.crossinterpolate2(f, ld, tolerance=1e-10, maxbonddim=10)
    import itertools
    for v in itertools.product(*[range(1, d + 1) for d in ld]):
        ref = np.linalg.multi_dot([cores[i][:, v[i] - 1, :] for i in range(4)])[0, 0]
        assert abs(t.evaluate(list(v)) - ref) < 1e-10 * max(1, abs(ref))


def test_convergencecriterion_truth_table():
    """test_tensorci2.jl:504-554."""
    C = O.convergencecriterion
    assert C([1, 2], [1e-2, 1e-5], [0, 0], 1e-4, 4, 3) is False
    assert C([1, 2, 2, 2], [1e-2, 1e-5, 1e-5, 1e-5], [0, 0, 0, 0], 1e-4, 4, 3) is True
    assert C([1, 2, 2, 2], [1e-2, 1e-2, 1e-5, 1e-5], [0, 0, 0, 0], 1e-4, 4, 3) is False
    assert C([1, 2, 2, 2], [1e-2] * 4, [0, 0, 0, 0], 1e-4, 2, 3) is True
    assert C([1, 2, 2, 2], [1e-2] * 4, [0, 1, 1, 1], 1e-4, 2, 3) is True


def test_tci_argument_errors():
    """tolerance <= 0 with unlimited maxbonddim → ArgumentError (tensorci2.jl:750-754);
    f = 0 at the initial pivot → maxsamplevalue is zero! (tensorci2.jl:50)."""
    with pytest.raises(O.OracleError) as e:
        O.crossinterpolate2(O.Functor(1, [1.0]), [2] * 5, tolerance=0.0)
    assert e.value.code == 3
    with pytest.raises(O.OracleError) as e:
        O.crossinterpolate2(O.Functor(1, [0.0]), [2] * 5)
    assert e.value.code == 4


def test_tci_lorentz8_golden():
    with open(os.path.join(GOLD, "tci_c1.json")) as fh:
        g = json.load(fh)
    for ns in (0, 5):
        t, ranks, errors = O.crossinterpolate2(O.Functor(1, [1.0]), [10] * 8, tolerance=1e-8,
                                               nsearchglobalpivot=ns)
        gg = g[f"nsearch{ns}"]
        assert list(ranks) == gg["ranks"] and list(errors) == gg["errors"]
        assert [t.Iset(b) for b in range(8)] == gg["Iset"]


def test_tci_interpolation_property():
    """The TT reproduces f exactly (to rounding) on every pivot cross, and its
    normalized max error on random points stays at the tolerance scale."""
    f = O.Functor(1, [1.0])
    t, ranks, errors = O.crossinterpolate2(f, [10] * 8, tolerance=1e-8)
    rng = np.random.default_rng(11)
    pts = rng.integers(1, 11, (2000, 8))
    err = max(abs(t.evaluate(list(p)) - 1.0 / (1 + int(np.sum(p ** 2)))) for p in pts)
    assert err / t.maxsamplevalue < 1e-7
    for b in range(7):
        for I in t.Iset(b + 1)[:3]:
            for J in t.Jset(b + 1)[:3]:
                for s in range(1, 11):
                    v = I + [s] + J
                    assert abs(t.evaluate(v) - 1.0 / (1 + sum(x * x for x in v))) < 1e-12
