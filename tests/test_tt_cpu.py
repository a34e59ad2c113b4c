"""Tensor-train sweeps on the CPU: the oracle's compress! (tensortrain.jl:95-183)
and sweep1sitegetindices! (conversion.jl:73-156) pinned against the
mathematics, and the host-side index bookkeeping of TensorCI2(tt)
(conversion.jl:158-176) run over the oracle's sweep.  No GPU.

Pins of the oracle sweep:
* one bond against the independent numpy rrLU (oracle/rrlu_np.py): the
  factored site, pivots and pivot errors bit for bit (a permutation of L or
  U), the carried product to rounding;
* invariants every sweep keeps: the TT's values unchanged without
  truncation, [I; ·] at the pivot rows of a :CI left factor (unit lower
  triangular for :LU), the transposed statements backwards;
* the reference's compress! assertions (test_tensortrain.jl:37-47) and
  the NaN error of rrlu with a zero pivot (matrixlu.jl:164-166).
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from conftest import load_pkg
from problems import LORENTZ, TT, random_tt
from rrlu_np import rrlu_np

tci = load_pkg()


def _rtt(bd, ld, seed):
    return random_tt(bd, ld, seed)[2]


def _grid(ld):
    return np.array(np.meshgrid(*[np.arange(1, d + 1) for d in ld], indexing="ij")).reshape(len(ld), -1).T


def _tci_tt(f, ld, **kw):
    t, _, _ = O.crossinterpolate2(f, ld, **kw)
    return t.sitetensors()


# ------------------------------------------------------------ one bond
@pytest.mark.parametrize("method", ["LU", "CI"])
@pytest.mark.parametrize("forward", [True, False])
def test_single_bond_against_numpy_rrlu(method, forward):
    rng = np.random.default_rng(7)
    cores = [rng.standard_normal((1, 5, 4)), rng.standard_normal((4, 3, 1))]
    new, piv, pe = O.tt_sweep(cores, method, forward, maxrank=3, reltol=1e-14, abstol=0.0)
    s = 0 if forward else 1
    A = cores[s].reshape((5, 4), order="F") if forward else cores[1].reshape((4, 3), order="F")
    r = rrlu_np(A, maxrank=3, leftorthogonal=forward)
    p = r["npivot"]
    rp, cp, L, U = r["rowpermutation"], r["colpermutation"], r["L"], r["U"]
    assert list(piv[0][0]) == list(rp[:p]) and list(piv[0][1]) == list(cp[:p])
    d = np.abs(np.diag(U[:p, :p]) if forward else np.diag(L[:p, :p]))
    assert np.array_equal(pe[:p + 1], np.concatenate([d, [r["error"]]]))
    if method == "LU":
        lft = np.empty_like(L)
        lft[rp - 1, :] = L
        rgt = np.empty_like(U)
        rgt[:, cp - 1] = U
    else:
        # MatrixLUCI (matrixluci.jl:40-84) through numpy solves
        if forward:
            lft = np.zeros((A.shape[0], p))
            lft[rp - 1, :] = np.vstack([np.eye(p), np.linalg.solve(L[:p, :p].T, L[p:, :].T).T])
            rgt = np.zeros((p, A.shape[1]))
            rgt[:, cp - 1] = L[:p, :p] @ U
        else:
            lft = np.zeros((A.shape[0], p))
            lft[rp - 1, :] = L @ U[:, :p]
            rgt = np.zeros((p, A.shape[1]))
            rgt[:, cp - 1] = np.hstack([np.eye(p), np.linalg.solve(U[:, :p], U[:, p:])])
    if forward:
        got_site = new[0].reshape((5, p), order="F")
        nxt = rgt @ cores[1].reshape((4, 3), order="F")
        got_next = new[1].reshape((p, 3), order="F")
        site = lft
    else:
        got_site = new[1].reshape((p, 3), order="F")
        nxt = cores[0].reshape((5, 4), order="F") @ lft
        got_next = new[0].reshape((5, p), order="F")
        site = rgt
    if method == "LU":
        assert np.array_equal(got_site, site)      # a permutation of L or U: exact
    else:
        assert np.allclose(got_site, site, rtol=1e-13, atol=1e-13)
    assert np.allclose(got_next, nxt, rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------ invariants
@pytest.mark.parametrize("method", ["LU", "CI"])
def test_sweep_invariants(method):
    bd, ld = [1, 3, 7, 6, 4, 1], [3, 4, 2, 5, 3]
    cores = _rtt(bd, ld, 11)
    pts = _grid(ld)
    v0 = O.tt_evaluate(cores, pts)
    fw, piv, _ = O.tt_sweep(cores, method, True, reltol=0.0)
    assert np.allclose(O.tt_evaluate(fw, pts), v0, rtol=1e-12, atol=1e-12 * np.abs(v0).max())
    for b in range(len(ld) - 1):
        p = fw[b].shape[2]
        M = fw[b].reshape((-1, p), order="F")[np.array(piv[b][0]) - 1, :]
        if method == "CI":
            assert np.array_equal(M, np.eye(p))
        else:
            assert np.array_equal(np.triu(M), np.eye(p))
    bw, piv, _ = O.tt_sweep(fw, method, False, reltol=0.0)
    assert np.allclose(O.tt_evaluate(bw, pts), v0, rtol=1e-12, atol=1e-12 * np.abs(v0).max())
    for b in range(len(ld) - 1):
        p = bw[b + 1].shape[0]
        M = bw[b + 1].reshape((p, -1), order="F")[:, np.array(piv[b][1]) - 1]
        if method == "CI":
            assert np.array_equal(M, np.eye(p))
        else:
            assert np.array_equal(np.tril(M), np.eye(p))


# ------------------------------------------------------------ compress!
@pytest.mark.parametrize("method", ["LU", "CI"])
def test_compress_reference_assertions(method):
    """test_tensortrain.jl:37-47 on a real function of the same grid."""
    f = O.Functor(LORENTZ, [1.0])
    ld = [6] * 4
    tt = _tci_tt(f, ld, tolerance=1e-8)
    rank = max(c.shape[2] for c in tt[:-1])
    c5 = O.tt_compress(tt, method, maxbonddim=5)
    assert max(c.shape[2] for c in c5[:-1]) <= 5
    c1 = O.tt_compress(tt, method, tolerance=1.0)
    assert max(c.shape[2] for c in c1[:-1]) <= rank
    # and without truncation the values stay
    c0 = O.tt_compress(tt, method, tolerance=1e-14)
    pts = _grid(ld)
    v0 = O.tt_evaluate(tt, pts)
    assert np.max(np.abs(O.tt_evaluate(c0, pts) - v0)) <= 1e-12 * np.abs(v0).max()


@pytest.mark.parametrize("method", ["LU", "CI"])
def test_compress_truncates_a_redundant_train(method):
    # a rank-2 function written with bonds of 6: compress! finds rank 2
    rng = np.random.default_rng(3)
    ld = [4, 4, 4, 4]
    cores = _rtt([1, 2, 2, 2, 1], ld, 5)
    # t_b' = A_b^+ · t_b · A_{b+1} with random full-rank A: same values, bonds 6
    A = [None] + [rng.standard_normal((2, 6)) for _ in range(3)] + [None]
    big = []
    for b, c in enumerate(cores):
        t = c
        if b > 0:
            t = np.einsum("ij,jkl->ikl", np.linalg.pinv(A[b]), t)
        if b < len(cores) - 1:
            t = np.einsum("ijk,kl->ijl", t, A[b + 1])
        big.append(t)
    pts = _grid(ld)
    v0 = O.tt_evaluate(cores, pts)
    assert np.allclose(O.tt_evaluate(big, pts), v0, rtol=1e-10)
    c = O.tt_compress(big, method, tolerance=1e-10)
    assert [x.shape[2] for x in c[:-1]] == [2, 2, 2]
    assert np.allclose(O.tt_evaluate(c, pts), v0, rtol=1e-8, atol=1e-9 * np.abs(v0).max())


def test_compress_zero_pivot_is_the_nan_error():
    # identical columns of powers of two: the second pivot is exactly 0 and
    # 0/0 fills L (matrixlu.jl:164-166), for :LU and :CI alike
    col = np.array([1.0, 2.0, 0.5, 4.0])
    c0 = np.stack([col, col], 1).reshape((1, 4, 2), order="F")
    c1 = np.arange(1.0, 7.0).reshape((2, 3, 1), order="F")
    for m in ("LU", "CI"):
        with pytest.raises(O.OracleError) as e:
            O.tt_compress([c0, c1], m)
        assert e.value.code == 1


def test_compress_normalizeerror_false_is_absolute():
    f = O.Functor(LORENTZ, [1.0])
    ld = [6] * 4
    tt = _tci_tt(f, ld, tolerance=1e-8)
    v = O.tt_evaluate(tt, _grid(ld))
    scale = np.abs(v).max()
    a = O.tt_compress(tt, "LU", tolerance=1e-4 * scale, normalizeerror=False)
    r = O.tt_compress(tt, "LU", tolerance=1e-4, normalizeerror=True)
    # the same truncation in absolute and relative terms is close but the
    # relative one uses the largest pivot, not max|f|: both cut the rank
    assert max(x.shape[2] for x in a[:-1]) < max(x.shape[2] for x in tt[:-1])
    assert max(x.shape[2] for x in r[:-1]) < max(x.shape[2] for x in tt[:-1])


# ------------------------------------------------------------ TensorCI2(tt)
def _oracle_sweep(tt, method="CI", forward=True, maxrank=O.INT64_MAX, reltol=1e-14, abstol=0.0, ctx=None):
    new, piv, pe = O.tt_sweep(tt.sitetensors, method, forward, maxrank, reltol, abstol)
    tt.sitetensors = new
    return [list(map(int, p[0])) for p in piv], [list(map(int, p[1])) for p in piv], pe


def test_tensorci2_from_tt_bookkeeping(monkeypatch):
    """conversion.jl:158-176 (test_conversion.jl:75-99 with a real f): the
    host's index bookkeeping over the oracle's sweeps gives nested sets whose
    cross interpolation of f reproduces the train, rank and bonds kept."""
    monkeypatch.setattr(tci, "tt_sweep", _oracle_sweep)
    f = O.Functor(LORENTZ, [1.0])
    ld = [4] * 4
    t, _, _ = O.crossinterpolate2(f, ld, tolerance=1e-14, maxbonddim=5)
    tt = tci.TensorTrain(t.sitetensors())
    links = tt.linkdims()
    assert max(links) == 5
    pts = _grid(ld)
    vtt = O.tt_evaluate(tt.sitetensors, pts)
    tcib = tci.tensorci2_from_tt(tt, tolerance=1e-14)
    assert tcib.linkdims() == links
    assert tcib.rank() == 5
    L = len(ld)
    for b in range(L):
        assert all(len(i) == b for i in tcib.Iset[b]) and all(len(j) == L - 1 - b for j in tcib.Jset[b])
    for b in range(1, L):
        # nesting: every Iset[b+1] entry extends an Iset[b] entry
        assert all(i[:-1] in tcib.Iset[b - 1] for i in tcib.Iset[b])
        assert all(j[1:] in tcib.Jset[b] for j in tcib.Jset[b - 1])
    assert np.allclose(O.tt_evaluate(tcib.sitetensors, pts), vtt, rtol=0, atol=1e-13)
    # the cross interpolation of the train itself from these sets is the train
    flat = np.concatenate([c.ravel(order="F") for c in t.sitetensors()])
    g = O.Functor(TT, flat, [1] + links + [1] + ld)
    o = O.TCI(ld)
    for b in range(L):
        o.setIset(b, tcib.Iset[b])
        o.setJset(b, tcib.Jset[b])
    O.lib().orc_tci_set_maxsample(o.h, 1.0)
    assert O.lib().orc_fillsitetensors(o.h, C.byref(g.st)) == 0
    assert np.allclose([o.evaluate(list(x)) for x in pts], vtt, rtol=0, atol=1e-12)
    # and optimize! from them converges to f (test_conversion.jl:95-99)
    o2 = O.TCI(ld)
    for b in range(L):
        o2.setIset(b, tcib.Iset[b])
        o2.setJset(b, tcib.Jset[b])
    O.lib().orc_tci_set_maxsample(o2.h, tcib.maxsamplevalue)
    o2.optimize(f, tolerance=1e-14)
    fv = O.eval_points(f, pts.astype(np.int32))
    assert np.max(np.abs(np.array([o2.evaluate(list(x)) for x in pts]) - fv)) < 1e-13


def test_tensorci2_from_tt_needs_two_sites():
    with pytest.raises(tci.TCIError):
        tci.tensorci2_from_tt(tci.TensorTrain([np.ones((1, 3, 1))]))


def test_compress_method_errors():
    tt = tci.TensorTrain(_rtt([1, 2, 1], [2, 2], 0))
    with pytest.raises(tci.TCIError, match="Not implemented yet"):
        tci.compress(tt, "QR")


# ------------------------------------------------------------ estimatetrueerror
def _zone_case():
    f = O.Functor(LORENTZ, [1.0])
    ld = [6] * 5
    t, _, _ = O.crossinterpolate2(f, ld, tolerance=1e-3, maxbonddim=3)
    return f, ld, t.sitetensors()


def test_floatingzone_fixed_point():
    """_floatingzone (globalsearch.jl:43-95): on return the error is
    |f − TT| at the found pivot, and no single-site move from it finds a
    larger error (the sweep that ended the search moved nothing upward)."""
    f, ld, cores = _zone_case()
    rng = np.random.default_rng(9)
    starts = np.stack([rng.integers(1, d + 1, 12) for d in ld], 1)
    piv, errs = O.floatingzone(f, cores, starts)
    fmax = np.abs(O.eval_points(f, _grid(ld).astype(np.int32))).max()
    atol = 1e-13 * fmax
    for q in range(len(errs)):
        x = piv[q]
        e = abs(O.eval_points(f, x[None, :])[0] - O.tt_evaluate(cores, x[None, :])[0])
        assert abs(e - errs[q]) <= atol
        for p in range(len(ld)):
            nb = np.repeat(x[None, :], ld[p], 0)
            nb[:, p] = np.arange(1, ld[p] + 1)
            en = np.abs(O.eval_points(f, nb) - O.tt_evaluate(cores, nb))
            assert en.max() <= errs[q] + atol


def test_floatingzone_against_brute_force():
    f, ld, cores = _zone_case()
    grid = _grid(ld).astype(np.int32)
    allerr = np.abs(O.eval_points(f, grid) - O.tt_evaluate(cores, grid))
    gmax = allerr.max()
    fmax = np.abs(O.eval_points(f, grid)).max()
    # from anywhere the search stays below the true maximum; from the
    # maximizer it returns the maximum itself
    rng = np.random.default_rng(2)
    starts = np.stack([rng.integers(1, d + 1, 20) for d in ld], 1)
    _, errs = O.floatingzone(f, cores, starts)
    assert errs.max() <= gmax + 1e-13 * fmax
    _, e1 = O.floatingzone(f, cores, grid[np.argmax(allerr)][None, :])
    assert abs(e1[0] - gmax) <= 1e-13 * fmax


def test_floatingzone_sweep_limits():
    f, ld, cores = _zone_case()
    rng = np.random.default_rng(4)
    starts = np.stack([rng.integers(1, d + 1, 8) for d in ld], 1)
    _, e_all = O.floatingzone(f, cores, starts)
    p1, e_one = O.floatingzone(f, cores, starts, nsweeps=1)
    assert np.all(e_one <= e_all)
    # earlystoptol 0 stops after the first sweep whose error is positive
    p0, e0 = O.floatingzone(f, cores, starts, earlystoptol=0.0)
    assert np.array_equal(p0, p1) and np.array_equal(e0, e_one)


def test_estimatetrueerror_host_order(monkeypatch):
    """globalsearch.jl:37-40: descending error, repeats dropped, over the
    oracle's per-point search."""
    f, ld, cores = _zone_case()

    def orc(tt, g, initialpoints, nsweeps=O.INT64_MAX, earlystoptol=float("inf"), ctx=None):
        return O.floatingzone(f, tt.sitetensors, initialpoints, nsweeps, earlystoptol)

    monkeypatch.setattr(tci, "floatingzone", orc)
    g = tci.Lorentz(1.0)
    starts = np.stack([np.random.default_rng(8).integers(1, d + 1, 30) for d in ld], 1)
    starts = np.vstack([starts, starts[:5]])          # repeats collapse
    res = tci.estimatetrueerror(tci.TensorTrain(cores), g, initialpoints=starts)
    errs = [e for _, e in res]
    assert errs == sorted(errs, reverse=True)
    assert len({(tuple(p), e) for p, e in res}) == len(res)
    piv, e = O.floatingzone(f, cores, starts)
    assert len(res) == len({(tuple(map(int, piv[q])), float(e[q])) for q in range(len(e))})
    with pytest.raises(tci.TCIError, match="No search"):
        tci.estimatetrueerror(tci.TensorTrain(cores), g, nsearch=0)


def test_kronecker_entries_follow_the_reference_order():
    """kronecker(Iset, d) and kronecker(d, Jset) (tensorci2.jl:315-327) are
    Julia comprehensions over (set, local index) flattened column-major: the
    set index runs fastest on the left, the local index on the right.  The
    index sweep reads single entries of them by index arithmetic."""
    I = [[1], [2]]
    left = [[1, 1], [2, 1], [1, 2], [2, 2], [1, 3], [2, 3]]
    assert [tci._kron_left_at(I, 3, q) for q in range(6)] == left
    J = [[4, 5], [6, 7]]
    right = [[1, 4, 5], [2, 4, 5], [3, 4, 5], [1, 6, 7], [2, 6, 7], [3, 6, 7]]
    assert [tci._kron_right_at(3, J, q) for q in range(6)] == right
    assert tci._kron_left_at([[]], 4, 2) == [3]
    assert tci._kron_right_at(4, [[]], 3) == [4]
