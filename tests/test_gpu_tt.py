"""Tensor-train sweeps on the device against the oracle, bit for bit:
compress! (tci_tt_compress_f64, tensortrain.jl:95-183) for :LU and :CI with
every truncation mode, one sweep of sweep1sitegetindices!
(tci_tt_sweep_f64, conversion.jl:73-156) with its pivots and pivot errors,
the products (tci_gemm_f64), and TensorCI2(tt) (conversion.jl:158-176)
as test_conversion.jl:75-99 checks it, with a real f.  The sweeps are
checked both in the fused one-workgroup kernel and driven bond by bond from
the host."""
import numpy as np
import pytest

import oracle as O
from problems import LORENTZ, random_tt

pytestmark = pytest.mark.gpu


def _grid(ld):
    return np.array(np.meshgrid(*[np.arange(1, d + 1) for d in ld], indexing="ij")).reshape(len(ld), -1).T


def _same(dev, orc):
    assert [t.shape for t in dev] == [t.shape for t in orc]
    for b, (x, y) in enumerate(zip(dev, orc)):
        assert np.array_equal(x, y), b


def _trains():
    f = O.Functor(LORENTZ, [1.0])
    t8, _, _ = O.crossinterpolate2(f, [10] * 8, tolerance=1e-10)
    t4, _, _ = O.crossinterpolate2(f, [6] * 4, tolerance=1e-8)
    rng = np.random.default_rng(5)
    big_bd = [1, 9, 60, 150, 150, 60, 9, 1]
    return {
        "lorentz8": t8.sitetensors(),
        "lorentz4": t4.sitetensors(),
        "random": random_tt([1, 3, 7, 6, 4, 1], [3, 4, 2, 5, 3], 11)[2],
        "random_big": [rng.standard_normal((big_bd[b], 9 if b < 2 else 4, big_bd[b + 1])) for b in range(7)],
    }


_TRAINS = _trains()


@pytest.fixture(params=["fused", "host"])
def route(request, monkeypatch):
    """The TT sweeps run either in one fused workgroup or driven from the
    host bond by bond (TCI_TT_FUSED); both must give the oracle's bits."""
    monkeypatch.setenv("TCI_TT_FUSED", "1" if request.param == "fused" else "0")
    return request.param


def test_gemm_equals_oracle_order(tci, ctx):
    rng = np.random.default_rng(2)
    for M, N, K in ((1, 1, 1), (7, 5, 3), (130, 70, 257), (3, 200, 1)):
        A = np.asfortranarray(rng.standard_normal((M, K)))
        B = np.asfortranarray(rng.standard_normal((K, N)))
        out = np.zeros((M, N), order="F")
        tci._check(tci._abi.lib().tci_gemm_f64(ctx.handle, M, N, K, tci._abi.dptr(A), M, tci._abi.dptr(B), K,
                                               tci._abi.dptr(out), M))
        ref = np.zeros((M, N))
        for k in range(K):   # ascending k, a rounding per product and per sum
            ref = ref + np.multiply.outer(A[:, k], B[k, :])
        assert np.array_equal(out, ref)


@pytest.mark.parametrize("name", sorted(_TRAINS))
@pytest.mark.parametrize("method", ["LU", "CI"])
def test_compress_equals_oracle(tci, ctx, name, method, route):
    cores = _TRAINS[name]
    for kw in (dict(tolerance=1e-12), dict(maxbonddim=5), dict(tolerance=1.0),
               dict(tolerance=1e-6, normalizeerror=False), dict(tolerance=1e-9, maxbonddim=12)):
        tt = tci.TensorTrain(cores)
        tci.compress(tt, method, ctx=ctx, **kw)
        _same(tt.sitetensors, O.tt_compress(cores, method, **kw))
    # the reference's assertions (test_tensortrain.jl:37-47)
    tt = tci.TensorTrain(cores)
    tci.compress(tt, method, maxbonddim=5, ctx=ctx)
    assert tt.rank() <= 5


@pytest.mark.parametrize("name", sorted(_TRAINS))
@pytest.mark.parametrize("method", ["LU", "CI"])
@pytest.mark.parametrize("forward", [True, False])
def test_sweep_equals_oracle(tci, ctx, name, method, forward, route):
    cores = _TRAINS[name]
    for kw in (dict(maxrank=O.INT64_MAX, reltol=1e-14, abstol=0.0), dict(maxrank=4, reltol=1e-14, abstol=0.0),
               dict(maxrank=O.INT64_MAX, reltol=1e-14, abstol=1e-7)):
        tt = tci.TensorTrain(cores)
        rows, cols, pe = tci.tt_sweep(tt, method, forward, ctx=ctx, **kw)
        new, piv, ope = O.tt_sweep(cores, method, forward, **kw)
        _same(tt.sitetensors, new)
        assert rows == [list(map(int, p[0])) for p in piv]
        assert cols == [list(map(int, p[1])) for p in piv]
        assert np.array_equal(pe, ope)


def test_zero_pivot_error_matches(tci, ctx, route):
    col = np.array([1.0, 2.0, 0.5, 4.0])
    c0 = np.stack([col, col], 1).reshape((1, 4, 2), order="F")
    c1 = np.arange(1.0, 7.0).reshape((2, 3, 1), order="F")
    for m in ("LU", "CI"):
        with pytest.raises(tci.TCIError) as e:
            tci.compress(tci.TensorTrain([c0, c1]), m, ctx=ctx)
        assert e.value.status == 1
        with pytest.raises(O.OracleError) as eo:
            O.tt_compress([c0, c1], m)
        assert eo.value.code == 1


def test_compress_svd_and_bad_arguments(tci, ctx):
    tt = tci.TensorTrain(_TRAINS["random"])
    with pytest.raises(tci.ArgumentError, match="SVD"):
        tci.compress(tt, "SVD", ctx=ctx)
    with pytest.raises(tci.ArgumentError):
        tci.compress(tt, "LU", maxbonddim=0, ctx=ctx)
    assert [t.shape for t in tt.sitetensors] == [t.shape for t in _TRAINS["random"]]


def test_tensorci2_from_tt(tci, ctx):
    """test_conversion.jl:75-99 with f(v) = 1/(Σv²+1)."""
    f = tci.Lorentz(1.0)
    ld = [4] * 4
    t, _, _ = tci.crossinterpolate2(float, f, ld, tolerance=1e-14, maxbonddim=5, ctx=ctx)
    tt = tci.TensorTrain(t)
    assert tt.rank() == 5 and tt.linkdims() == t.linkdims()
    assert tt.sitedims() == [[4]] * 4
    pts = _grid(ld)
    vt = t.evaluate(pts, ctx)
    tcib = tci.tensorci2_from_tt(tt, tolerance=1e-14, ctx=ctx)
    assert tcib.rank() == 5 and tcib.linkdims() == tt.linkdims()
    assert np.max(np.abs(tt.evaluate(pts, ctx) - vt)) < 1e-13
    assert np.max(np.abs(tcib.evaluate(pts, ctx) - vt)) < 1e-13
    tci.optimize(tcib, f, tolerance=1e-14, ctx=ctx)
    fv = O.eval_points(O.Functor(LORENTZ, [1.0]), pts.astype(np.int32))
    assert np.max(np.abs(tcib.evaluate(pts, ctx) - fv)) < 1e-13


def test_tensorci2_from_tt_equals_oracle_sweeps(tci, ctx, monkeypatch):
    """The device conversion equals the same bookkeeping over the oracle's
    sweeps (sets, site tensors and pivot errors bit for bit)."""
    cores = _TRAINS["lorentz8"]
    a = tci.tensorci2_from_tt(tci.TensorTrain(cores), tolerance=1e-12, maxiter=5, ctx=ctx)

    def orc(tt, method="CI", forward=True, maxrank=O.INT64_MAX, reltol=1e-14, abstol=0.0, ctx=None):
        new, piv, pe = O.tt_sweep(tt.sitetensors, method, forward, maxrank, reltol, abstol)
        tt.sitetensors = new
        return [list(map(int, p[0])) for p in piv], [list(map(int, p[1])) for p in piv], pe

    monkeypatch.setattr(tci, "tt_sweep", orc)
    b = tci.tensorci2_from_tt(tci.TensorTrain(cores), tolerance=1e-12, maxiter=5)
    assert a.Iset == b.Iset and a.Jset == b.Jset
    _same(a.sitetensors, b.sitetensors)
    assert np.array_equal(a.pivoterrors, b.pivoterrors)
    assert a.maxsamplevalue == b.maxsamplevalue


# ------------------------------------------------------------ estimatetrueerror
def _zone_cases(tci):
    f = O.Functor(LORENTZ, [1.0])
    t5, _, _ = O.crossinterpolate2(f, [6] * 5, tolerance=1e-3, maxbonddim=3)
    t8, _, _ = O.crossinterpolate2(f, [10] * 8, tolerance=1e-6, maxbonddim=6)
    rt = random_tt([1, 3, 4, 3, 1], [3, 5, 4, 2], 21)
    # a TT functor of one train against another train
    other = random_tt([1, 2, 5, 2, 1], [3, 5, 4, 2], 22)[2]
    g_tt = O.Functor(7, rt[0], rt[1])
    d_tt = tci.TTFunctor(rt[2])
    return [("lorentz5", f, tci.Lorentz(1.0), t5.sitetensors(), [6] * 5),
            ("lorentz8", f, tci.Lorentz(1.0), t8.sitetensors(), [10] * 8),
            ("tt_vs_tt", g_tt, d_tt, other, [3, 5, 4, 2])]


def test_floatingzone_equals_oracle(tci, ctx):
    for name, of, df, cores, ld in _zone_cases(tci):
        rng = np.random.default_rng(len(name))
        starts = np.stack([rng.integers(1, d + 1, 300) for d in ld], 1)
        for kw in (dict(), dict(nsweeps=1), dict(earlystoptol=1e-4)):
            piv, errs = tci.floatingzone(tci.TensorTrain(cores), df, starts, ctx=ctx, **kw)
            opiv, oerrs = O.floatingzone(of, cores, starts, **kw)
            assert np.array_equal(piv, opiv), name
            assert np.array_equal(errs, oerrs), name


def test_estimatetrueerror_device(tci, ctx):
    name, of, df, cores, ld = _zone_cases(tci)[0]
    res = tci.estimatetrueerror(tci.TensorTrain(cores), df, nsearch=100, seed=3, ctx=ctx)
    errs = [e for _, e in res]
    assert errs == sorted(errs, reverse=True) and len(res) >= 1
    grid = _grid(ld).astype(np.int32)
    allerr = np.abs(O.eval_points(of, grid) - O.tt_evaluate(cores, grid))
    assert errs[0] <= allerr.max() * (1 + 1e-12)
    # the reference's assertion shape: the estimate is a real error of the train
    p0 = np.array(res[0][0])[None, :]
    assert abs(abs(O.eval_points(of, p0)[0] - O.tt_evaluate(cores, p0)[0]) - errs[0]) <= 1e-13


def test_floatingzone_nan_stops(tci, ctx):
    # a table with a NaN: the search that meets it ends with NaN, on both
    tab = np.arange(1.0, 13.0).reshape(3, 4)
    tab[2, 1] = np.nan
    of = O.Functor(6, tab.ravel(order="F"), [3, 4])
    df = tci.TableFunctor(tab)
    cores = [np.ones((1, 3, 1)), np.ones((1, 4, 1))]
    starts = np.array([[1, 1], [3, 3], [2, 4]])
    piv, errs = tci.floatingzone(tci.TensorTrain(cores), df, starts, ctx=ctx)
    opiv, oerrs = O.floatingzone(of, cores, starts)
    assert np.array_equal(piv, opiv)
    assert np.array_equal(errs, oerrs, equal_nan=True)
    assert np.isnan(errs).any()
