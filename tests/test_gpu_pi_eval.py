"""The standalone Pi-tensor evaluation entry point (tci_pi_eval_f64, the
device counterpart of _batchevaluate_dispatch, batcheval.jl:32-61) and its
kernels: the closed forms of test_batcheval.jl:13-35, the oracle's
batcheval bit for bit on several functors and M = 0, 1, 2, 3, the empty
sets, and the NaN-propagating maxsamplevalue update (updatemaxsample!,
tensorci2.jl:397-399)."""
import numpy as np
import pytest

import oracle as O
from problems import GSUM, LORENTZ, QGC, TABLE, WLORENTZ, gauss_sum, quantics_gc

pytestmark = pytest.mark.gpu


def test_closed_form_M1(tci, ctx):
    # test_batcheval.jl:14-24: f = sum, 100 left [1,1], 100 right [1,1], d = 2
    ld = [2] * 5
    left = [[1, 1]] * 100
    right = [[1, 1]] * 100
    out, mx = tci.batchevaluate(tci.SumFunctor(1.0), ld, left, right, 1)
    ref = np.array([[[sum(l + [c] + r) for r in right] for c in (1, 2)] for l in left], float)
    assert out.shape == (100, 2, 100)
    assert np.array_equal(out, ref)
    assert mx == 6.0


def test_closed_form_M2(tci, ctx):
    # test_batcheval.jl:26-35
    ld = [2] * 5
    left = [[1]] * 100
    right = [[1, 1]] * 100
    out, mx = tci.batchevaluate(tci.SumFunctor(1.0), ld, left, right, 2)
    ref = np.array([[[[sum(l + [c, cp] + r) for r in right] for cp in (1, 2)] for c in (1, 2)] for l in left],
                   float)
    assert out.shape == (100, 2, 2, 100)
    assert np.array_equal(out, ref)


def _oracle(kind, params, ip, ld, left, right):
    nl = len(left[0]) if len(left) else 0
    nr = len(right[0]) if len(right) else 0
    return O.batcheval(O.Functor(kind, params, ip), ld, np.array(left, np.int32).reshape(len(left), nl),
                       np.array(right, np.int32).reshape(len(right), nr))


_FUNCTORS = [
    ("lorentz", LORENTZ, [1.0], [], [10] * 6),
    ("weighted_lorentz", WLORENTZ, np.concatenate([[16.0], np.linspace(0.5, 2, 6)]), [], [16] * 6),
    ("gauss_sum", GSUM, gauss_sum(3, 6, 8, 1), [], [8] * 6),
    ("quantics_gc", QGC, quantics_gc(3, 3, F=2.0 ** 20), [], [2] * 6),
]


@pytest.mark.parametrize("fn", _FUNCTORS, ids=lambda f: f[0])
@pytest.mark.parametrize("M", [0, 1, 2, 3])
def test_equals_oracle(tci, ctx, fn, M):
    name, kind, params, ip, ld = fn
    rng = np.random.default_rng(M + 17)
    nl = 1 if M < 3 else 1
    nr = len(ld) - nl - M
    left = [[int(rng.integers(1, ld[t] + 1)) for t in range(nl)] for _ in range(7)]
    right = [[int(rng.integers(1, ld[nl + M + t] + 1)) for t in range(nr)] for _ in range(5)]
    out, mx = tci.batchevaluate(tci.DeviceFunctor(kind, np.asarray(params, float), np.asarray(ip, np.int32)),
                                ld, left, right, M)
    want = _oracle(kind, params, ip, ld, left, right)
    assert out.shape == want.shape
    assert np.array_equal(out, want)
    assert mx == np.max(np.abs(want))


def test_empty_sets_give_empty_tensors(tci, ctx):
    out, mx = tci.batchevaluate(tci.SumFunctor(1.0), [2] * 4, [], [[1, 1]], 1)
    assert out.size == 0 and mx == 0.0
    out, mx = tci.batchevaluate(tci.SumFunctor(1.0), [2] * 4, [[1]], [], 2)
    assert out.size == 0


def test_full_domain_no_left_no_right(tci, ctx):
    # M = L with empty multi-indices on both sides: the whole tensor
    ld = [3, 4, 2]
    out, mx = tci.batchevaluate(tci.SumFunctor(2.0), ld, [[]], [[]], 3)
    grid = np.add.outer(np.add.outer(np.arange(1, 4), np.arange(1, 5)), np.arange(1, 3)) * 2.0
    assert np.array_equal(out.reshape(3, 4, 2), grid)
    assert mx == np.max(grid)


def test_nan_propagates_into_maxsample(tci, ctx):
    import ctypes as C
    tab = np.arange(1.0, 10.0).reshape(3, 3)
    tab[2, 1] = np.nan
    f = tci.TableFunctor(tab)
    out, mx = tci.batchevaluate(f, [3, 3], [[1], [2], [3]], [[1], [2], [3]], 0)
    assert np.isnan(mx)
    assert np.isnan(out[2, 1]) and out[0, 0] == 1.0
    # the running value is folded in as |value| (updatemaxsample! takes abs)
    lib = tci._abi.lib()
    fs = f.struct()
    ld = np.array([3, 3], np.int32)
    lft = np.array([[1]], np.uint8)
    rgt = np.array([[1]], np.uint8)
    o = np.zeros(1)
    m = C.c_double(-50.0)
    assert lib.tci_pi_eval_f64(ctx.handle, C.byref(fs), 2, tci._abi.i32ptr(ld), tci._abi.u8ptr(lft), 1, 1,
                               tci._abi.u8ptr(rgt), 1, 1, tci._abi.dptr(o), C.byref(m)) == 0
    assert m.value == 50.0 and o[0] == 1.0


def test_argument_errors(tci, ctx):
    with pytest.raises(tci.ArgumentError):
        tci.batchevaluate(tci.SumFunctor(1.0), [2] * 4, [[3]], [[1, 1]], 1)      # index out of range
    with pytest.raises(tci.ArgumentError):
        tci.batchevaluate(tci.SumFunctor(1.0), [2] * 4, [[1]], [[1, 1]], 2)      # wrong M
    with pytest.raises(tci.ArgumentError):
        # a weighted Lorentzian with a too-short parameter block on one site
        tci.batchevaluate(tci.DeviceFunctor(WLORENTZ, np.array([16.0])), [4], [[]], [[]], 1)
