"""Pins of two oracle functions that the reference's own fixtures do not
cover: the site tensors of fillsitetensors! (T_b = Pi1·P⁻¹,
tensorci2.jl:367-394) and the selection rule of DefaultGlobalPivotFinder
(globalpivotfinder.jl:143-195), including the project PRNG that replaces
Julia's default RNG.  No GPU."""
import numpy as np
import pytest

import oracle as O
from problems import c4_spec, specs

MASK = (1 << 64) - 1


# ---------------------------------------------------------------- PRNG
def _splitmix64(x):
    """SplitMix64 (Steele, Lea, Flood 2014): returns (new state, output)."""
    x = (x + 0x9E3779B97F4A7C15) & MASK
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return x, z ^ (z >> 31)


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & MASK


def _xoshiro_next(s):
    """xoshiro256** (Blackman, Vigna 2018), state list s of four words."""
    r = (_rotl((s[1] * 5) & MASK, 7) * 9) & MASK
    t = (s[1] << 17) & MASK
    s[2] ^= s[0]
    s[3] ^= s[1]
    s[1] ^= s[2]
    s[0] ^= s[3]
    s[2] ^= t
    s[3] = _rotl(s[3], 45)
    return r


def _seed(seed):
    x, s = seed & MASK, []
    for _ in range(4):
        x, z = _splitmix64(x)
        s.append(z)
    return s


def test_prng_known_answers():
    # SplitMix64 from state 0: its first output is the published 0xe220a8397b1dcdaf
    assert _splitmix64(0)[1] == 0xE220A8397B1DCDAF
    # xoshiro256** from state {1, 2, 3, 4}: rotl(2·5, 7)·9 = 11520, then 0
    s = [1, 2, 3, 4]
    assert [_xoshiro_next(s), _xoshiro_next(s)] == [11520, 0]
    # the oracle's C PRNG draws the same words
    import ctypes as C
    st = (C.c_uint64 * 4)()
    O.lib().orc_rng_seed.argtypes = [C.POINTER(C.c_uint64), C.c_uint64]
    O.lib().orc_rng_next.argtypes = [C.POINTER(C.c_uint64)]
    O.lib().orc_rng_next.restype = C.c_uint64
    O.lib().orc_rng_seed(st, 12345)
    py = _seed(12345)
    assert list(st) == py
    for _ in range(50):
        assert O.lib().orc_rng_next(st) == _xoshiro_next(py)


# ------------------------------------------------------ global pivot finder
def _finder_brute_force(t, f, ld, abstol, nsearch, maxng, margin, seed):
    """globalpivotfinder.jl:156-188 in plain Python: start points rand(1:d)
    (here a 32-bit draw of the project PRNG, modulo d), a scan over every
    site p and value v with the strict `>` rule, the threshold
    abstol·margin and the truncation to maxng."""
    L = len(ld)
    s = _seed(seed)
    pts = [[1 + ((_xoshiro_next(s) >> 32) % d) for d in ld] for _ in range(nsearch)]
    cores = t.sitetensors()
    found = []
    for pt in pts:
        cand = []
        for p in range(L):
            for v in range(1, ld[p] + 1):
                x = list(pt)
                x[p] = v
                cand.append(x)
        xs = np.array(cand, np.int32)
        errs = np.abs(O.eval_points(f, xs) - O.tt_evaluate(cores, xs))
        best_err, best = 0.0, list(pt)
        for e, x in zip(errs, cand):
            if e > best_err:
                best_err, best = e, x
        if best_err > abstol * margin:
            found.append(best)
    return found[:maxng]


_FINDER_CASES = [
    # (spec name, abstol factor, nsearch, maxng, margin, seed)
    ("lorentz5_tight", 1e-12, 5, 5, 10.0, 0),
    ("lorentz5_tight", 1e-9, 10, 3, 10.0, 7),      # truncation to maxng
    ("lorentz5_tight", 0.0, 6, 6, 10.0, 3),        # every point with a nonzero error
    ("gauss_sum", 1e-10, 8, 4, 2.0, 11),
    ("random_tt", 1e-14, 6, 6, 10.0, 5),
    ("weighted_lorentz", 1e-12, 5, 5, 10.0, 1),
]


@pytest.mark.parametrize("case", _FINDER_CASES, ids=lambda c: f"{c[0]}-{c[5]}")
def test_global_finder_matches_brute_force(case):
    name, tolf, nsearch, maxng, margin, seed = case
    spec = [s for s in specs() if s[0] == name][0]
    _, kind, params, ip, ld, piv, opts = spec
    f = O.Functor(kind, params, ip)
    t = O.TCI(ld)
    t.init(f, piv if piv is not None else [[1] * len(ld)])
    # one sweep2site! with loose tolerance leaves errors for the search to find
    t.sweep2site(f, niter=2, abstol=1e-4 * t.maxsamplevalue, maxbonddim=opts.get("maxbonddim", O.INT64_MAX),
                 fillsitetensors=True)
    abstol = tolf * t.maxsamplevalue
    got = t.globalpivotfinder(f, abstol, nsearch, maxng, margin, seed)
    want = _finder_brute_force(t, f, ld, abstol, nsearch, maxng, margin, seed)
    assert got == want
    if name == "lorentz5_tight" and seed == 7:
        assert len(got) == 3            # the truncation was exercised


def test_global_finder_threshold_is_strict():
    # an exact interpolation (rank-1 f): every error is 0, which is not > 0
    f = O.Functor(8, [1.0, -1.0, 0.0, 0.0])      # exp(-x): rank 1 in quantics
    ld = [2] * 6
    t = O.TCI(ld)
    t.init(f, [[1] * 6])
    t.sweep2site(f, niter=2, abstol=0.0, fillsitetensors=True)
    assert t.globalpivotfinder(f, 0.0, 5, 5, 10.0, 0) == _finder_brute_force(t, f, ld, 0.0, 5, 5, 10.0, 0)


# --------------------------------------------------------- fillsitetensors!
_FILL = [s for s in specs() if s[0] in {"lorentz8", "lorentz5_tight", "random_tt", "weighted_lorentz",
                                         "gauss_sum"}] + [c4_spec(0)]


@pytest.mark.parametrize("spec", _FILL, ids=lambda s: s[0])
def test_fillsitetensors_residual(spec):
    """T_b·P = Pi1 to backward-stable rounding, and T_b equals numpy's solve
    to the accuracy the conditioning of P allows; the last site is Pi1."""
    name, kind, params, ip, ld, piv, opts = spec
    f = O.Functor(kind, params, ip)
    t = O.TCI(ld)
    t.init(f, piv if piv is not None else [[1] * len(ld)])
    t.sweep2site(f, niter=2, abstol=opts["tolerance"] * t.maxsamplevalue,
                 maxbonddim=opts.get("maxbonddim", O.INT64_MAX), fillsitetensors=True)
    L = len(ld)
    eps = np.finfo(float).eps
    for b in range(L):
        Ib, Jb = t.Iset(b), t.Jset(b)
        nI, d, nJ = len(Ib), ld[b], len(Jb)
        left = Ib if b > 0 else [[]]
        right = Jb if b < L - 1 else [[]]
        Pi1 = O.batcheval(f, ld, np.array(left, np.int32).reshape(len(left), b),
                          np.array(right, np.int32).reshape(len(right), L - 1 - b)).reshape(nI * d, nJ, order="F")
        T = t.sitetensor(b).reshape(nI * d, -1, order="F")
        if b == L - 1:
            assert np.array_equal(T, Pi1)
            continue
        In = t.Iset(b + 1)
        P = O.batcheval(f, ld, np.array(In, np.int32).reshape(len(In), b + 1),
                        np.array(Jb, np.int32).reshape(nJ, L - 1 - b)).reshape(len(In), nJ, order="F")
        c = P.shape[0]
        res = np.abs(T @ P - Pi1)
        bound = 8 * c * eps * (np.abs(T) @ np.abs(P)) + 1e-300
        assert np.all(res <= 64 * bound), (name, b, float(np.max(res / bound)))
        ref = np.linalg.solve(P.T, Pi1.T).T
        cond = np.linalg.cond(P)
        assert np.allclose(T, ref, rtol=0, atol=100 * c * eps * cond * np.max(np.abs(ref)) + 1e-300), (name, b)
