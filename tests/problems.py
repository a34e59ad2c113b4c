"""Problem specs shared by the CPU (engine build) and GPU parity tests.

Each spec: (name, kind, params, iparams, localdims, initialpivots, options).
Functor kinds and parameter layouts: include/tci_functors.h.
"""
import numpy as np

LORENTZ, WLORENTZ, QGC, GSUM, SUM, TABLE, TT, QEXP = 1, 2, 3, 4, 5, 6, 7, 8


def gauss_sum(K, n, d, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 2.0, K)
    w = rng.uniform(0.2, 0.6, K)
    c = rng.uniform(0, 1, (K, n))
    return np.concatenate([[K, d], a, w, c.ravel()])


def quantics_gc(K, seed, A=0.1, F=2.0 ** 6):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 2, K)
    c = rng.uniform(0, 1, K)
    w = 10.0 ** rng.uniform(-2, -1, K)
    trip = np.stack([a, c, w], 1).ravel()
    return np.concatenate([[K], trip, [A, F]])


def random_tt(bd, ld, seed):
    rng = np.random.default_rng(seed)
    cores = [rng.random((bd[i], ld[i], bd[i + 1])) for i in range(len(ld))]
    return np.concatenate([c.ravel(order="F") for c in cores]), list(bd) + list(ld), cores


def specs():
    tt_params, tt_ip, _ = random_tt([1, 2, 3, 2, 1], [2, 3, 3, 2], 5)
    tab = np.diag([1.0, 1e-5, 0.0])
    rngw = np.random.default_rng(4)
    return [
        ("lorentz8", LORENTZ, [1.0], [], [10] * 8, None, dict(tolerance=1e-8)),
        ("lorentz8_nosearch", LORENTZ, [1.0], [], [10] * 8, None, dict(tolerance=1e-8, nsearchglobalpivot=0)),
        ("lorentz5_tight", LORENTZ, [1.0], [], [10] * 5, None, dict(tolerance=1e-12, maxiter=200)),
        ("lorentz5_pivots", LORENTZ, [1.0], [], [10] * 5,
         [[1, 1, 1, 1, 1], [10, 8, 10, 4, 4], [5, 4, 8, 9, 3], [7, 7, 10, 5, 9], [7, 7, 10, 5, 9]],
         dict(tolerance=1e-12, maxiter=200)),
        ("lorentz5_forward", LORENTZ, [0.5], [], [10] * 5, None,
         dict(tolerance=1e-8, maxiter=8, sweepstrategy="forward")),
        ("lorentz6_backward_strict", LORENTZ, [1.0], [], [6] * 6, None,
         dict(tolerance=1e-9, sweepstrategy="backward", strictlynested=True, nsearchglobalpivot=0)),
        ("weighted_lorentz", WLORENTZ, np.concatenate([[16.0], rngw.uniform(0.5, 2, 6)]), [], [16] * 6, None,
         dict(tolerance=1e-10, maxbonddim=40)),
        ("gauss_sum", GSUM, gauss_sum(3, 5, 8, 1), [], [8] * 5, None, dict(tolerance=1e-8, maxbonddim=32)),
        ("gauss_sum_d2", GSUM, gauss_sum(2, 9, 2, 2), [], [2] * 9, None, dict(tolerance=1e-10)),
        ("quantics_gc", QGC, quantics_gc(3, 3), [], [2] * 12, None, dict(tolerance=1e-10, maxbonddim=50)),
        ("quantics_exp_maxbond1", QEXP, [1.0, -1.0, 0.0, 0.0], [], [2] * 8, [[1] * 8, [1] + [2] * 7],
         dict(tolerance=1e-4, maxbonddim=1, maxiter=2, normalizeerror=False)),
        ("quantics_exp_small_maxbond", QEXP, [1.0, -1.0, 1e-4, -2.0], [], [2] * 8, [[1] * 8, [1] + [2] * 7],
         dict(tolerance=1e-10, maxbonddim=1, maxiter=10, normalizeerror=False, nsearchglobalpivot=10,
              maxnglobalpivot=5)),
        ("random_tt", TT, tt_params, tt_ip, [2, 3, 3, 2], None, dict(tolerance=1e-10, maxbonddim=10)),
        ("diag_table", TABLE, tab.ravel(order="F"), [3, 3], [3, 3], [[1, 1]], dict(tolerance=1e-8)),
        ("sum_20sites", SUM, [1.0], [], [2] * 20, None, dict(tolerance=1e-8)),
    ]


def _mode(nsearch):
    return "_parity" if nsearch == 0 else ""


def c1_spec(nsearch=5):
    """Plain TCI: 1/(1+v.v) on {1..10}^8, tol 1e-8, initial pivot ones (README.md:24-28)."""
    return ("c1_lorentz8" + _mode(nsearch), LORENTZ, [1.0], [], [10] * 8, None,
            dict(tolerance=1e-8, nsearchglobalpivot=nsearch))


def c3_params(seed=3):
    """Quantics workload: 6 Gaussians with centers U(0,1), widths 10^U(-6,-1),
    amplitudes U(0.5,2), plus 0.1 cos(2^20 x) (seeded numpy Generator)."""
    rng = np.random.default_rng(seed)
    K = 6
    a = rng.uniform(0.5, 2, K)
    c = rng.uniform(0, 1, K)
    w = 10.0 ** rng.uniform(-6, -1, K)
    return np.concatenate([[K], np.stack([a, c, w], 1).ravel(), [0.1, 2.0 ** 20]])


def c3_spec(nsearch=5):
    """Quantics TCI: 30 sites of d = 2, tol 1e-10, maxbonddim 200."""
    return ("c3_quantics30" + _mode(nsearch), QGC, c3_params(), [], [2] * 30, None,
            dict(tolerance=1e-10, maxbonddim=200, nsearchglobalpivot=nsearch))


def c4_params(seed=4, L=40):
    """High-dimensional workload: 1/(1 + sum c_i (v_i/16)^2), c_i ~ U(0.5, 2)."""
    rng = np.random.default_rng(seed)
    return np.concatenate([[16.0], rng.uniform(0.5, 2, L)])


def c4_spec(nsearch=5):
    """High-dimensional TCI: 40 sites of d = 16, tol 1e-10, maxbonddim 256."""
    return ("c4_highdim40" + _mode(nsearch), WLORENTZ, c4_params(), [], [16] * 40, None,
            dict(tolerance=1e-10, maxbonddim=256, nsearchglobalpivot=nsearch))


def config_specs():
    """The solve configurations at their full sizes, in parity mode (no
    global search, fully deterministic) and with the optimize!
    defaults (global search with the project PRNG)."""
    return [c1_spec(0), c1_spec(5), c3_spec(0), c3_spec(5), c4_spec(0), c4_spec(5)]


def oracle_opts(opts):
    return {k: v for k, v in opts.items() if k != "chi_cap"}


def served_mix(count, seed=5, chi_cap=0):
    """Requests of the served workload: each a sum of K ~ U{1..8} separable
    Gaussians in n ~ U{4..12} dims, one local dimension d ∈ {2,4,8,16,32} for
    all sites, tolerance 10^U(-10,-6), maxbonddim 128.  Returns specs in the
    format of specs()."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        K = int(rng.integers(1, 9))
        n = int(rng.integers(4, 13))
        d = int(rng.choice([2, 4, 8, 16, 32]))
        tol = float(10.0 ** rng.uniform(-10, -6))
        params = gauss_sum(K, n, d, int(rng.integers(0, 2 ** 31)))
        out.append((f"served{q}", GSUM, params, [], [d] * n, None,
                    dict(tolerance=tol, maxbonddim=128, chi_cap=chi_cap)))
    return out


def rook_specs():
    """The specs above with pivotsearch = :rook (tensorci2.jl:552-595), for
    those whose reference tests also run with rook (test_tensorci2.jl:55,
    247, 343) and a few more shapes."""
    keep = {"lorentz5_tight", "lorentz5_pivots", "lorentz5_forward", "lorentz6_backward_strict", "gauss_sum",
            "quantics_gc", "quantics_exp_maxbond1", "random_tt", "diag_table", "lorentz8_nosearch"}
    return [(name + "_rook", kind, params, ip, ld, piv, dict(opts, pivotsearch="rook"))
            for name, kind, params, ip, ld, piv, opts in specs() if name in keep]
