"""Write the golden fixtures for the rrLU / LUCI / TCI2 parity tests.

Run in a container that has the reference checkout (read as TEXT only; Julia
is not available, nothing of the reference is executed):

    python tests/golden/make_goldens.py /root/reference

1. reference_fixtures.json — the matrices hard-coded in the reference's own
   tests (test/test_matrixlu.jl, test/test_matrixluci.jl), copied as data with
   their file:line, together with the known answers those tests assert.
2. rrlu_pivots.json — golden pivot sequences on those matrices and on seeded
   random / tie-heavy matrices, produced by the C restatement (oracle/) and
   accepted only where the independent numpy restatement (oracle/rrlu_np.py)
   agrees bit for bit.  The reference publishes no pivot sequences.
3. tci_c1.json — the 8-site Lorentzian (README.md:24-28, tolerance 1e-8)
   solved by the oracle: ranks, errors, pivot sets, maxsamplevalue.
4. configs.json (`make_goldens.py --configs`, needs no reference checkout) —
   the full-size solve configurations of tests/problems.py solved by the
   oracle: ranks, link dims, errors and a digest of the pivot sets; and
   config_errors.json, their max|TT - f| on 10^5 seeded random points.
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def julia_matrix(path, first_line):
    """Parse the first `[ ... ]` matrix literal starting at 1-based line first_line."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    text = "\n".join(lines[first_line - 1:])
    i = text.index("[")
    depth, j = 0, i
    while True:
        if text[j] == "[":
            depth += 1
        elif text[j] == "]":
            depth -= 1
            if depth == 0:
                break
        j += 1
    body = text[i + 1:j]
    rows = [r for r in re.split(r"[;\n]", body) if r.strip()]
    mat = [[float(t) for t in r.split()] for r in rows]
    last_line = first_line + text[:j].count("\n")
    return np.array(mat), last_line


def fixtures(ref):
    t_lu = os.path.join(ref, "test", "test_matrixlu.jl")
    t_luci = os.path.join(ref, "test", "test_matrixluci.jl")
    out = {}

    def add(name, path, line, expect, scale=1.0, note=""):
        A, last = julia_matrix(path, line)
        rel = os.path.relpath(path, ref)
        out[name] = dict(source=f"{rel}:{line}-{last}", A=(scale * A).tolist(), expect=expect, note=note)

    add("argmax_10x8", t_lu, 8, {"submatrixargmax_3_5": [3, 5], "note": "test_matrixlu.jl:21-28"})
    add("rrlu_4x4", t_lu, 55, {"unit_lower_L": True, "upper_U": True, "reconstructs": True})
    add("truncated_3x3", t_lu, 89, {"npivot": 1})
    add("approx_8x6", t_lu, 100, {"maxrank4_rows": 4, "maxrank4_cols": 4})
    add("lowrank_p", t_lu, 142, {})
    add("lowrank_q", t_lu, 154, {})
    add("identity_2x2", t_lu, 168, {"pivoterrors": [1.0, 1.0, 0.0], "lastpivoterror": 0.0})
    add("lastpivot_5x5", t_lu, 178, {"maxrank2_len_pivoterrors": 3, "maxrank2_last_positive": True,
                                     "abstol05_last_below": 0.5, "abstol0_last": 0.0})
    add("tiny_4x4", t_lu, 198, {"npivots_abstol_1e-3": 1}, scale=1e-13,
        note="the test multiplies the literal by 1e-13 (test_matrixlu.jl:198)")
    add("luci_8x6", t_luci, 7, {"maxrank4_rows": 4})
    add("luci_p", t_luci, 49, {})
    add("luci_q", t_luci, 61, {})
    return out


def golden_pivots(fx):
    import oracle as O
    from rrlu_np import rrlu_np
    cases = []

    def case(name, A, **kw):
        A = np.asarray(A, float)
        a = O.rrlu(A, **kw)
        npkw = dict(kw)
        if "maxrank" in npkw and npkw["maxrank"] >= (1 << 62):
            npkw.pop("maxrank")
        b = rrlu_np(A, **npkw)
        assert list(a.rowpermutation) == list(b["rowpermutation"]), name
        assert list(a.colpermutation) == list(b["colpermutation"]), name
        assert a.npivot == b["npivot"], name
        assert np.array_equal(a.L, b["L"]) and np.array_equal(a.U, b["U"]), name
        assert (a.error == b["error"]) or (np.isnan(a.error) and np.isnan(b["error"])), name
        cases.append(dict(name=name, A=A.tolist(), kwargs=kw, rowpermutation=a.rowpermutation.tolist(),
                          colpermutation=a.colpermutation.tolist(), npivot=a.npivot,
                          error=a.error if np.isfinite(a.error) else repr(a.error),
                          L=a.L.tolist(), U=a.U.tolist()))

    for k, v in fx.items():
        A = np.array(v["A"])
        case(k, A)
        case(k + "_right", A, leftorthogonal=False)
    case("approx_8x6_maxrank4", np.array(fx["approx_8x6"]["A"]), maxrank=4)
    case("lastpivot_5x5_maxrank2", np.array(fx["lastpivot_5x5"]["A"]), maxrank=2)
    case("lastpivot_5x5_abstol05", np.array(fx["lastpivot_5x5"]["A"]), abstol=0.5)
    case("tiny_4x4_abstol", np.array(fx["tiny_4x4"]["A"]), abstol=1e-3)
    pq = np.array(fx["lowrank_p"]["A"]) @ np.array(fx["lowrank_q"]["A"])
    case("lowrank_pq", pq)
    # seeded random and tie-heavy matrices (project PRNG via numpy Generator)
    rng = np.random.default_rng(20240917)
    for t in range(6):
        m, n = rng.integers(3, 24, 2)
        case(f"random_{t}", rng.standard_normal((m, n)))
        case(f"ties_{t}", rng.integers(-2, 3, (m, n)).astype(float))
        case(f"ties_right_{t}", rng.integers(-2, 3, (m, n)).astype(float), leftorthogonal=False)
    return cases


def tci_c1():
    import oracle as O
    f = O.Functor(1, [1.0])
    res = {}
    for ns in (0, 5):
        t, r, e = O.crossinterpolate2(f, [10] * 8, tolerance=1e-8, nsearchglobalpivot=ns)
        res[f"nsearch{ns}"] = dict(ranks=r.tolist(), errors=e.tolist(), linkdims=t.linkdims(),
                                   maxsamplevalue=t.maxsamplevalue,
                                   Iset=[t.Iset(b) for b in range(8)], Jset=[t.Jset(b) for b in range(8)])
    return res


def set_digest(t, L):
    """sha256 of every Iset and Jset of a solve, in site order."""
    import hashlib
    h = hashlib.sha256()
    for b in range(L):
        h.update(json.dumps([t.Iset(b), t.Jset(b)]).encode())
    return h.hexdigest()


def configs():
    """The solve configurations at full size (tests/problems.py config_specs),
    solved by the oracle: ranks, link dims, errors, global pivots, max sample
    and a digest of the pivot sets."""
    import oracle as O
    sys.path.insert(0, os.path.dirname(HERE))
    from problems import config_specs, oracle_opts
    res = {}
    for name, kind, params, ip, ld, piv, opts in config_specs():
        t, r, e = O.crossinterpolate2(O.Functor(kind, params, ip), ld, piv, **oracle_opts(opts))
        res[name] = dict(ranks=r.tolist(), errors=e.tolist(), linkdims=t.linkdims(),
                         nglobalpivots=t.nglobalpivots.tolist(), maxsamplevalue=t.maxsamplevalue,
                         sets_sha256=set_digest(t, len(ld)), workstats=t.workstats())
    return res


def config_errors(npts=100_000, seed=12345):
    """max|TT - f| over npts seeded random points (numpy Generator(seed), one
    integers(1, d+1, npts) draw per site), normalized by max|f| over the
    points and by the solve's maxsamplevalue (the reference's normalization,
    tensorci2.jl:769-770); f in host FP64."""
    import oracle as O
    sys.path.insert(0, os.path.dirname(HERE))
    from problems import config_specs, oracle_opts
    res = {}
    for name, kind, params, ip, ld, piv, opts in config_specs():
        f = O.Functor(kind, params, ip)
        t, r, e = O.crossinterpolate2(f, ld, piv, **oracle_opts(opts))
        rng = np.random.default_rng(seed)
        x = np.stack([rng.integers(1, int(d) + 1, npts) for d in ld], 1).astype(np.int32)
        tt = O.tt_evaluate(t.sitetensors(), x)
        fv = O.eval_points(f, x)
        me, mf = float(np.max(np.abs(tt - fv))), float(np.max(np.abs(fv)))
        res[name] = dict(npts=npts, seed=seed, max_abs_err=me, max_abs_f=mf, maxsamplevalue=t.maxsamplevalue,
                         err_maxf=me / mf, err_maxsample=me / t.maxsamplevalue, tolerance=opts["tolerance"],
                         meets_tol=bool(me / t.maxsamplevalue <= opts["tolerance"]))
    return res


def main():
    if sys.argv[1:] == ["--configs"]:
        with open(os.path.join(HERE, "configs.json"), "w") as fh:
            json.dump(configs(), fh, indent=1)
        with open(os.path.join(HERE, "config_errors.json"), "w") as fh:
            json.dump(config_errors(), fh, indent=1)
        print("wrote configs.json")
        return
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    fx = fixtures(ref)
    with open(os.path.join(HERE, "reference_fixtures.json"), "w") as fh:
        json.dump(fx, fh, indent=1)
    with open(os.path.join(HERE, "rrlu_pivots.json"), "w") as fh:
        json.dump(golden_pivots(fx), fh)
    with open(os.path.join(HERE, "tci_c1.json"), "w") as fh:
        json.dump(tci_c1(), fh)
    with open(os.path.join(HERE, "configs.json"), "w") as fh:
        json.dump(configs(), fh, indent=1)
    with open(os.path.join(HERE, "config_errors.json"), "w") as fh:
        json.dump(config_errors(), fh, indent=1)
    print("wrote", len(fx), "reference matrices")


if __name__ == "__main__":
    main()
