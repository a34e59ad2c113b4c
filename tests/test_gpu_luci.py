"""MatrixLUCI left/right on the device (tci_luci_f64, matrixluci.jl:40-84)
against the oracle's restatement bit for bit, in both orientations, and the
reference's own checks of test_matrixluci.jl: the cross identities
colstimespivotinv = A[:,J]·A[I,J]⁻¹ and pivotinvtimesrows = A[I,J]⁻¹·A[I,:]
(29-30), left·right equal to the cross (37), the low-rank approximation
bound (41-46) and exact rank 3 (66-73)."""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
FX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_fixtures.json")))


def _both(tci, A, **kw):
    d = tci.MatrixLUCI(A, **kw)
    lu, left, right = O.luci_factors(A, **kw)
    assert d.npivot == lu.npivot
    assert np.array_equal(d.rowpermutation, lu.rowpermutation)
    assert np.array_equal(d.colpermutation, lu.colpermutation)
    assert np.array_equal(d.left(), left) and np.array_equal(d.right(), right)
    return d


@pytest.mark.parametrize("lo", [True, False])
def test_luci_equals_oracle(tci, ctx, lo):
    rng = np.random.default_rng(31)
    for t in range(12):
        m, n = (int(x) for x in rng.integers(1, 160, 2))
        r = int(rng.integers(1, 15))
        A = rng.standard_normal((m, r)) @ rng.standard_normal((r, n)) + 1e-11 * rng.standard_normal((m, n))
        _both(tci, A, reltol=1e-9, leftorthogonal=lo)
        _both(tci, A, maxrank=3, leftorthogonal=lo)
    for name in ("luci_8x6", "approx_8x6", "rrlu_4x4", "lastpivot_5x5"):
        _both(tci, np.array(FX[name]["A"]), leftorthogonal=lo)


def test_luci_cross_identities(tci, ctx):
    # test_matrixluci.jl:18-37 on its 8×6 matrix with maxrank 4
    A = np.array(FX["luci_8x6"]["A"])
    lo = tci.MatrixLUCI(A, maxrank=4)
    I, J = lo.rowindices() - 1, lo.colindices() - 1
    assert len(I) == 4 and len(J) == 4
    P = A[np.ix_(I, J)]
    np.testing.assert_allclose(lo.left(), A[:, J] @ np.linalg.inv(P), rtol=1e-10, atol=1e-12)
    ro = tci.MatrixLUCI(A, maxrank=4, leftorthogonal=False)
    np.testing.assert_allclose(ro.right(), np.linalg.inv(P) @ A[I, :], rtol=1e-10, atol=1e-12)
    cross = A[:, J] @ np.linalg.inv(P) @ A[I, :]
    np.testing.assert_allclose(lo.left() @ lo.right(), cross, rtol=1e-10, atol=1e-12)
    assert lo.left().shape == (8, 4) and lo.right().shape == (4, 6)


def test_luci_lowrank_bound_and_exact_rank(tci, ctx):
    A = np.array(FX["luci_8x6"]["A"])
    rng = np.random.default_rng(0)
    B = np.hstack([A, A + 1e-3 * rng.random((8, 6))])
    lu = tci.MatrixLUCI(B, reltol=1e-2)
    assert len(lu.rowindices()) < 8 and len(lu.colindices()) < 12
    assert np.max(np.abs(lu.left() @ lu.right() - B)) < 1e-2
    pq = np.array(FX["luci_p"]["A"]) @ np.array(FX["luci_q"]["A"])
    e = tci.MatrixLUCI(pq)
    assert e.npivots() == 3
    np.testing.assert_allclose(e.left() @ e.right(), pq, rtol=1e-12, atol=1e-12)
