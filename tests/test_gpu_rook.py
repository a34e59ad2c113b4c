"""pivotsearch = :rook (arrlu on the lazily evaluated Pi, tensorci2.jl:552-595,
matrixlu.jl:227-293) in whole device solves against the oracle, bit for
bit: each rook spec alone, all in one launch, and a C1-sized solve (L=8,
d=10, tolerance 1e-8) whose sweep evaluates only the submatrices the rook
search asks for."""
import numpy as np
import pytest

from problems import LORENTZ, rook_specs
from test_gpu_solves import _check, _tci_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spec", rook_specs(), ids=lambda s: s[0])
def test_rook_solve_alone(tci, ctx, spec):
    print("rook solve", spec[0], flush=True)
    (t,) = tci.crossinterpolate2_batch([_tci_problem(tci, spec)])
    _check(t, spec)


def test_rook_all_in_one_launch(tci, ctx):
    sp = rook_specs()
    ts = tci.crossinterpolate2_batch([_tci_problem(tci, s) for s in sp])
    for t, s in zip(ts, sp):
        _check(t, s)


def test_rook_lorentz8_evaluates_less(tci, ctx):
    """The rook solve of 1/(1+v.v) on {1..10}^8 matches the oracle and
    evaluates fewer Pi entries than the full search of the same problem."""
    full = ("c1_full", LORENTZ, [1.0], [], [10] * 8, None, dict(tolerance=1e-8, nsearchglobalpivot=0))
    rook = ("c1_rook", LORENTZ, [1.0], [], [10] * 8, None,
            dict(tolerance=1e-8, nsearchglobalpivot=0, pivotsearch="rook"))
    tf, tr = tci.crossinterpolate2_batch([_tci_problem(tci, full), _tci_problem(tci, rook)])
    _check(tf, full)
    _check(tr, rook)
    print("Pi entries: full", tf.workstats["pi_elements"], "rook", tr.workstats["pi_elements"], flush=True)
    assert tr.workstats["pi_elements"] < tf.workstats["pi_elements"]
