"""The solve engine's team choice (tci_ctx_autotune_solve,
tci_ctx_set_solve_tuning): a batch gives the same bits with 64, 256, 512 or 1024
threads per solve, the autotuner keeps the fastest team, and bad tables are
refused."""
import pytest

from problems import rook_specs, specs
from test_gpu_serving import _req, _same

pytestmark = pytest.mark.gpu

NAMES = ("lorentz5_tight", "lorentz8", "weighted_lorentz", "gauss_sum", "quantics_gc", "random_tt", "diag_table",
         "sum_20sites")


@pytest.fixture
def own_ctx(tci, ctx):
    c = tci.Context(0)
    yield c
    c.close()


def test_every_team_gives_the_same_bits(tci, own_ctx):
    reqs = [_req(tci, s) for s in specs() if s[0] in NAMES]
    reqs += [_req(tci, s) for s in rook_specs()[:3]]
    assert len(reqs) >= len(NAMES)
    own_ctx.set_solve_tuning([])
    base = tci.crossinterpolate2_batch(reqs, ctx=own_ctx)
    for team in (64, 512, 1024):
        own_ctx.set_solve_tuning([(0, 0, team)])
        got = tci.crossinterpolate2_batch(reqs, ctx=own_ctx)
        for i, (a, b) in enumerate(zip(got, base)):
            _same(a, b, f"team {team}, request {i}")
        # one at a time as well (each its own launch and plan)
        for i, r in enumerate(reqs[:4]):
            (a,) = tci.crossinterpolate2_batch([r], ctx=own_ctx)
            _same(a, base[i], f"team {team}, alone {i}")


def test_autotune_keeps_the_fastest_team(tci, own_ctx):
    (spec,) = [s for s in specs() if s[0] == "weighted_lorentz"]
    team, ms = own_ctx.autotune_solve(_req(tci, spec), batch=128)
    print("solve tune weighted_lorentz: team", team, "ms", ms, flush=True)
    assert team in (64, 256, 512, 1024)
    assert all(m > 0 for m in ms)
    assert ms[(64, 256, 512, 1024).index(team)] == min(ms)
    # the tuned context still solves to the same bits
    own_ctx.set_solve_tuning([])
    (a,) = tci.crossinterpolate2_batch([_req(tci, spec)], ctx=own_ctx)
    own_ctx.set_solve_tuning([(0, 0, team)])
    (b,) = tci.crossinterpolate2_batch([_req(tci, spec)], ctx=own_ctx)
    _same(a, b, "tuned")


def test_bad_solve_tables_are_refused(tci, own_ctx):
    for bad in ([(0, 0, 300)], [(-1, 0, 256)], [(0, -1, 256)]):
        with pytest.raises(tci.TCIError):
            own_ctx.set_solve_tuning(bad)
