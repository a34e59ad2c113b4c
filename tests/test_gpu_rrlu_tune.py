"""The standalone rrLU's autotuner (tci_ctx_autotune_rrlu) and tuning tables
(tci_ctx_set_rrlu_tuning): the measured table is well formed, and whatever
kernel a table sends a job to, the pivots, counts and errors equal the
oracle's (rrlu!, matrixlu.jl:141-204)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(8, 5, 3), (64, 64, 20), (128, 100, 40), (200, 150, 30), (517, 611, 8), (640, 640, 64), (1000, 37, 20)]


def _mats(seed):
    rng = np.random.default_rng(seed)
    out = []
    for m, n, r in SHAPES:
        X = rng.standard_normal((m, r))
        Y = rng.standard_normal((r, n))
        out.append(X @ Y + 1e-12 * rng.standard_normal((m, n)))
    return out


def _check_vs_oracle(tci, ctx, mats, what):
    res = tci.rrlu_batched(mats, reltol=1e-10, ctx=ctx)
    for A, g in zip(mats, res):
        o = O.rrlu(A, reltol=1e-10)
        assert g["status"] == 0, (what, A.shape)
        assert g["npivot"] == o.npivot, (what, A.shape)
        assert np.array_equal(g["rowpermutation"], o.rowpermutation), (what, A.shape)
        assert np.array_equal(g["colpermutation"], o.colpermutation), (what, A.shape)
        assert g["error"] == o.error, (what, A.shape)


@pytest.fixture
def own_ctx(tci, ctx):
    c = tci.Context(0)
    yield c
    c.close()


def test_autotune_table_and_auto_parity(tci, own_ctx):
    table = own_ctx.autotune_rrlu()
    for row in table:
        print("rrlu tune", row, flush=True)
    sides = [r["side"] for r in table]
    assert len(table) >= 6 and sides == sorted(sides) and sides[0] == 32
    for r in table:
        assert r["grid_team"] in (256, 512, 1024)
        assert 0 <= r["grid_max_jobs"] <= 64
        assert r["ms_stream"] > 0 and r["ms_grid"] > 0
        assert r["tile"] in (0, 1) and (r["tile"] == 0 or r["side"] <= 128)
        if r["side"] <= 128:
            assert r["ms_tile_batch"] > 0 and r["ms_stream_batch"] > 0
    _check_vs_oracle(tci, own_ctx, _mats(3), "tuned AUTO")
    # the table can be read back into another context
    own_ctx.set_rrlu_tuning(table)
    _check_vs_oracle(tci, own_ctx, _mats(4), "installed table")


def _row(side, tile, team, maxjobs):
    return dict(side=side, tile=tile, grid_team=team, grid_max_jobs=maxjobs)


@pytest.mark.parametrize("table", [
    [_row(128, 0, 256, 64), _row(8192, 0, 512, 64)],      # every job to the cooperative kernel
    [_row(128, 1, 1024, 0), _row(8192, 0, 1024, 0)],      # tile, then streaming, never cooperative
    [_row(64, 1, 512, 0), _row(256, 0, 256, 64), _row(8192, 0, 1024, 1)],   # mixed routes
])
def test_tables_route_jobs_without_changing_results(tci, own_ctx, table):
    own_ctx.set_rrlu_tuning(table)
    _check_vs_oracle(tci, own_ctx, _mats(5), str(table))
    own_ctx.set_rrlu_tuning([])
    _check_vs_oracle(tci, own_ctx, _mats(6), "fixed rules")


def test_bad_tables_are_refused(tci, own_ctx):
    for bad in ([_row(256, 0, 256, 0), _row(128, 0, 256, 0)],     # sides not ascending
                [_row(128, 0, 300, 0)],                            # team
                [_row(256, 1, 256, 0)],                            # tile beyond 128
                [_row(128, 0, 256, -1)]):                          # negative job count
        with pytest.raises(tci.TCIError):
            own_ctx.set_rrlu_tuning(bad)
