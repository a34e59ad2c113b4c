"""Slot-level serving (tci_server_*): a persistent grid that takes each
request from a device queue as a workgroup frees.  Every served request
equals its solve alone, bit for bit; the grid leaves when idle and starts
again on the next request; stop runs every published request first; a full
queue is refused with TCI_ERR_OOM."""
import ctypes as C
import time

import numpy as np
import pytest

from problems import served_mix, specs
from test_gpu_serving import _req, _same

pytestmark = pytest.mark.gpu


def _serving():
    from importlib import import_module
    return import_module("tci_amd.serving")


def _shape(tci):
    return (tci.GaussSum([1.0], [0.3], np.zeros((1, 12)), 32), [32] * 12, None, dict(tolerance=1e-8))


def test_slot_served_equals_alone(tci, ctx):
    serving = _serving()
    reqs = [_req(tci, s) for s in served_mix(40, seed=13)]
    reqs += [_req(tci, s) for s in specs() if s[0] in ("lorentz8", "gauss_sum", "random_tt", "quantics_gc")]
    arrivals = serving.poisson_arrivals(len(reqs), rate_per_s=2000.0, seed=4)
    srv = serving.SlotServer(ctx, _shape(tci), slots=64, idle_ms=20)
    try:
        res, stats = srv.run(reqs, arrivals, keep_tci=True)
    finally:
        srv.close()
    for i, r in enumerate(reqs):
        (alone,) = tci.crossinterpolate2_batch([r], ctx=ctx)
        _same(res[i].tci, alone, f"request {i}")
    print("slot serving p50/p99 ms", stats.p50_ms, stats.p99_ms, flush=True)


def test_slot_server_idles_and_restarts(tci, ctx):
    serving = _serving()
    reqs = [_req(tci, s) for s in served_mix(6, seed=17)]
    srv = serving.SlotServer(ctx, _shape(tci), slots=8, idle_ms=2)
    try:
        res1, _ = srv.run(reqs[:3], np.zeros(3), keep_tci=True)
        time.sleep(0.05)                     # the grid has left by now
        res2, _ = srv.run(reqs[3:], np.zeros(3), keep_tci=True)
    finally:
        srv.close()
    for r, q in zip(res1 + res2, reqs):
        (alone,) = tci.crossinterpolate2_batch([q], ctx=ctx)
        _same(r.tci, alone, "restart")


def test_slot_server_stop_drains_and_queue_full(tci, ctx):
    lib = tci._abi.lib()
    reqs = [_req(tci, s) for s in served_mix(3, seed=19)]
    holders = [tci._holder(*q) for q in reqs]
    shape = tci._holder(*_shape(tci))
    h = C.c_void_p()
    assert lib.tci_server_start(ctx.handle, C.byref(shape.st), 2, 50, 0, C.byref(h)) == 0
    bs = []
    try:
        for k in range(2):
            b = C.c_void_p()
            assert lib.tci_server_submit(h, C.byref(holders[k].st), C.byref(b)) == 0
            bs.append(b)
        b = C.c_void_p()
        assert lib.tci_server_submit(h, C.byref(holders[2].st), C.byref(b)) == 11   # ring of 2 is full
    finally:
        assert lib.tci_server_stop(h) == 0                                         # runs both first
    for k, b in enumerate(bs):
        assert lib.tci_batch_query(b) == 0
        t = tci.TensorCI2(b, 0, reqs[k][1])
        (alone,) = tci.crossinterpolate2_batch([reqs[k]], ctx=ctx)
        _same(t, alone, f"drained {k}")
        lib.tci_batch_free(b)


def test_slot_server_poll_returns_each_once(tci, ctx):
    lib = tci._abi.lib()
    reqs = [_req(tci, s) for s in served_mix(12, seed=23)]
    holders = [tci._holder(*q) for q in reqs]
    shape = tci._holder(*_shape(tci))
    h = C.c_void_p()
    assert lib.tci_server_start(ctx.handle, C.byref(shape.st), 16, 50, 0, C.byref(h)) == 0
    handles = {}
    try:
        for k, hd in enumerate(holders):
            b = C.c_void_p()
            assert lib.tci_server_submit(h, C.byref(hd.st), C.byref(b)) == 0
            handles[b.value] = (k, b)
        out = (C.c_void_p * 4)()
        n = C.c_int32()
        got = []
        t_end = time.time() + 60
        while len(got) < len(reqs) and time.time() < t_end:
            assert lib.tci_server_poll(h, out, 4, C.byref(n)) == 0
            assert 0 <= n.value <= 4
            got += [out[q] for q in range(n.value)]
            time.sleep(1e-3)
        assert sorted(got) == sorted(handles)          # every request, once
        assert lib.tci_server_poll(h, out, 4, C.byref(n)) == 0 and n.value == 0
        for v in got:
            k, b = handles[v]
            assert lib.tci_batch_query(b) == 0
            t = tci.TensorCI2(b, 0, reqs[k][1])
            (alone,) = tci.crossinterpolate2_batch([reqs[k]], ctx=ctx)
            _same(t, alone, f"polled {k}")
    finally:
        for _, b in handles.values():
            lib.tci_batch_free(b)
        assert lib.tci_server_stop(h) == 0


def test_slot_server_ring_wraps_and_reuses_blocks(tci, ctx):
    """More requests than ring slots: slots and memory blocks are reused as
    requests retire, and every result still equals its solve alone."""
    serving = _serving()
    reqs = [_req(tci, s) for s in served_mix(30, seed=29)]
    srv = serving.SlotServer(ctx, _shape(tci), slots=8, idle_ms=20)
    try:
        res, stats = srv.run(reqs, np.zeros(len(reqs)), keep_tci=True)
    finally:
        srv.close()
    assert stats.oom_waits > 0                        # the ring of 8 was full at times
    for i, r in enumerate(reqs):
        (alone,) = tci.crossinterpolate2_batch([r], ctx=ctx)
        _same(res[i].tci, alone, f"request {i}")


def test_slot_server_submit_many_partial_bursts(tci, ctx):
    """A burst larger than the ring is taken in part (TCI_OK, accepted < count);
    the rest goes in after retirements; every result equals its solve alone."""
    lib = tci._abi.lib()
    reqs = [_req(tci, s) for s in served_mix(10, seed=31)]
    holders = [tci._holder(*q) for q in reqs]
    shape = tci._holder(*_shape(tci))
    h = C.c_void_p()
    assert lib.tci_server_start(ctx.handle, C.byref(shape.st), 4, 50, 512, C.byref(h)) == 0
    got = {}
    try:
        nxt = 0
        out = (C.c_void_p * 10)()
        acc = C.c_int32()
        polled = (C.c_void_p * 10)()
        n = C.c_int32()
        handles = {}
        t_end = time.time() + 60
        while len(got) < len(reqs) and time.time() < t_end:
            if nxt < len(reqs):
                rest = len(reqs) - nxt
                arr = (tci._abi.tci_problem * rest)(*[hd.st for hd in holders[nxt:]])
                assert lib.tci_server_submit_many(h, rest, arr, out, C.byref(acc)) == 0
                assert 0 <= acc.value <= min(rest, 4)
                for q in range(acc.value):
                    handles[out[q]] = (nxt + q, C.c_void_p(out[q]))
                nxt += acc.value
            assert lib.tci_server_poll(h, polled, 10, C.byref(n)) == 0
            for q in range(n.value):
                k, b = handles.pop(polled[q])
                got[k] = tci.TensorCI2(b, 0, reqs[k][1])
                lib.tci_batch_free(b)
            time.sleep(1e-3)
        assert sorted(got) == list(range(len(reqs)))
    finally:
        assert lib.tci_server_stop(h) == 0
    for k, r in enumerate(reqs):
        (alone,) = tci.crossinterpolate2_batch([r], ctx=ctx)
        _same(got[k], alone, f"burst {k}")


def test_slot_server_capacity_retry(tci, ctx):
    """A request whose rank outgrows its default set capacity (a random 10^4
    table, middle rank 100) ends with TCI_ERR_CAPACITY on the device and is
    submitted again at twice the capacity; the served result equals
    crossinterpolate2 with its own retries, and the stream's other requests
    are unaffected."""
    serving = _serving()
    rng = np.random.default_rng(7)
    table = (tci.TableFunctor(rng.random((10, 10, 10, 10))), [10] * 4, None, dict(tolerance=1e-12))
    reqs = [table] + [_req(tci, s) for s in served_mix(5, seed=37)]
    srv = serving.SlotServer(ctx, _shape(tci), slots=16, idle_ms=20)
    try:
        res, stats = srv.run(reqs, np.zeros(len(reqs)), keep_tci=True)
    finally:
        srv.close()
    assert stats.retries >= 1
    assert res[0].status == 0 and max(res[0].tci.linkdims()) > 64
    t, _, _ = tci.crossinterpolate2(float, *table[:1], table[1], **table[3])
    _same(res[0].tci, t, "table")
    for i, r in enumerate(reqs[1:], 1):
        (alone,) = tci.crossinterpolate2_batch([r], ctx=ctx)
        _same(res[i].tci, alone, f"request {i}")
