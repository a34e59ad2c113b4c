"""Host-side logic that needs no GPU: the capacity plan of a problem, the
serving loop's handling of device failures (with a stubbed library), and the
functor checks of the Pi-evaluation entry point (which run before any device
call)."""
import ctypes as C

import numpy as np
import pytest

from conftest import load_pkg


def _holder(pkg, ld, **kw):
    return pkg._abi.ProblemHolder(ld, pkg._abi.F_LORENTZ_INT, [1.0], **kw)


def test_capacity_plan_defaults():
    pkg = load_pkg()
    # 8 sites of d = 10: rank bound 10^4, default capacity 32, limit 10^4
    assert pkg.capacity_plan(_holder(pkg, [10] * 8)) == (32, 10 ** 4)
    # maxbonddim below 64 caps both
    assert pkg.capacity_plan(_holder(pkg, [10] * 8, maxbonddim=20)) == (20, 20)
    # a small domain caps both at its rank bound: 2^2 = 4 at the middle bond
    assert pkg.capacity_plan(_holder(pkg, [2] * 4)) == (4, 4)
    assert pkg.capacity_plan(_holder(pkg, [3, 5])) == (3, 3)
    # an explicit chi_cap is used as given
    assert pkg.capacity_plan(_holder(pkg, [10] * 8, chi_cap=7)) == (7, 10 ** 4)
    # saturating products for long chains
    used, limit = pkg.capacity_plan(_holder(pkg, [16] * 40, maxbonddim=256))
    assert (used, limit) == (32, 256)
    # d = 64: the Pi buffer of (65·(chi+5))² doubles is kept within 8 MiB
    used, limit = pkg.capacity_plan(_holder(pkg, [64] * 6))
    assert used == 10 and (65 * 15) ** 2 * 8 <= 8 << 20 < (65 * 16) ** 2 * 8
    # d = 128: never below 8
    assert pkg.capacity_plan(_holder(pkg, [128] * 6))[0] == 8


def test_capacity_plan_rejects_bad_problems():
    pkg = load_pkg()
    h = _holder(pkg, [10] * 8, chi_cap=-1)
    with pytest.raises(pkg.ArgumentError):
        pkg.capacity_plan(h)


class _StubLib:
    """Stands in for libtci_hip.so: submits succeed, queries answer from a
    script of statuses."""

    def __init__(self, query_codes, wait_code=0):
        self.query_codes = list(query_codes)
        self.wait_code = wait_code
        self.freed = []
        self.nsub = 0

    def tci_ctx_set_streams(self, h, n):
        return 0

    def tci_solve_submit(self, h, cnt, arr, out):
        self.nsub += 1
        out._obj.value = 1000 + self.nsub
        return 0

    def tci_batch_query(self, b):
        return self.query_codes.pop(0) if self.query_codes else 0

    def tci_batch_wait(self, b):
        return self.wait_code

    def tci_batch_result(self, b, k, r):
        r._obj.status = 0
        r._obj.niter = 0
        return 0

    def tci_batch_history(self, *a):
        return 0

    def tci_batch_free(self, b):
        self.freed.append(b.value)
        return 0

    def tci_last_error(self):
        return b"stub device failure"


class _Ctx:
    handle = C.c_void_p(1)


def _serve(monkeypatch, stub, n=3):
    pkg = load_pkg()
    serving = __import__("tci_amd.serving", fromlist=["TCIServer"])
    monkeypatch.setattr(pkg._abi, "lib", lambda: stub)
    srv = serving.TCIServer(_Ctx(), streams=1)
    reqs = [(pkg.Lorentz(1.0), [4] * 4, None, {})] * n
    return serving, srv.run(reqs, np.zeros(n))


def test_serving_raises_on_device_error(monkeypatch):
    # pending, then TCI_ERR_DEVICE: the loop must stop, free the batch and raise
    stub = _StubLib([13, 12])
    serving = __import__("tci_amd.serving", fromlist=["ServeError"])
    with pytest.raises(serving.ServeError) as e:
        _serve(monkeypatch, stub)
    assert e.value.status == 12
    assert stub.freed == [1001]


def test_serving_raises_when_report_fetch_fails(monkeypatch):
    stub = _StubLib([0], wait_code=12)
    serving = __import__("tci_amd.serving", fromlist=["ServeError"])
    with pytest.raises(serving.ServeError):
        _serve(monkeypatch, stub)
    assert stub.freed == [1001]


def test_serving_completes_with_pending_then_ok(monkeypatch):
    stub = _StubLib([13, 13, 0])
    _, (results, stats) = _serve(monkeypatch, stub)
    assert stats.n == 3 and all(r.status == 0 for r in results)
    assert stub.freed == [1001]


def test_pi_eval_checks_functor_for_one_site():
    # L = 1 used to skip the functor checks; a too-short parameter block must
    # be refused before any device work (TCI_ERR_ARG), even with no GPU
    pkg = load_pkg()
    lib = pkg._abi.lib()
    ld = np.array([4], np.int32)
    f = pkg._abi.tci_functor(pkg._abi.F_LORENTZ_WEIGHTED, 1, pkg._abi.dptr(np.array([16.0])), 0, None)
    out = np.zeros(4)
    st = lib.tci_pi_eval_f64(C.c_void_p(1), C.byref(f), 1, pkg._abi.i32ptr(ld), None, 1, 0, None, 1, 0,
                             pkg._abi.dptr(out), None)
    assert st == 3
    assert b"weighted" in lib.tci_last_error()
    # an unknown kind as well
    f2 = pkg._abi.tci_functor(99, 0, None, 0, None)
    assert lib.tci_pi_eval_f64(C.c_void_p(1), C.byref(f2), 1, pkg._abi.i32ptr(ld), None, 1, 0, None, 1, 0,
                               pkg._abi.dptr(out), None) == 3


class _StubServerLib:
    """Stands in for the slot server of libtci_hip.so: a ring of `ring`
    entries, requests finishing `per_poll` at a time in submission order;
    result statuses from a script keyed by submission number."""

    def __init__(self, ring=2, per_poll=1, statuses=None, poll_code=0, room_limit=None):
        self.ring, self.per_poll = ring, per_poll
        self.statuses = dict(statuses or {})
        self.poll_code = poll_code
        self.room_limit = room_limit
        self.inflight = []          # handles submitted, not yet polled
        self.unfreed = set()
        self.freed = []
        self.nsub = 0
        self.status_of = {}
        self.chi_caps = []
        self.bursts = []

    def tci_server_start(self, ctx, shape, slots, idle, team, out):
        out._obj.value = 77
        return 0

    def tci_server_stop(self, h):
        return 0

    def tci_server_submit_many(self, h, count, arr, out, acc):
        self.bursts.append(count)
        room = self.ring - len(self.unfreed)
        if self.room_limit is not None:
            room = min(room, self.room_limit)
        k = max(0, min(count, room))
        for q in range(k):
            self.nsub += 1
            hv = 5000 + self.nsub
            out[q] = hv
            self.inflight.append(hv)
            self.unfreed.add(hv)
            self.status_of[hv] = self.statuses.get(self.nsub, 0)
            self.chi_caps.append(arr[q].opts.chi_cap)
        acc._obj.value = k
        return 0

    def tci_server_poll(self, h, out, cap, n):
        if self.poll_code:
            return self.poll_code
        k = min(self.per_poll, len(self.inflight), cap)
        for q in range(k):
            out[q] = self.inflight.pop(0)
        n._obj.value = k
        return 0

    def tci_batch_result(self, b, i, r):
        r._obj.status = self.status_of[b.value]
        r._obj.niter = 0
        return 0

    def tci_batch_history(self, *a):
        return 0

    def tci_capacity_plan(self, p, used, limit):
        used._obj.value = 32
        limit._obj.value = 1000
        return 0

    def tci_batch_free(self, b):
        self.freed.append(b.value)
        self.unfreed.discard(b.value)
        return 0

    def tci_last_error(self):
        return b"stub failure"


def _slot(monkeypatch, stub, n=5):
    pkg = load_pkg()
    serving = __import__("tci_amd.serving", fromlist=["SlotServer"])
    monkeypatch.setattr(pkg._abi, "lib", lambda: stub)
    shape = (pkg.Lorentz(1.0), [4] * 4, None, {})
    srv = serving.SlotServer(_Ctx(), shape, slots=stub.ring, idle_ms=5)
    reqs = [(pkg.Lorentz(1.0), [4] * 4, None, {}) for _ in range(n)]
    return serving, srv, reqs


def test_slot_serving_bursts_fill_the_ring_and_free_each_once(monkeypatch):
    stub = _StubServerLib(ring=2, per_poll=1)
    serving, srv, reqs = _slot(monkeypatch, stub)
    results, stats = srv.run(reqs, np.zeros(len(reqs)), keep_history=False)
    srv.close()   # while the stub stands in for the library
    assert stats.n == 5 and all(r is not None and r.status == 0 for r in results)
    assert max(stub.bursts) <= 2                       # bursts never exceed the ring's free entries
    assert sorted(stub.freed) == sorted(set(stub.freed)) and len(stub.freed) == 5
    assert not stub.unfreed and stub.nsub == 5


def test_slot_serving_resubmits_a_capacity_failure_at_twice_the_capacity(monkeypatch):
    # the second submission ends with TCI_ERR_CAPACITY (10): resubmitted with chi_cap 64
    stub = _StubServerLib(ring=4, per_poll=2, statuses={2: 10})
    serving, srv, reqs = _slot(monkeypatch, stub, n=3)
    results, stats = srv.run(reqs, np.zeros(3))
    srv.close()
    assert stats.retries == 1 and stub.nsub == 4
    assert all(r.status == 0 for r in results)
    assert stub.chi_caps[:3] == [0, 0, 0] and stub.chi_caps[3] == 64
    assert not stub.unfreed


def test_slot_serving_poll_error_frees_everything_and_raises(monkeypatch):
    stub = _StubServerLib(ring=4, poll_code=12)
    serving, srv, reqs = _slot(monkeypatch, stub, n=3)
    with pytest.raises(serving.ServeError) as e:
        srv.run(reqs, np.zeros(3))
    srv.close()
    assert e.value.status == 12
    assert not stub.unfreed and len(stub.freed) == 3


def test_slot_serving_refuses_a_request_that_never_fits(monkeypatch):
    # nothing in flight and still no room: raise instead of waiting forever
    stub = _StubServerLib(ring=4, room_limit=0)
    serving, srv, reqs = _slot(monkeypatch, stub, n=2)
    with pytest.raises(serving.ServeError) as e:
        srv.run(reqs, np.zeros(2))
    srv.close()
    assert e.value.status == 11
