"""Served TCI: continuous batching of independent crossinterpolate2 requests.

Requests arrive over time (an arrival time per request).  The loop admits
every request that has arrived into a new batch as soon as one of the
context's streams is free and the HBM budget allows, so requests that arrive
while other batches are on the GPU start without waiting for them; a batch's
solves run one per workgroup and each solve keeps its pivot sets, Pi buffers
and site tensors resident on the device until it finishes.  Latency is
measured from a request's arrival to its results being on the host.

The reference runs one solve per call (tensorci2.jl:943-953); this module has
no counterpart there.  A solve's result does not depend on what else is in its
batch or on the GPU at the same time (each solve is one workgroup with its own
state), which tests/test_gpu_serving.py checks.
"""
import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _abi


PENDING = 13          # TCI_ERR_PENDING
TCI_ERR_DEVICE = 12
TCI_ERR_CAPACITY = 10


class ServeError(RuntimeError):
    """A batch failed on the device (tci_status in .status)."""

    def __init__(self, status, msg=""):
        super().__init__(f"status {status}: {msg}")
        self.status = status


@dataclass
class ServedResult:
    status: int
    ranks: np.ndarray
    errors: np.ndarray
    maxsamplevalue: float
    arrival: float
    done: float
    tci: object = None          # TensorCI2 when keep_tci

    @property
    def latency(self):
        return self.done - self.arrival


@dataclass
class ServeStats:
    n: int
    wall_s: float
    solves_per_s: float
    p50_ms: float
    p99_ms: float
    batches: int
    mean_batch: float
    oom_waits: int = 0
    retries: int = 0
    latencies_ms: list = field(default_factory=list)
    host_ms: dict = field(default_factory=dict)   # host time per phase of the loop


class TCIServer:
    """Continuous-batching front end over one context.

    streams: batches that may run on the GPU at the same time.
    max_batch: most requests admitted into one batch."""

    def __init__(self, ctx, streams=4, max_batch=4096, idle_sleep_s=20e-6):
        self.ctx = ctx
        self.streams = int(streams)
        self.max_batch = int(max_batch)
        self.idle_sleep_s = idle_sleep_s
        st = _abi.lib().tci_ctx_set_streams(ctx.handle, self.streams)
        if st:
            raise RuntimeError(_abi.lib().tci_last_error().decode())

    def run(self, requests, arrivals, keep_tci=False):
        """requests[i] = (f, localdims, initialpivots_or_None, kwargs); arrivals
        sorted, seconds from the start.  Returns (results, ServeStats).

        A request whose solve outgrew its capacity (TCI_ERR_CAPACITY) while
        the caller left chi_cap unset is admitted again, ahead of new
        arrivals, at twice the capacity (as crossinterpolate2_batch does);
        its latency counts both runs."""
        from . import TensorCI2, _holder, capacity_plan
        lib = _abi.lib()
        n = len(requests)
        holders = [_holder(f, ld, piv, kw) for f, ld, piv, kw in requests]
        results = [None] * n
        inflight = []                       # (batch handle, [request ids], arrays kept alive)
        retry = []                          # request ids to admit again (larger capacity)
        nxt = 0
        nbatches = 0
        nretries = 0
        oom_waits = 0
        t0 = time.perf_counter()
        while nxt < n or inflight or retry:
            now = time.perf_counter() - t0
            progressed = False
            # admit the retries and everything that has arrived, when a stream is free
            ready = nxt < n and arrivals[nxt] <= now
            if (retry or ready) and len(inflight) < self.streams:
                ids = retry[:self.max_batch]
                hi = nxt
                while hi < n and arrivals[hi] <= now and len(ids) + (hi - nxt) < self.max_batch:
                    hi += 1
                ids = ids + list(range(nxt, hi))
                cnt = len(ids)
                while cnt > 0:
                    arr = (_abi.tci_problem * cnt)(*[holders[i].st for i in ids[:cnt]])
                    b = C.c_void_p()
                    st = lib.tci_solve_submit(self.ctx.handle, cnt, arr, C.byref(b))
                    if st == 0:
                        taken = ids[:cnt]
                        inflight.append((b, taken, arr))
                        nr = min(len(retry), cnt)
                        retry = retry[nr:]
                        nxt += cnt - nr
                        nbatches += 1
                        progressed = True
                        break
                    if st == 11 and inflight:       # HBM budget: wait for a batch to finish
                        oom_waits += 1
                        break
                    if st == 11 and cnt > 1:        # nothing running: admit fewer
                        cnt //= 2
                        continue
                    msg = lib.tci_last_error().decode()
                    for b2, _, _ in inflight:
                        lib.tci_batch_free(b2)
                    raise ServeError(st, f"submit of {cnt} requests failed: {msg}")
            # retire finished batches; only TCI_ERR_PENDING means "still
            # running": any other failure of the query or of the report
            # fetch frees the batch and raises
            still = []
            for pos, (b, ids, arr) in enumerate(inflight):
                q = lib.tci_batch_query(b)
                if q == PENDING:
                    still.append((b, ids, arr))
                    continue
                w = lib.tci_batch_wait(b) if q == 0 else q   # reports on the host
                if w == 0:
                    r = _abi.tci_result()
                    for k in range(len(ids)):
                        if lib.tci_batch_result(b, k, C.byref(r)) != 0:
                            w = TCI_ERR_DEVICE
                            break
                if w != 0:
                    msg = lib.tci_last_error().decode()
                    # this batch, the ones still running and the ones not yet visited
                    for b2, _, _ in [(b, ids, arr)] + still + inflight[pos + 1:]:
                        lib.tci_batch_free(b2)
                    raise ServeError(w, f"batch of requests {ids[0]}..{ids[-1]} failed: {msg}")
                done = time.perf_counter() - t0
                for k, i in enumerate(ids):
                    lib.tci_batch_result(b, k, C.byref(r))
                    if r.status == TCI_ERR_CAPACITY and not requests[i][3].get("chi_cap"):
                        used, limit = capacity_plan(holders[i])
                        chi = int(min(2 * used, limit, 8192))
                        if chi > used:
                            f, ld, piv, kw = requests[i]
                            holders[i] = _holder(f, ld, piv, dict(kw, chi_cap=chi))
                            retry.append(i)
                            nretries += 1
                            continue
                    m = max(r.niter, 1)
                    ranks = np.zeros(m, np.int64)
                    errs = np.zeros(m)
                    lib.tci_batch_history(b, k, _abi.i64ptr(ranks), _abi.dptr(errs), None)
                    t = TensorCI2(b, k, requests[i][1]) if keep_tci else None
                    results[i] = ServedResult(r.status, ranks[:r.niter], errs[:r.niter],
                                              r.maxsamplevalue, float(arrivals[i]), done, t)
                lib.tci_batch_free(b)
                progressed = True
            inflight = still
            if not progressed:
                if nxt < n and not inflight and not retry:
                    wait = arrivals[nxt] - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 1e-3))
                else:
                    time.sleep(self.idle_sleep_s)
        wall = time.perf_counter() - t0
        lat = np.array([r.latency for r in results]) * 1e3
        stats = ServeStats(n=n, wall_s=wall, solves_per_s=n / wall if wall > 0 else 0.0,
                           p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                           batches=nbatches, mean_batch=n / max(nbatches, 1), oom_waits=oom_waits,
                           retries=nretries,
                           latencies_ms=lat.tolist())
        return results, stats


class SlotServer:
    """Slot-level serving (tci_server_*): a persistent grid on the device takes
    each request from a queue as soon as a workgroup is free, so a request
    waits neither for a batch to form nor for a batch's slowest solve.

    shape: the largest request expected, as (f, localdims, initialpivots,
    kwargs) — it sizes the grid's on-chip plan.  slots: the queue's ring
    (requests in flight at most).  idle_ms: the grid leaves after this long
    without a request (and starts again with the next one).  team: threads
    per solve, 256 (two solves per CU: the most in flight) or 512 (one per
    CU, each solve faster: for streams with fewer requests in flight than
    CUs); 0 takes the context's default."""

    def __init__(self, ctx, shape, slots=16384, idle_ms=50, team=0):
        from . import _holder
        self.ctx = ctx
        h = _holder(*shape)
        self.slots = int(slots)
        self.h = C.c_void_p()
        st = _abi.lib().tci_server_start(ctx.handle, C.byref(h.st), int(slots), int(idle_ms), int(team),
                                         C.byref(self.h))
        if st:
            raise ServeError(st, _abi.lib().tci_last_error().decode())

    def close(self):
        if self.h:
            st = _abi.lib().tci_server_stop(self.h)
            self.h = C.c_void_p()
            if st:
                raise ServeError(st, _abi.lib().tci_last_error().decode())

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, requests, arrivals, keep_tci=False, keep_history=True):
        """requests[i] = (f, localdims, initialpivots_or_None, kwargs); arrivals
        sorted, seconds from the start.  The requests that have arrived are
        submitted together (tci_server_submit_many); a full queue or an
        exhausted HBM budget holds the rest back until results are retired.
        A request that outgrew its capacity with chi_cap unset is submitted
        again at twice the capacity.  keep_history=False skips the per-request
        rank/error history.  Returns (results, ServeStats)."""
        from . import TensorCI2, _holder, capacity_plan
        lib = _abi.lib()
        n = len(requests)
        results = [None] * n
        inflight = {}                       # batch handle value -> (request id, batch handle, holder)
        retry = []
        retry_holder = {}                   # request id -> holder at the larger capacity
        hcache = {}                         # id(request) -> holder: repeated requests share one
        nxt = 0
        nretries = 0
        oom_waits = 0
        pcap = 4096
        polled = (C.c_void_p * pcap)()
        npolled = C.c_int32()
        accepted = C.c_int32()
        host = {"submit": 0.0, "poll": 0.0, "retire": 0.0}
        clock = time.perf_counter
        t0 = time.perf_counter()

        def fail_all(st, msg):
            for _, b2, _ in inflight.values():
                lib.tci_batch_free(b2)
            raise ServeError(st, msg)

        while nxt < n or inflight or retry:
            now = time.perf_counter() - t0
            progressed = False
            ts = clock()
            ids = list(retry)
            hs = [retry_holder[i] for i in retry]
            j = nxt
            room = min(4096, self.slots - len(inflight))   # the ring's free entries at most
            while j < n and arrivals[j] <= now and len(ids) < room:
                key = id(requests[j])
                h = hcache.get(key)
                if h is None:
                    h = hcache[key] = _holder(*requests[j])
                ids.append(j)
                hs.append(h)
                j += 1
            if ids:
                arr = (_abi.tci_problem * len(ids))(*[h.st for h in hs])
                outb = (C.c_void_p * len(ids))()
                st = lib.tci_server_submit_many(self.h, len(ids), arr, outb, C.byref(accepted))
                if st:
                    fail_all(st, f"submit failed: {lib.tci_last_error().decode()}")
                acc = accepted.value
                for q in range(acc):
                    inflight[outb[q]] = (ids[q], C.c_void_p(outb[q]), hs[q])
                taken = min(acc, len(retry))
                for i in retry[:taken]:
                    del retry_holder[i]
                retry = retry[taken:]
                nxt += max(0, acc - taken)
                progressed = acc > 0
                if acc < len(ids):
                    if not inflight:
                        fail_all(11, f"request {ids[acc]} does not fit the queue or the HBM budget alone")
                    oom_waits += 1
            tp = clock()
            host["submit"] += tp - ts
            st = lib.tci_server_poll(self.h, polled, pcap, C.byref(npolled))
            tr = clock()
            host["poll"] += tr - tp
            if st:
                fail_all(st, f"poll failed: {lib.tci_last_error().decode()}")
            done = time.perf_counter() - t0
            for q in range(npolled.value):
                i, b, h = inflight.pop(polled[q])
                progressed = True
                r = _abi.tci_result()
                lib.tci_batch_result(b, 0, C.byref(r))
                if r.status == TCI_ERR_CAPACITY and not requests[i][3].get("chi_cap"):
                    used, limit = capacity_plan(h)
                    chi = int(min(2 * used, limit, 8192))
                    if chi > used:
                        f, ld, piv, kw = requests[i]
                        retry_holder[i] = _holder(f, ld, piv, dict(kw, chi_cap=chi))
                        retry.append(i)
                        nretries += 1
                        lib.tci_batch_free(b)
                        continue
                if keep_history:
                    m = max(r.niter, 1)
                    ranks = np.zeros(m, np.int64)
                    errs = np.zeros(m)
                    lib.tci_batch_history(b, 0, _abi.i64ptr(ranks), _abi.dptr(errs), None)
                    ranks, errs = ranks[:r.niter], errs[:r.niter]
                else:
                    ranks, errs = None, None
                t = TensorCI2(b, 0, requests[i][1]) if keep_tci else None
                results[i] = ServedResult(r.status, ranks, errs, r.maxsamplevalue, float(arrivals[i]), done, t)
                lib.tci_batch_free(b)
            host["retire"] += clock() - tr
            if not progressed:
                if nxt < n and not inflight and not retry:
                    wait = arrivals[nxt] - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 1e-3))
                else:
                    time.sleep(20e-6)
        wall = time.perf_counter() - t0
        lat = np.array([r.latency for r in results]) * 1e3
        stats = ServeStats(n=n, wall_s=wall, solves_per_s=n / wall if wall > 0 else 0.0,
                           p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                           batches=n, mean_batch=1.0, oom_waits=oom_waits, retries=nretries,
                           latencies_ms=lat.tolist(), host_ms={k: v * 1e3 for k, v in host.items()})
        return results, stats


def poisson_arrivals(n, rate_per_s, seed=5):
    """n arrival times of a Poisson stream (exponential gaps), seconds."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, n))
