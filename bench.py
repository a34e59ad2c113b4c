"""Benchmark: TCI solves/s and p50/p99 latency at tolerance 1e-8, batched and
served vs plain, on one MI355X.

Headline workload: crossinterpolate2(Float64, f, fill(10, 8); tolerance=1e-8)
with f(v) = 1/(1 + v·v) and the initial pivot ones(8) (the README example,
README.md:24-28), every optimize! default kept (maxiter 20, back-and-forth
sweeps, full pivot search, global pivot search with 5 points).  A "step" is
one batched launch of B such solves, each a complete solve including the
final sweep1site!, results returned to the host.  The plain path is the same
solve submitted alone (one solve per launch).

The same JSON line carries, as sub-objects, the other workloads at their full
sizes: the batched standalone rrLU (4096 uneven matrices in one launch, with
a seeded sample of pivot sequences checked against the oracle outside the
timed region), the quantics (L=30, d=2) and high-dimensional (L=40, d=16)
solves, each with its error max|TT - f| on 10^5 seeded random points, and the
served Gaussian-sum stream (10 000 requests, Poisson arrivals, rate swept).
The CPU baseline beside each is the oracle (the C restatement of the
reference's CPU path) on this host's cores.

Prints ONE JSON line.  --gpus > 1 is refused: this build runs on one GPU.
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import statistics
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "tensorcrossinterpolation.jl_amd")
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def load_pkg():
    spec = importlib.util.spec_from_file_location("tci_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["tci_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    return O


def _problems():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import problems
    return problems


# ------------------------------------------------------------------ host CPU
def host_cpu():
    """CPU model, nproc, the CPUs this process may run on and the cgroup CPU
    quota; `threads` = the CPUs usable at once (affinity, capped by the quota)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = nproc
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    threads = aff if quota is None else max(1, min(aff, int(quota)))
    return {"model": model, "nproc": nproc, "affinity": aff, "cgroup_quota_cpus": quota, "threads": threads}


def cpu_throughput(fn, threads, seconds):
    """fn() run repeatedly on `threads` threads (the oracle releases the GIL
    in its C calls) for about `seconds`; returns (calls/s, calls, wall s)."""
    counts = [0] * threads
    stop = time.perf_counter() + seconds

    def work(k):
        while time.perf_counter() < stop:
            fn()
            counts[k] += 1

    t0 = time.perf_counter()
    th = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    dt = time.perf_counter() - t0
    return sum(counts) / dt, sum(counts), dt


def cpu_latency(fn, reps):
    fn()   # warm-up
    lat = []
    for _ in range(reps):
        s = time.perf_counter()
        fn()
        lat.append(time.perf_counter() - s)
    return statistics.median(lat)


# ------------------------------------------------------------------ solves
def run_batch(tci, ctx, arr, count):
    """One launch of `count` solves; returns (wall s, kernel ms, ranks of
    solve 0, [workstats of solve 0 + phase profile])."""
    lib = tci._abi.lib()
    b = C.c_void_p()
    t0 = time.perf_counter()
    st = lib.tci_solve_submit(ctx.handle, count, arr, C.byref(b))
    if st:
        raise RuntimeError(lib.tci_last_error().decode())
    st = lib.tci_batch_wait(b)
    t1 = time.perf_counter()
    if st:
        raise RuntimeError(lib.tci_last_error().decode())
    ms = C.c_float()
    lib.tci_batch_elapsed_ms(b, C.byref(ms))
    r = tci._abi.tci_result()
    bad = 0
    for i in (0, count // 2, count - 1):
        lib.tci_batch_result(b, i, C.byref(r))
        bad += r.status != 0
    lib.tci_batch_result(b, 0, C.byref(r))
    ranks = np.zeros(max(r.niter, 1), np.int64)
    lib.tci_batch_history(b, 0, tci._abi.i64ptr(ranks), None, None)
    ws = list(r.workstats)
    prof = np.zeros(8)
    lib.tci_batch_profile(b, 0, tci._abi.dptr(prof))
    ws.append(prof.tolist())
    lib.tci_batch_free(b)
    if bad:
        raise RuntimeError("a solve failed")
    return t1 - t0, ms.value, ranks[:r.niter].tolist(), ws


def profiled(name, key):
    """A figure recorded in profiles/<name>.md as a `key: value` line, or None."""
    path = os.path.join(ROOT, "profiles", name + ".md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        line = line.strip()
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def sample_error(tci, spec, npts=100_000, seed=12345):
    """max|TT - f| of one device solve on npts seeded random points (f by
    the device functor, whose bits equal the host FP64 evaluation), divided
    by max|f| over the points and by the solve's maxsamplevalue."""
    name, kind, params, ip, ld, piv, opts = spec
    f = tci.DeviceFunctor(kind, np.asarray(params, float), np.asarray(ip, np.int32))
    (t,) = tci.crossinterpolate2_batch([(f, ld, piv, dict(opts))], error_sample=(npts, seed))
    me, mf = t.sample_error
    return {"points": npts, "seed": seed, "max_abs_err": me, "err_over_max_f": me / mf,
            "err_over_maxsample": me / t.maxsamplevalue, "tolerance": opts["tolerance"],
            "meets_tolerance": bool(me / t.maxsamplevalue <= opts["tolerance"]),
            "ranks": list(map(int, t.ranks))}


def true_error_search(tci, ctx, spec, cpu_seconds, nsearch=2000, seed=7):
    """estimatetrueerror (globalsearch.jl:18-41) on one solve's train: the
    floating-zone search from nsearch seeded random start points on the
    device (one workgroup per start), the largest |f - TT| it finds, and
    the oracle's _floatingzone on a bounded sample of the same starts (its
    pivots and errors compared with the device's)."""
    name, kind, params, ip, ld, piv, opts = spec
    f = tci.DeviceFunctor(kind, np.asarray(params, float), np.asarray(ip, np.int32))
    (t,) = tci.crossinterpolate2_batch([(f, ld, piv, dict(opts))], ctx=ctx)
    tt = tci.TensorTrain(t)
    starts = tci.sample_points(ld, nsearch, seed)
    tci.floatingzone(tt, f, starts[:8], ctx=ctx)   # warm-up
    t0 = time.perf_counter()
    pv, errs = tci.floatingzone(tt, f, starts, ctx=ctx)
    dt = time.perf_counter() - t0
    worst = float(np.nanmax(errs))
    out = {"nsearch": nsearch, "seed": seed, "ms": round(dt * 1e3, 3), "starts_per_s": round(nsearch / dt, 1),
           "max_abs_err_found": worst, "over_maxsample": worst / t.maxsamplevalue,
           "tolerance": opts["tolerance"]}
    if cpu_seconds > 0:
        O = _oracle()
        of = O.Functor(kind, params, ip)
        cores = [np.asarray(c) for c in t.sitetensors]
        k, t1 = 0, time.perf_counter()
        same = True
        while k < nsearch and time.perf_counter() - t1 < cpu_seconds:
            op, oe = O.floatingzone(of, cores, starts[k:k + 1])
            same = same and np.array_equal(op[0], pv[k]) and oe[0] == errs[k]
            k += 1
        ct = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": round(k / ct, 2), "unit": "starts/s", "cores": 1, "kind": "port",
                               "sample": f"the first {k} starts in {ct:.1f} s on one thread"}
        out["oracle_equal_on_sample"] = bool(same)
    return out


def tt_post_section(tci, ctx):
    """compress!(tt, :LU) and TensorCI2(tt) (tensortrain.jl:147-183,
    conversion.jl:158-176) on the device for two trains: the C4 solve's
    (L=40, d=16, bond dimensions up to 16) and a random train with L=20,
    d=16, bond dimension 128.  Each result is compared with the oracle's on
    the same input (cores bitwise for compress!, pivot sets for the
    conversion), and the oracle is timed beside it on one thread."""
    P = _problems()
    O = _oracle()
    name, kind, params, ip, ld, piv, opts = P.c4_spec(5)
    f = tci.DeviceFunctor(kind, np.asarray(params, float), np.asarray(ip, np.int32))
    (t,) = tci.crossinterpolate2_batch([(f, ld, piv, dict(opts))], ctx=ctx)
    rng = np.random.default_rng(11)
    chi, L, d = 128, 20, 16
    dims = [1] + [chi] * (L - 1) + [1]
    rand = [rng.standard_normal((dims[b], d, dims[b + 1])) / np.sqrt(dims[b] * d) for b in range(L)]
    trains = {"c4_solve_train": [np.asarray(c) for c in t.sitetensors], "random_L20_d16_chi128": rand}
    out = {}
    for tname, cores in trains.items():
        row = {"linkdims": [int(c.shape[2]) for c in cores[:-1]]}
        # compress!(:LU) with a truncating tolerance
        tt = tci.TensorTrain(cores)
        tci.compress(tci.TensorTrain(cores), "LU", tolerance=1e-8, ctx=ctx)   # warm-up
        t0 = time.perf_counter()
        tci.compress(tt, "LU", tolerance=1e-8, ctx=ctx)
        row["compress_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
        row["compressed_linkdims_max"] = int(max(tt.linkdims()))
        t0 = time.perf_counter()
        oc = O.tt_compress([np.array(c, order="F") for c in cores], "LU", tolerance=1e-8)
        row["compress_cpu_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
        row["compress_equal_oracle"] = bool(all(np.array_equal(a, b) for a, b in zip(tt.sitetensors, oc)))
        # TensorCI2(tt)
        t0 = time.perf_counter()
        conv = tci.tensorci2_from_tt(tci.TensorTrain(cores), tolerance=1e-12, ctx=ctx)
        row["to_tensorci2_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
        row["to_tensorci2_linkdims_max"] = int(max(len(s_) for s_ in conv.Iset[1:])) if hasattr(conv, "Iset") else None
        out[tname] = row
    out["note"] = ("host-driven per-bond sweeps: one factorization and one product per bond, each a device call; "
                   "the oracle runs the same steps on one host thread")
    return out


def tune_team(tci, ctx, holder, batch):
    """The solve team for batches of this workload of at least `batch`
    solves (tci_ctx_autotune_solve: each team timed on such a batch, outside
    the timed region), used until the table is reset."""
    lib = tci._abi.lib()
    ms = (C.c_float * 4)()
    team = C.c_int32()
    if lib.tci_ctx_autotune_solve(ctx.handle, C.byref(holder.st), int(batch), ms, C.byref(team)):
        raise RuntimeError(lib.tci_last_error().decode())
    return {"team": team.value, "kernel_ms_by_team": {str(t): round(ms[q], 3) for q, t in enumerate((64, 256, 512, 1024))}}


def solve_config(tci, ctx, spec, batch, steps, plain_reps, cpu, cpu_seconds, cpu_reps):
    """Batched and plain GPU rates of one solve configuration, its sample
    error, and the oracle's single-thread latency and multi-thread rate."""
    name, kind, params, ip, ld, piv, opts = spec
    holder = tci._abi.ProblemHolder(ld, kind, params, ip, piv, **opts)
    arr = (tci._abi.tci_problem * batch)(*([holder.st] * batch))
    one = (tci._abi.tci_problem * 1)(holder.st)
    tuned = tune_team(tci, ctx, holder, batch)
    tuned_plain = tune_team(tci, ctx, holder, 1)
    run_batch(tci, ctx, one, 1)
    plain = [run_batch(tci, ctx, one, 1)[0] for _ in range(plain_reps)]
    run_batch(tci, ctx, arr, batch)
    walls, kms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        w, k, ranks, ws = run_batch(tci, ctx, arr, batch)
        walls.append(w)
        kms.append(k)
    elapsed = time.perf_counter() - t0
    value = batch * steps / elapsed
    out = {"workload": name, "batch_solves": batch, "solves_per_s": round(value, 2),
           "ms_per_launch": round(elapsed / steps * 1e3, 3), "kernel_ms": round(float(np.mean(kms)), 3),
           "plain": {"solves_per_s": round(1.0 / statistics.median(plain), 2),
                     "latency_ms_p50": round(statistics.median(plain) * 1e3, 3)},
           "batched_vs_plain": round(value * statistics.median(plain), 1), "ranks": ranks,
           "solve_team": tuned, "solve_team_plain": tuned_plain,
           "solve_phases_us": dict(zip(["eval", "rrlu", "sets", "fill", "gsearch", "sweep1", "total", "pinv"],
                                       [round(v, 1) for v in ws[6][:8]])),
           "error": sample_error(tci, spec),
           "true_error_search": true_error_search(tci, ctx, spec, min(cpu_seconds, 5.0))}
    ctx.set_solve_tuning([])
    if cpu_seconds > 0:
        O = _oracle()
        of = O.Functor(kind, params, ip)
        kw = {k: v for k, v in opts.items() if k != "chi_cap"}

        def one_solve():
            O.crossinterpolate2(of, ld, piv, **kw)

        lat = cpu_latency(one_solve, cpu_reps)
        cps, n, dt = cpu_throughput(one_solve, cpu["threads"], cpu_seconds)
        out["cpu_baseline"] = {"value": round(cps, 3), "unit": "solves/s", "cores": cpu["threads"], "kind": "port",
                               "sample": f"{n} solves of the same problem in {dt:.1f} s on {cpu['threads']} threads",
                               "latency_ms_1thread": round(lat * 1e3, 3)}
        out["vs_cpu"] = round(value / cps, 1)
    return out


# ------------------------------------------------------------------ rrLU
def c2_shapes(count, seed=2024):
    """Shapes of the batched-rrLU workload: rows, cols ~ U{32..1024}, rank ~ U{8..128}."""
    rng = np.random.default_rng(seed)
    return rng.integers(32, 1025, count), rng.integers(32, 1025, count), rng.integers(8, 129, count)


def c2_matrices_device(count, seed=2024):
    """A_i = X·Y + 1e-12·E with N(0,1) factors (seed 1000+i), FP64, made on the
    device and stored column-major back to back.  Returns (A, offsets, shapes)."""
    import torch
    ms, ns, rs = c2_shapes(count, seed)
    offs, o = [], 0
    for m, n in zip(ms, ns):
        offs.append(o)
        o += int(m) * int(n)
    A = torch.empty(o, dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda")
    for i, (m, n, r) in enumerate(zip(ms, ns, rs)):
        g.manual_seed(1000 + i)
        X = torch.randn(int(m), int(r), dtype=torch.float64, device="cuda", generator=g)
        Y = torch.randn(int(r), int(n), dtype=torch.float64, device="cuda", generator=g)
        E = torch.randn(int(m), int(n), dtype=torch.float64, device="cuda", generator=g)
        M = torch.addmm(E * 1e-12, X, Y)
        A[offs[i]:offs[i] + int(m) * int(n)] = M.t().contiguous().view(-1)   # column-major
    torch.cuda.synchronize()
    return A, offs, [(int(m), int(n)) for m, n in zip(ms, ns)]


def rrlu_section(args, tci, ctx, cpu):
    """Batched standalone rrLU: every matrix in ONE launch, full pivoting to
    relative tolerance 1e-10, matrices resident in HBM (a pristine copy
    restores them, untimed, before each step).  Afterwards a seeded sample of
    the matrices is factored by the oracle and its pivot sequences compared."""
    import torch
    cnt = args.rrlu_count
    A, offs, shapes = c2_matrices_device(cnt)
    pristine = A.clone()
    lib = tci._abi.lib()
    perm = torch.zeros(sum(m + n for m, n in shapes), dtype=torch.int32, device="cuda")
    jobs = (tci._abi.tci_lu_job * cnt)()
    po = 0
    poff = []
    for i, (m, n) in enumerate(shapes):
        j = jobs[i]
        j.A = A.data_ptr() + 8 * offs[i]
        j.rowperm = perm.data_ptr() + 4 * po
        j.colperm = perm.data_ptr() + 4 * (po + m)
        j.m, j.n, j.maxrank = m, n, tci.TCI_UNLIMITED
        j.reltol, j.abstol, j.leftorth = 1e-10, 0.0, 1
        poff.append(po)
        po += m + n
    jb = torch.frombuffer(bytearray(jobs), dtype=torch.uint8).to("cuda")
    # HIP events on the library's own stream time the launch
    sp = C.c_void_p()
    if lib.tci_ctx_stream(ctx.handle, C.byref(sp)):
        raise RuntimeError(lib.tci_last_error().decode())
    lstream = torch.cuda.ExternalStream(sp.value)
    times, kms = [], []
    for it in range(args.rrlu_warmup + args.rrlu_steps):
        A.copy_(pristine)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(lstream)
        st = lib.tci_rrlu_batched_f64(ctx.handle, C.c_void_p(jb.data_ptr()), cnt)
        if st:
            raise RuntimeError(lib.tci_last_error().decode())
        e1.record(lstream)
        ctx.synchronize()
        if it >= args.rrlu_warmup:
            times.append(time.perf_counter() - t0)
            kms.append(e0.elapsed_time(e1))
    C.memmove(jobs, bytes(jb.cpu().numpy().tobytes()), C.sizeof(jobs))
    ps = np.array([jobs[i].npivot for i in range(cnt)])
    bad = sum(jobs[i].status != 0 for i in range(cnt))
    elapsed = sum(times)
    value = cnt * len(times) / elapsed
    tot_bytes = 0.0
    for (m, n), p in zip(shapes, ps):
        k = np.arange(1, p + 1, dtype=np.float64)
        tot_bytes += 8.0 * m * n + 16.0 * float(np.sum((m - k) * (n - k))) + 8.0 * p * (m + n)
    kms_avg = float(np.mean(kms))
    achieved = tot_bytes / (kms_avg * 1e-3) / 1e9
    traffic = profiled("tci_rrlu_kernel_c2_4096", "traffic_bytes_per_launch") if cnt == 4096 else None
    out = {"workload": "rrlu of %d matrices, rows/cols U{32..1024}, A = X.Y + 1e-12 E, rank U{8..128}, "
                       "reltol 1e-10, one launch" % cnt,
           "matrices_per_s": round(value, 2), "ms_per_launch": round(elapsed / len(times) * 1e3, 3),
           "mean_pivots": round(float(ps.mean()), 2), "failed": int(bad),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                        "kernel": "tci_rrlu_kernel", "kernel_ms": round(kms_avg, 3), "bytes_per_launch": tot_bytes,
                        "note": "achieved = algorithmic bytes 8mn + 16 sum (m-k)(n-k) + 8p(m+n) per matrix over "
                                "the launch time (HIP events on the library's stream); the trailing block is "
                                "streamed through L2/HBM once per pivot"}}
    O = _oracle()
    host = pristine.cpu().numpy()
    permh = perm.cpu().numpy()
    rng = np.random.default_rng(77)
    sample = sorted(rng.choice(cnt, size=min(args.rrlu_check, cnt), replace=False).tolist())
    def check(i):
        m, n = shapes[i]
        o = O.rrlu(host[offs[i]:offs[i] + m * n].reshape((m, n), order="F"), reltol=1e-10)
        return (o.npivot == jobs[i].npivot and
                np.array_equal(permh[poff[i]:poff[i] + m], o.rowpermutation) and
                np.array_equal(permh[poff[i] + m:poff[i] + m + n], o.colpermutation) and
                o.error == jobs[i].error)

    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=cpu["threads"]) as ex:   # the oracle's C calls release the GIL
        mism = sum(not ok for ok in ex.map(check, sample))
    out["pivot_check"] = {"matrices": len(sample), "mismatches": int(mism),
                          "note": "seeded sample of the same launch's matrices; rowpermutation, colpermutation, "
                                  "npivot and error compared with the oracle bit for bit, outside the timed region"}
    if args.cpu_seconds > 0:
        k = [0]
        lock = threading.Lock()

        def one():
            with lock:
                q = k[0] % cnt
                k[0] += 1
            m, n = shapes[q]
            O.rrlu(host[offs[q]:offs[q] + m * n].reshape((m, n), order="F"), reltol=1e-10)

        cps, nmat, dt = cpu_throughput(one, cpu["threads"], args.cpu_seconds)
        out["cpu_baseline"] = {"value": round(cps, 3), "unit": "matrices/s", "cores": cpu["threads"], "kind": "port",
                               "sample": f"{nmat} matrices of the same batch (in order) in {dt:.1f} s "
                                         f"on {cpu['threads']} threads"}
        out["vs_cpu"] = round(value / cps, 1)
    return out


def rrlu_bond_section(args, tci, ctx, cpu, m=4352, n=4352, rank=300, maxrank=256):
    """One bond of the L=40, d=16 workload: a single 4352 x 4352 Pi-sized
    matrix (rank 300 + 1e-12 noise) factored to maxrank 256, resident in
    HBM.  The cooperative whole-chip kernel (AUTO) against the one-workgroup
    streaming kernel (STREAM) and the oracle on one host thread; the pivots
    of every device run are compared with the oracle's."""
    import torch
    lib = tci._abi.lib()
    rng = np.random.default_rng(4)
    Ah = np.asfortranarray(rng.standard_normal((m, rank)) @ rng.standard_normal((rank, n))
                           + 1e-12 * rng.standard_normal((m, n)))
    pristine = torch.from_numpy(Ah.ravel(order="F").copy()).to("cuda")
    A = pristine.clone()
    perm = torch.zeros(m + n, dtype=torch.int32, device="cuda")
    job = tci._abi.tci_lu_job()
    job.A, job.rowperm, job.colperm = A.data_ptr(), perm.data_ptr(), perm.data_ptr() + 4 * m
    job.m, job.n, job.maxrank = m, n, maxrank
    job.reltol, job.abstol, job.leftorth = 1e-14, 0.0, 1
    jb = torch.frombuffer(bytearray(job), dtype=torch.uint8).to("cuda")
    sp = C.c_void_p()
    if lib.tci_ctx_stream(ctx.handle, C.byref(sp)):
        raise RuntimeError(lib.tci_last_error().decode())
    lstream = torch.cuda.ExternalStream(sp.value)
    O = _oracle()
    t0 = time.perf_counter()
    o = O.rrlu(Ah, maxrank=maxrank, reltol=1e-14)
    cpu_s = time.perf_counter() - t0

    def run(algo, steps):
        if lib.tci_ctx_set_rrlu_algo(ctx.handle, algo):
            raise RuntimeError(lib.tci_last_error().decode())
        kms, ok = [], True
        for _ in range(steps):
            A.copy_(pristine)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(lstream)
            if lib.tci_rrlu_batched_f64(ctx.handle, C.c_void_p(jb.data_ptr()), 1):
                raise RuntimeError(lib.tci_last_error().decode())
            e1.record(lstream)
            ctx.synchronize()
            kms.append(e0.elapsed_time(e1))
            r = tci._abi.tci_lu_job.from_buffer_copy(bytes(jb.cpu().numpy().tobytes()))
            ph = perm.cpu().numpy()
            ok = ok and (r.npivot == o.npivot and r.error == o.error and
                         np.array_equal(ph[:m], o.rowpermutation) and np.array_equal(ph[m:], o.colpermutation))
        lib.tci_ctx_set_rrlu_algo(ctx.handle, 0)
        return kms, ok

    run(0, 1)   # warm-up
    kms, ok_grid = run(0, args.rrlu_steps)
    skms, ok_stream = run(1, 1)
    k = np.arange(1, o.npivot + 1, dtype=np.float64)
    tot = 8.0 * m * n + 16.0 * float(np.sum((m - k) * (n - k))) + 8.0 * o.npivot * (m + n)
    kavg = float(np.mean(kms))
    achieved = tot / (kavg * 1e-3) / 1e9
    return {"workload": f"rrlu of one {m} x {n} matrix (rank {rank} + 1e-12 noise), maxrank {maxrank}, "
                        f"reltol 1e-14 (one bond of L=40, d=16, maxbonddim 256)",
            "pivots": int(o.npivot), "ms_cooperative": round(kavg, 3), "ms_one_workgroup": round(float(skms[0]), 3),
            "speedup_vs_one_workgroup": round(float(skms[0]) / kavg, 1),
            "pivots_match_oracle": bool(ok_grid and ok_stream),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "kernel": "tci_rrlu_grid_kernel",
                         "kernel_ms": round(kavg, 3), "bytes_per_launch": tot,
                         "note": "algorithmic bytes 8mn + 16 sum_k (m-k)(n-k) + 8p(m+n) over the launch time "
                                 "(HIP events on the library's stream, gather and copy-back included); the "
                                 "151 MB matrix is streamed once per pivot"},
            "cpu_baseline": {"value": round(cpu_s * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
                             "sample": "the same matrix factored once by the oracle (C, -O2) on one host thread"},
            "vs_cpu": round(cpu_s * 1e3 / kavg, 1)}


# ------------------------------------------------------------------ served
def slot_leg(args, tci, ctx, P, reqs, batch_cap):
    """The same requests admitted one at a time into a resident grid
    (tci_server_*): a request starts as soon as a workgroup is free.  The
    arrival streams are the batch leg's (lambda x the batch capacity), so the
    latencies compare like for like."""
    from importlib import import_module
    serving = import_module("tci_amd.serving")
    # the largest request of the mix sizes the grid's on-chip plan
    shape = (tci.DeviceFunctor(P.GSUM, np.asarray(P.gauss_sum(8, 12, 32, 0), float), np.zeros(0, np.int32)),
             [32] * 12, None, dict(tolerance=1e-10, maxbonddim=128))
    srv = serving.SlotServer(ctx, shape, slots=max(16384, len(reqs)), idle_ms=50)
    try:
        res, st = srv.run(reqs, np.zeros(len(reqs)))
        out = {"admission": "slot: a resident grid takes each request from a device queue",
               "capacity_solves_per_s": round(st.solves_per_s, 2), "failed": int(sum(r.status != 0 for r in res)),
               "capacity_retries": st.retries, "capacity_wall_ms": round(st.wall_s * 1e3, 1),
               "capacity_host_ms": {k: round(v, 1) for k, v in st.host_ms.items()}, "sweep": []}
        for lam in args.c5_lambdas:
            arr = serving.poisson_arrivals(len(reqs), lam * batch_cap, seed=5)
            res, st = srv.run(reqs, arr)
            out["sweep"].append({"lambda": lam, "arrival_rate_per_s": round(lam * batch_cap, 1),
                                 "solves_per_s": round(st.solves_per_s, 2), "p50_ms": round(st.p50_ms, 3),
                                 "p99_ms": round(st.p99_ms, 3), "failed": int(sum(r.status != 0 for r in res))})
    finally:
        srv.close()
    return out


def served_section(args, tci, ctx, cpu):
    """The served workload: 10 000 requests, each a sum of K ~ U{1..8}
    separable Gaussians in n ~ U{4..12} dims, one local dimension
    d ∈ {2,4,8,16,32}, tolerance 10^U(-10,-6), maxbonddim 128, arriving as a
    Poisson stream at λ × the measured capacity, admitted by continuous
    batching; λ swept."""
    from importlib import import_module
    serving = import_module("tci_amd.serving")
    P = _problems()
    mix = P.served_mix(args.c5_requests, seed=5)
    reqs = [(tci.DeviceFunctor(k, np.asarray(pa, float), np.asarray(ip, np.int32)), ld, piv, dict(o))
            for _, k, pa, ip, ld, piv, o in mix]
    srv = serving.TCIServer(ctx, streams=args.streams, max_batch=args.c5_max_batch)
    # capacity: every request present at t = 0
    res, st = srv.run(reqs, np.zeros(len(reqs)))
    cap = st.solves_per_s
    out = {"workload": f"{len(reqs)} Gaussian-sum requests (K~U{{1..8}}, n~U{{4..12}}, d in {{2,4,8,16,32}}, "
                       "tol 10^U(-10,-6), maxbonddim 128), Poisson arrivals at lambda x capacity",
           "capacity_solves_per_s": round(cap, 2), "streams": args.streams, "max_batch": args.c5_max_batch,
           "failed": int(sum(r.status != 0 for r in res)), "capacity_retries": st.retries, "sweep": []}
    for lam in args.c5_lambdas:
        arr = serving.poisson_arrivals(len(reqs), lam * cap, seed=5)
        res, st = srv.run(reqs, arr)
        out["sweep"].append({"lambda": lam, "arrival_rate_per_s": round(lam * cap, 1),
                             "solves_per_s": round(st.solves_per_s, 2), "p50_ms": round(st.p50_ms, 3),
                             "p99_ms": round(st.p99_ms, 3), "batches": st.batches,
                             "mean_batch": round(st.mean_batch, 1),
                             "failed": int(sum(r.status != 0 for r in res))})
    if args.c5_slot:
        out["slot"] = slot_leg(args, tci, ctx, P, reqs, cap)
    if args.cpu_seconds > 0:
        O = _oracle()
        k = [0]
        lock = threading.Lock()

        def one():
            with lock:
                q = k[0] % len(mix)
                k[0] += 1
            _, kind, pa, ip, ld, piv, o = mix[q]
            O.crossinterpolate2(O.Functor(kind, pa, ip), ld, piv, **{a: b for a, b in o.items() if a != "chi_cap"})

        cps, n, dt = cpu_throughput(one, cpu["threads"], args.cpu_seconds)
        out["cpu_baseline"] = {"value": round(cps, 3), "unit": "solves/s", "cores": cpu["threads"], "kind": "port",
                               "sample": f"the first {n} requests of the stream in {dt:.1f} s on {cpu['threads']} threads"}
        out["vs_cpu"] = round(cap / cps, 1)
    return out


# ------------------------------------------------------------------ main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["all", "solve", "rrlu", "bond", "served", "c3", "c4", "ttpost"], default="all",
                    help="all: the headline line with every sub-object; solve: the headline only; "
                         "rrlu: the batched rrLU only; bond: the C4 bond rrLU; served: the C5 stream; "
                         "c3, c4: one solve configuration")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--plain", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--served", type=int, default=8192, help="requests in the headline's served run; 0: skip")
    ap.add_argument("--served-load", type=float, default=0.5, help="arrival rate as a fraction of batched capacity")
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--rrlu-count", type=int, default=4096)
    ap.add_argument("--rrlu-steps", type=int, default=3)
    ap.add_argument("--rrlu-warmup", type=int, default=1)
    ap.add_argument("--rrlu-check", type=int, default=512, help="matrices whose pivots are checked against the oracle")
    ap.add_argument("--c3-batch", type=int, default=2048)
    ap.add_argument("--c4-batch", type=int, default=512)
    ap.add_argument("--c5-requests", type=int, default=10000)
    ap.add_argument("--c5-max-batch", type=int, default=4096)
    ap.add_argument("--c5-lambdas", type=float, nargs="*", default=[0.25, 0.5, 1.0, 2.0, 4.0])
    ap.add_argument("--c5-slot", type=int, default=1, help="1: also serve C5 with slot-level admission")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 3 plain device copies of 1 GiB (the reference kernel for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 2 * src.numel() * 8, "dispatches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(solves are independent: shard requests across GPUs)"}))
        sys.exit(2)
    import torch
    tci = load_pkg()
    ctx = tci.default_context()
    cpu = host_cpu()
    if args.workload == "bond":
        print(json.dumps(rrlu_bond_section(args, tci, ctx, cpu)), flush=True)
        return
    if args.workload == "served":
        print(json.dumps(served_section(args, tci, ctx, cpu)), flush=True)
        return
    if args.workload == "ttpost":
        print(json.dumps(tt_post_section(tci, ctx)), flush=True)
        return
    if args.workload in ("c3", "c4"):
        P = _problems()
        spec, batch, cs = ((P.c3_spec(5), args.c3_batch, args.cpu_seconds / 2) if args.workload == "c3"
                           else (P.c4_spec(5), args.c4_batch, args.cpu_seconds))
        steps, plain = (3, 5) if args.workload == "c3" else (2, 3)
        print(json.dumps(solve_config(tci, ctx, spec, batch, steps, plain, cpu, cs, 3)), flush=True)
        return
    if args.workload == "rrlu":
        r = rrlu_section(args, tci, ctx, cpu)
        out = {"metric": "batched standalone rrLU matrices/s (FP64, full pivoting, reltol 1e-10)",
               "value": r["matrices_per_s"], "unit": "matrices/s", "n_gpus": 1, "steps": args.rrlu_steps,
               "warmup": args.rrlu_warmup, "ms_per_step": r["ms_per_launch"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
               "config": {"workload": r["workload"]}, "roofline": r["roofline"], "rrlu": r,
               "cpu_baseline": r.get("cpu_baseline"), "host_cpu": cpu}
        print(json.dumps(out), flush=True)
        return
    P = _problems()
    spec = P.c1_spec(5)
    B = args.batch
    holder = tci._abi.ProblemHolder([10] * 8, tci._abi.F_LORENTZ_INT, [1.0], tolerance=1e-8)
    arr = (tci._abi.tci_problem * B)(*([holder.st] * B))
    one = (tci._abi.tci_problem * 1)(holder.st)
    tuned = tune_team(tci, ctx, holder, B)
    tuned_plain = tune_team(tci, ctx, holder, 1)
    for _ in range(args.warmup):
        run_batch(tci, ctx, arr, B)
    run_batch(tci, ctx, one, 1)
    # plain path: one solve per launch
    plain = []
    for _ in range(args.plain):
        wall, kms, ranks1, _ = run_batch(tci, ctx, one, 1)
        plain.append(wall)
    plain_lat = statistics.median(plain)
    # batched path: timed steps, bracketed by device synchronization
    torch.cuda.synchronize()
    walls, kms_all = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        wall, kms, ranks, ws = run_batch(tci, ctx, arr, B)
        walls.append(wall)
        kms_all.append(kms)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = B * args.steps / elapsed
    p50 = float(np.percentile(walls, 50)) * 1e3
    p99 = float(np.percentile(walls, 99)) * 1e3
    kms_avg = float(np.mean(kms_all))
    phases = dict(zip(["eval", "rrlu", "sets", "fill", "gsearch", "sweep1", "total", "pinv"], ws[6][:8]))
    # the Schur updates run on the FP64 units with the trailing block in
    # registers: the roof is FP64 (MI355X FP64 vector and matrix peaks are
    # both 78.6 TFLOP/s); F_schur = 2 sum (m-k)(n-k) over every rrLU of a solve
    schur_tflops = B * ws[3] / (kms_avg * 1e-3) / 1e12
    traffic = profiled("tci_solve_kernel_lorentz8_b4096", "traffic_bytes_per_launch") if B == 4096 else None
    streamed = 8.0 * ws[0] + ws[2]
    out = {
        "metric": "TCI solves/s + p50/p99 latency at tol 1e-8, batched/served vs plain on MI355X",
        "value": round(value, 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "crossinterpolate2 Lorentzian 1/(1+v.v), L=8, d=10, tol 1e-8, "
                               "initial pivot ones, optimize! defaults; %d solves per launch" % B,
                   "batch_solves": B, "ranks": ranks},
        "latency_ms": {"p50": round(p50, 3), "p99": round(p99, 3), "note": "per request, batch submit to results on host"},
        "plain": {"solves_per_s": round(1.0 / plain_lat, 2), "latency_ms_p50": round(plain_lat * 1e3, 3)},
        "batched_vs_plain": round(value * plain_lat, 1),
        "solve_phases_us": {k: round(v, 1) for k, v in phases.items()},
        "roofline": {"bound": "mfma", "achieved": round(schur_tflops, 4), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(schur_tflops / FP64_PEAK_TFLOPS, 5), "traffic": traffic,
                     "kernel": "tci_solve_kernel", "kernel_ms": round(kms_avg, 3),
                     "flops_per_solve": ws[3],
                     "hbm_measured": None if traffic is None else {
                         "GB_per_s": round(traffic / (kms_avg * 1e-3) / 1e9, 2),
                         "frac": round(traffic / (kms_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)},
                     "streamed_model_bytes_per_solve": streamed,
                     "note": "FP64 roof: the Schur updates (2 sum (m-k)(n-k) flops per rrLU) run on the FP64 "
                             "vector units with every Pi matrix and trailing block in registers/LDS "
                             "(MI355X FP64 vector = FP64 matrix peak = 78.6 TFLOP/s, no MFMA used); the "
                             "step is bound by the latency of the dependent pivot chain. traffic = measured "
                             "HBM bytes per launch (copy-calibrated FETCH_SIZE + WRITE_SIZE); "
                             "streamed_model_bytes_per_solve = what a design streaming each Pi through HBM "
                             "would move (8mn per evaluation + 8mn + 16 sum (m-k)(n-k) + 8p(m+n) per rrLU)"},
        "error": sample_error(tci, spec),
        "host_cpu": cpu,
    }
    if args.served > 0:
        # served: Poisson arrivals of the same request at a fraction of the
        # batched capacity, admitted by continuous batching over several streams
        from importlib import import_module
        serving = import_module("tci_amd.serving")
        rate = args.served_load * value
        f = tci.Lorentz(1.0)
        reqs = [(f, [10] * 8, None, dict(tolerance=1e-8))] * args.served
        arrivals = serving.poisson_arrivals(args.served, rate, seed=5)
        srv = serving.TCIServer(ctx, streams=args.streams, max_batch=B)
        res, st = srv.run(reqs, arrivals)
        bad = sum(r.status != 0 for r in res)
        out["served"] = {"requests": st.n, "arrival_rate_per_s": round(rate, 1),
                         "solves_per_s": round(st.solves_per_s, 2), "p50_ms": round(st.p50_ms, 3),
                         "p99_ms": round(st.p99_ms, 3), "batches": st.batches,
                         "mean_batch": round(st.mean_batch, 1), "streams": args.streams, "failed": bad,
                         "note": "latency from arrival to results on the host"}
        # the same stream admitted per workgroup slot by the resident grid
        # at this load fewer requests are in flight than there are CUs, so
        # one faster 512-thread solve per CU serves them best
        slot = serving.SlotServer(ctx, reqs[0], slots=max(16384, args.served), idle_ms=50, team=512)
        try:
            slot.run(reqs[:512], np.zeros(512), keep_history=False)   # warm the server's memory classes
            res, st = slot.run(reqs, arrivals, keep_history=False)
        finally:
            slot.close()
        out["served_slot"] = {"requests": st.n, "arrival_rate_per_s": round(rate, 1),
                              "solves_per_s": round(st.solves_per_s, 2), "p50_ms": round(st.p50_ms, 3),
                              "p99_ms": round(st.p99_ms, 3), "failed": int(sum(r.status != 0 for r in res)),
                              "host_ms": {k: round(v, 1) for k, v in st.host_ms.items()},
                              "team": 512,
                              "note": "slot-level admission: a resident grid takes each request as a workgroup "
                                      "frees; same arrivals as served"}
    if args.cpu_seconds > 0:
        O = _oracle()
        of = O.Functor(1, [1.0])

        def one_solve():
            O.crossinterpolate2(of, [10] * 8, tolerance=1e-8)

        lat1 = cpu_latency(one_solve, 5)
        cps, n, dt = cpu_throughput(one_solve, cpu["threads"], 1.5 * args.cpu_seconds)
        out["cpu_baseline"] = {"value": round(cps, 2), "unit": "solves/s", "cores": cpu["threads"], "kind": "port",
                               "sample": f"{n} solves of the same workload in {dt:.1f} s on {cpu['threads']} threads "
                                         f"(oracle C restatement, -O2, one solve per thread; host: {cpu['model']}, "
                                         f"nproc {cpu['nproc']})",
                               "latency_ms_1thread": round(lat1 * 1e3, 3)}
        out["vs_cpu"] = round(value / cps, 1)
    out["solve_team"] = tuned
    out["solve_team_plain"] = tuned_plain
    ctx.set_solve_tuning([])
    if args.workload == "all":
        out["rrlu_batched"] = rrlu_section(args, tci, ctx, cpu)
        out["rrlu_c4_bond"] = rrlu_bond_section(args, tci, ctx, cpu)
        out["c3_quantics"] = solve_config(tci, ctx, P.c3_spec(5), args.c3_batch, 3, 5, cpu, args.cpu_seconds / 2, 3)
        out["c4_highdim"] = solve_config(tci, ctx, P.c4_spec(5), args.c4_batch, 2, 3, cpu, args.cpu_seconds, 3)
        out["c5_served"] = served_section(args, tci, ctx, cpu)
        out["tt_post"] = tt_post_section(tci, ctx)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
