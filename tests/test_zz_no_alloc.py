"""The batched rrLU entry point that allocates nothing
(tci_rrlu_batched_noalloc_f64) against tci_rrlu_batched_f64, through the C
ABI: packed factors, permutations, pivot counts, errors and statuses bit for
bit, on two calls in a row (the second with fewer, smaller matrices), with
tile-sized and streamed matrices in each call; a batch whose staging would
exceed the fixed scratch is refused before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs(tci, mats, lo, A, perm):
    shapes = [m.shape for m in mats]
    jobs = (tci._abi.tci_lu_job * len(mats))()
    o = po = 0
    for i, (m, n) in enumerate(shapes):
        j = jobs[i]
        j.A = A.data_ptr() + 8 * o
        j.rowperm = perm.data_ptr() + 4 * po
        j.colperm = perm.data_ptr() + 4 * (po + m)
        j.m, j.n, j.maxrank = m, n, tci.TCI_UNLIMITED
        j.reltol, j.abstol, j.leftorth = 1e-11, 0.0, int(lo)
        o += m * n
        po += m + n
    return jobs


def _run(tci, ctx, mats, lo, entry):
    import torch
    flat = np.concatenate([m.ravel(order="F") for m in mats])
    A = torch.from_numpy(flat).to("cuda")
    perm = torch.zeros(sum(m.shape[0] + m.shape[1] for m in mats), dtype=torch.int32, device="cuda")
    jobs = _jobs(tci, mats, lo, A, perm)
    jb = torch.frombuffer(bytearray(jobs), dtype=torch.uint8).to("cuda")
    torch.cuda.synchronize()
    st = getattr(tci._abi.lib(), entry)(ctx.handle, C.c_void_p(jb.data_ptr()), len(mats))
    assert st == 0, tci._abi.lib().tci_last_error()
    ctx.synchronize()
    C.memmove(jobs, bytes(jb.cpu().numpy().tobytes()), C.sizeof(jobs))
    out = [(int(j.npivot), float(j.error), int(j.status)) for j in jobs]
    return A.cpu().numpy(), perm.cpu().numpy(), out


def _mats(seed, count, hi):
    rng = np.random.default_rng(seed)
    mats = []
    for t in range(count):
        big = t % 3 == 0
        m = int(rng.integers(129, hi)) if big else int(rng.integers(1, 129))
        n = int(rng.integers(1, hi)) if big else int(rng.integers(1, 129))
        r = int(rng.integers(1, 24))
        mats.append(rng.standard_normal((m, r)) @ rng.standard_normal((r, n)) + 1e-10 * rng.standard_normal((m, n)))
    return mats


def _same(a, b):
    assert np.array_equal(a[0], b[0])                 # packed A (L, U, trailing block)
    assert np.array_equal(a[1], b[1])                 # permutations
    for (pa, ea, sa), (pb, eb, sb) in zip(a[2], b[2]):
        assert pa == pb and sa == sb and (ea == eb or (np.isnan(ea) and np.isnan(eb)))


def test_noalloc_equals_allocating_entry_two_calls(tci, ctx):
    first = _mats(1, 36, 700)
    second = _mats(2, 13, 300)          # fewer and smaller: the scratch keeps the first call's contents
    for mats, lo in ((first, True), (second, False)):
        print("noalloc call", len(mats), lo, flush=True)
        ref = _run(tci, ctx, mats, lo, "tci_rrlu_batched_f64")
        got = _run(tci, ctx, mats, lo, "tci_rrlu_batched_noalloc_f64")
        _same(ref, got)


def test_noalloc_refuses_scratch_overflow(tci, ctx):
    import torch
    # sides above the LDS staging (3072) need 2·side doubles per job of
    # staging: 700 jobs of 3100×1 need about 35 MB, beyond the 32 MiB scratch
    mats = [np.ones((3100, 1))] * 700
    A = torch.zeros(3100 * 700, dtype=torch.float64, device="cuda")
    perm = torch.zeros(3101 * 700, dtype=torch.int32, device="cuda")
    jobs = _jobs(tci, mats, True, A, perm)
    jb = torch.frombuffer(bytearray(jobs), dtype=torch.uint8).to("cuda")
    torch.cuda.synchronize()
    lib = tci._abi.lib()
    assert lib.tci_rrlu_batched_noalloc_f64(ctx.handle, C.c_void_p(jb.data_ptr()), len(mats)) == 3
    assert b"scratch" in lib.tci_last_error()
    # nothing ran: the matrices are untouched
    assert float(A.sum().item()) == 0.0
