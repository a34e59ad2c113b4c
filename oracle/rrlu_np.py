"""Independent numpy restatement of the reference rrLU (test infrastructure).

TEST INFRASTRUCTURE ONLY (the checker): written separately from
tci_oracle.c as Gaussian elimination with complete pivoting (Golub & Van
Loan, Algorithm 3.4.2) with the reference's conventions, so that the two
restatements can pin each other bitwise:

* pivot = argmax of abs2 (x*x) over the trailing block, first maximum in
  column-major order, NaN never chosen, (k,k) if every entry is NaN
  (matrixlu.jl:1-32);
* full row then full column swap, then true division of the pivot column
  (leftorthogonal) or row, then A22 -= x·yᵀ with separate rounding of the
  product and the difference (matrixlu.jl:98-136);
* stop before a pivot whose |value| < reltol·max|accepted| or < abstol, but
  only once a pivot exists; error bookkeeping as matrixlu.jl:141-181.
"""
from __future__ import annotations

import numpy as np


def rrlu_np(A, maxrank=None, reltol=1e-14, abstol=0.0, leftorthogonal=True):
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = A.shape
    rowperm = np.arange(1, m + 1)
    colperm = np.arange(1, n + 1)
    mr = min(m, n) if maxrank is None else min(int(maxrank), m, n)
    maxerr = 0.0
    err = float("nan")
    p = 0
    while p < mr:
        k = p
        blk = A[k:, k:]
        v = blk * blk
        v = np.where(np.isnan(v), -np.inf, v)
        flat = int(np.argmax(v.ravel(order="F")))
        r = k + flat % (m - k)
        c = k + flat // (m - k)
        err = abs(A[r, c])
        if (err < reltol * maxerr or err < abstol) and p > 0:
            break
        maxerr = float("nan") if (np.isnan(maxerr) or np.isnan(err)) else max(maxerr, err)
        A[[k, r], :] = A[[r, k], :]
        rowperm[[k, r]] = rowperm[[r, k]]
        A[:, [k, c]] = A[:, [c, k]]
        colperm[[k, c]] = colperm[[c, k]]
        piv = A[k, k]
        if leftorthogonal:
            A[k + 1:, k] = A[k + 1:, k] / piv
        else:
            A[k, k + 1:] = A[k, k + 1:] / piv
        prod = np.multiply.outer(A[k + 1:, k], A[k, k + 1:])
        A[k + 1:, k + 1:] = A[k + 1:, k + 1:] - prod
        p += 1
    L = np.tril(A[:, :p])
    U = np.triu(A[:p, :])
    if np.isnan(L).any():
        raise FloatingPointError("lu.L contains NaNs")
    if np.isnan(U).any():
        raise FloatingPointError("lu.U contains NaNs")
    if leftorthogonal:
        L[np.arange(p), np.arange(p)] = 1.0
    else:
        U[np.arange(p), np.arange(p)] = 1.0
    if p >= min(m, n):
        err = 0.0
    return dict(rowpermutation=rowperm, colpermutation=colperm, L=L, U=U, npivot=p, error=err)
