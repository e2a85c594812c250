"""CPU restatements of calaman_G's balancing (?gebal) and 1-norm estimator (?lacn2)
-- TEST INFRASTRUCTURE.

Only ``tests/`` import this module, as the checker of the HIP path
(calaman_g_amd/csrc/balance.hip). numpy, FP64 / complex128.

* ``gebal``: src/lapack/gebal/interface.cppm + gebal.cu, which follow LAPACK's
  xGEBAL through 3.4 (EISPACK BALANC): the permutation that isolates eigenvalues
  (rows pushed down from the bottom, columns pushed left), then Parlett-Reinsch
  sweeps with off-diagonal 1-norms (|re| + |im| for complex), radix 2, the 0.95
  acceptance factor and the SFMIN/SFMAX guards; an index with non-finite norms is
  skipped (src/lapack/gebal/README.md "Differences").
* ``lacn2``: src/lapack/lacn2/interface.cppm, LAPACK's xLACN2 (Hager's method with
  Higham's refinements) as a driver that takes the matrix itself.
"""
from __future__ import annotations

import numpy as np

NONE, PERMUTE, SCALE, BOTH = 0, 1, 2, 3


def _abs1(x):
    return np.abs(x.real) + np.abs(x.imag) if np.iscomplexobj(x) else np.abs(x)


def _limits(single: bool):
    fi = np.finfo(np.float32 if single else np.float64)
    sfmin1 = float(fi.tiny) / float(fi.eps)  # DLAMCH('S') / DLAMCH('P')
    sfmax1 = 1.0 / sfmin1
    sfmin2 = sfmin1 * 2.0
    sfmax2 = 1.0 / sfmin2
    return sfmin1, sfmax1, sfmin2, sfmax2


def gebal(A, job: int = BOTH, single: bool = False, max_sweeps: int = 1000):
    """Returns (B, ilo, ihi, scale) with B = D^-1 P^T A P D; ilo, ihi 1-based;
    scale[i] the diagonal of D inside [ilo, ihi], the 1-based exchange index outside."""
    A = np.array(A, dtype=np.complex128 if np.iscomplexobj(A) else np.float64)
    n = A.shape[0]
    scale = np.ones(n)
    if n == 0:
        return A, 1, 0, scale
    if job == NONE:
        return A, 1, n, scale
    k, l = 0, n - 1  # 0-based window [k, l]
    if job in (PERMUTE, BOTH):
        # rows isolating an eigenvalue, pushed down
        while True:
            found = False
            for j in range(l, -1, -1):
                row = A[j, :l + 1].copy()
                row[j] = 0
                if np.all(row == 0):
                    scale[l] = j + 1
                    if j != l:
                        A[:l + 1, [j, l]] = A[:l + 1, [l, j]]
                        A[[j, l], k:] = A[[l, j], k:]
                    found = True
                    break
            if not found:
                break
            if l == 0:
                return A, k + 1, l + 1, scale
            l -= 1
        # columns isolating an eigenvalue, pushed left
        while True:
            found = False
            for j in range(k, l + 1):
                col = A[k:l + 1, j].copy()
                col[j - k] = 0
                if np.all(col == 0):
                    scale[k] = j + 1
                    if j != k:
                        A[:l + 1, [j, k]] = A[:l + 1, [k, j]]
                        A[[j, k], k:] = A[[k, j], k:]
                    found = True
                    break
            if not found:
                break
            k += 1
    scale[k:l + 1] = 1.0
    if job == PERMUTE:
        return A, k + 1, l + 1, scale
    sfmin1, sfmax1, sfmin2, sfmax2 = _limits(single)
    for _ in range(max_sweeps):
        noconv = False
        for i in range(k, l + 1):
            colv = _abs1(A[k:l + 1, i])
            rowv = _abs1(A[i, k:l + 1])
            colv[i - k] = 0.0  # the off-diagonal sums (src/lapack/gebal/gebal.cu:203-222)
            rowv[i - k] = 0.0
            c = float(colv.sum())
            r = float(rowv.sum())
            ca = float(_abs1(A[:l + 1, i]).max())
            ra = float(_abs1(A[i, k:]).max())
            if c == 0 or r == 0 or not np.isfinite(c + r + ca + ra):
                continue
            s0 = c + r
            f = 1.0
            g = r / 2
            while not (c >= g or max(f, c, ca) >= sfmax2 or min(r, g, ra) <= sfmin2):
                f *= 2; c *= 2; ca *= 2; r /= 2; g /= 2; ra /= 2
            g = c / 2
            while not (g < r or max(r, ra) >= sfmax2 or min(f, c, g, ca) <= sfmin2):
                f /= 2; c /= 2; g /= 2; ca /= 2; r *= 2; ra *= 2
            if c + r >= 0.95 * s0:
                continue
            if f < 1 and scale[i] < 1 and f * scale[i] <= sfmin1:
                continue
            if f > 1 and scale[i] > 1 and scale[i] >= sfmax1 / f:
                continue
            scale[i] *= f
            noconv = True
            A[i, k:] *= 1.0 / f
            A[:l + 1, i] *= f
        if not noconv:
            break
    return A, k + 1, l + 1, scale


def apply_balance_similarity(A, ilo, ihi, scale):
    """B = D^-1 P^T A P D from gebal's outputs, computed independently of gebal:
    the exchanges replayed as gebal made them (bottom from n down to ihi+1, then top
    from 1 up to ilo-1), then the diagonal scaling"""
    A = np.array(A, dtype=np.complex128 if np.iscomplexobj(A) else np.float64)
    n = A.shape[0]
    perm = np.arange(n)
    for m in range(n - 1, ihi - 1, -1):
        j = int(scale[m]) - 1
        perm[[j, m]] = perm[[m, j]]
    for m in range(0, ilo - 1):
        j = int(scale[m]) - 1
        perm[[j, m]] = perm[[m, j]]
    d = np.ones(n)
    d[ilo - 1:ihi] = scale[ilo - 1:ihi]
    B = A[np.ix_(perm, perm)]
    return B / d[:, None] * d[None, :]


def lacn2(A, max_iter: int = 5):
    """||A||_1 estimate (xLACN2 driven to completion with the matrix at hand):
    returns (est, v, kase sequence)"""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    seq = []
    x = np.full(n, 1.0 / n)
    x = A @ x
    seq.append(1)
    if n == 1:
        return abs(x[0]), x.copy(), seq
    est = float(np.abs(x).sum())
    isgn = np.where(x >= 0, 1, -1)
    x = A.T @ isgn.astype(np.float64)
    seq.append(2)
    j = int(np.argmax(np.abs(x)))
    it = 2
    v = None
    while True:
        x = np.zeros(n)
        x[j] = 1.0
        x = A @ x
        seq.append(1)
        v = x.copy()
        estold = est
        est = float(np.abs(v).sum())
        sgn = np.where(x >= 0, 1, -1)
        if np.all(sgn == isgn) or est <= estold:
            break
        isgn = sgn
        x = A.T @ isgn.astype(np.float64)
        seq.append(2)
        jlast = j
        j = int(np.argmax(np.abs(x)))
        if x[jlast] != abs(x[j]) and it < max_iter:
            it += 1
            continue
        break
    alt = np.array([(1 if i % 2 == 0 else -1) * (1.0 + i / (n - 1)) for i in range(n)])
    x = A @ alt
    seq.append(1)
    temp = 2.0 * (float(np.abs(x).sum()) / (3 * n))
    if temp > est:
        v = x.copy()
        est = temp
    return est, v, seq
