"""CPU restatement of calaman_G's expm_herm -- TEST INFRASTRUCTURE.

Only ``tests/`` import this module, as the checker of the HIP path
(calaman_g_amd/csrc/herm.hip). numpy, FP64 / complex128.

src/expm/expm_herm.cppm:145-199: the uplo triangle of A read (the other assumed to
mirror it; the diagonal's imaginary part ignored, as the vendor syevd/heevd do), the
eigendecomposition A = U diag(w) U^H by syevd/heevd, M = U diag(exp(w)) (the
herm_exp_scale kernel, src/expm/expm.cu:170-184), exp(A) = M U^H. numpy's eigh is
LAPACK's ?syevd/?heevd, the reference's own solver family.
"""
from __future__ import annotations

import numpy as np

LOWER, UPPER = 0, 1


def hermitian_from_triangle(A, uplo: int):
    """The full matrix the solver sees: the uplo triangle mirrored, a real diagonal"""
    A = np.asarray(A)
    cplx = np.iscomplexobj(A)
    T = np.triu(A, 1) if uplo == UPPER else np.tril(A, -1)
    H = T + T.conj().T
    H = H.astype(np.complex128 if cplx else np.float64)
    H[np.diag_indices_from(H)] = np.real(np.diag(A))
    return H


def expm_herm(A, uplo: int = UPPER):
    """exp(A) for a symmetric / Hermitian A given by its uplo triangle"""
    H = hermitian_from_triangle(A, uplo)
    if H.shape[0] == 0:
        return H.copy()
    w, U = np.linalg.eigh(H, UPLO="U" if uplo == UPPER else "L")
    M = U * np.exp(w)[None, :]
    return M @ U.conj().T
