"""SYMM restated on the CPU — TEST INFRASTRUCTURE ONLY (imported by tests as
the checker).

src/interface/symm_interface.hpp:36-80: C <- alpha A B + beta C (side 'l',
A is M x M) or alpha B A + beta C (side 'r', A is N x N), with A symmetric and
only its `uplo` triangle referenced ('u': A(i, j) for i <= j; 'l': i >= j);
side and uplo are case-insensitive.  The product is the FP64 GEMM oracle
(oracle.gemm_f64) on the mirrored matrix.

Pinned in tests/test_symm.py against OpenBLAS's dsymm (scipy.linalg.blas)
for both sides and triangles, and by a planted mistake (reading the other
triangle) that must fail that comparison.
"""
from __future__ import annotations

import numpy as np


def mirror(k: int, a: np.ndarray, lda: int, uplo: str) -> np.ndarray:
    """The full k x k symmetric matrix (column-major, ld k) from the stored
    triangle of column-major `a`."""
    A = np.asarray(a, np.float64)[: lda * k].reshape(k, lda).T[:k, :]  # A[i, j] = a[i + j lda]
    up = uplo.lower() == "u"
    i, j = np.indices((k, k))
    stored = (i <= j) if up else (i >= j)
    full = np.where(stored, A, A.T)
    return np.ascontiguousarray(full.T).reshape(-1)  # column-major flat


def symm_f64(side, uplo, M, N, alpha, a, lda, b, ldb, beta, c, ldc) -> np.ndarray:
    from . import gemm_f64

    s = side.lower()
    k = M if s == "l" else N
    full = mirror(k, a, lda, uplo)
    if s == "l":
        return gemm_f64("n", "n", M, N, k, alpha, full, k, np.asarray(b, np.float64), ldb, beta,
                        np.asarray(c, np.float64), ldc)
    return gemm_f64("n", "n", M, N, k, alpha, np.asarray(b, np.float64), ldb, full, k, beta,
                    np.asarray(c, np.float64), ldc)
