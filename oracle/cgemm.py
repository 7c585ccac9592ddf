"""Complex GEMM restated on the CPU — TEST INFRASTRUCTURE ONLY (imported by
tests as the checker).

The reference's _gemm on complex types (src/interface/gemm_interface.hpp:
140-198, with BLAS_ENABLE_COMPLEX): C <- alpha op(A) op(B) + beta C on
column-major interleaved complex matrices, op(X) = X for 'n' and X^T for
't' or 'c' (no conjugation: the reference only tests 'n' and 't' and maps
every non-'n' flag to a transpose).  Computed in complex128.

Pinned in tests/test_cgemm.py against OpenBLAS's zgemm (scipy.linalg.blas)
for 'n' / 't', and by a planted mistake (conjugating on 't') that must fail.
"""
from __future__ import annotations

import numpy as np


def dense(flat, rows, cols, ld):
    return np.asarray(flat, np.complex128)[: ld * cols].reshape(cols, ld).T[:rows, :]


def cgemm_c128(ta, tb, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc) -> np.ndarray:
    """Column-major flat complex128 copy of c after the update."""
    A = dense(a, M, K, lda) if ta.lower() == "n" else dense(a, K, M, lda).T
    B = dense(b, K, N, ldb) if tb.lower() == "n" else dense(b, N, K, ldb).T
    out = np.array(c, dtype=np.complex128, copy=True)
    view = out[: ldc * N].reshape(N, ldc).T
    Cm = view[:M, :]
    prod = A @ B if K > 0 else np.zeros((M, N), np.complex128)
    view[:M, :] = complex(alpha) * prod + (complex(beta) * Cm if complex(beta) != 0 else 0)
    return out
