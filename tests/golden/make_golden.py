"""Writes tests/golden/golden_cases.npz: small seeded cases with expected
outputs from CBLAS (OpenBLAS through scipy.linalg.blas — the library the
reference's tests call, common/include/common/system_reference_blas.hpp) and,
where CBLAS has no routine, from the reference tests' own hand-written rules:

* iamin: the first strict minimum of |x| (system_reference_blas.hpp:91-101);
* extension reduction: the hand loop of
  test/unittest/extension/reduction_test.cpp:155-180 and the worked example of
  doc/Reduction.md:522-546 (3 x 3 matrix 1..9: inner -> 6 15 24, outer -> 12 15 18);
* iamax/iamin known-answer patterns of
  test/unittest/blas1/blas1_iaminmax_common.hpp:31-63 (Limit, Incrementing,
  Decrementing; iamin replaces 0 by 1, blas1_iamin_test.cpp:45-50).

Inputs are splitmix64 streams (oracle.uniform); only their seeds are stored
(plus the head of stream 0, which pins the generator), so the fixture stays
small and the script is deterministic.  Run from the repo root: python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
from scipy.linalg import blas as sblas

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import oracle  # noqa: E402

OUT = Path(__file__).resolve().parent / "golden_cases.npz"


def colmajor(flat, rows, cols, ld):
    return np.asarray(flat).reshape(cols, ld).T[:rows, :]


def gemm_cases(d):
    cases = [
        # ta, tb, m, n, k, alpha, beta, lda_mul, ldb_mul, ldc_mul
        ("n", "n", 11, 16, 17, 1.5, 0.5, 1, 1, 1),
        ("t", "n", 32, 11, 63, 1.0, 0.0, 2, 1, 3),
        ("n", "t", 16, 32, 16, 1.5, 1.5, 1, 3, 1),
        ("t", "t", 33, 17, 40, -0.75, 1.0, 1, 2, 2),
        ("n", "n", 64, 64, 64, 1.5, 0.5, 1, 1, 1),
    ]
    for ci, (ta, tb, m, n, k, alpha, beta, lm, lbm, lcm) in enumerate(cases):
        ar, ac = (m, k) if ta == "n" else (k, m)
        br, bc = (k, n) if tb == "n" else (n, k)
        lda, ldb, ldc = ar * lm, br * lbm, m * lcm
        A = oracle.uniform(100 + ci, lda * ac)
        B = oracle.uniform(200 + ci, ldb * bc)
        C = oracle.uniform(300 + ci, ldc * n)
        Am, Bm = colmajor(A, ar, ac, lda), colmajor(B, br, bc, ldb)
        Cm = np.asfortranarray(colmajor(C, m, n, ldc).astype(np.float32))
        out = sblas.sgemm(alpha, Am, Bm, beta=beta, c=Cm, trans_a=ta == "t", trans_b=tb == "t")
        Cres = C.copy().reshape(n, ldc).T
        Cres[:m, :] = out
        p = f"gemm{ci}_"
        d[p + "args"] = np.array([ord(ta), ord(tb), m, n, k, lda, ldb, ldc], dtype=np.int64)
        d[p + "scal"] = np.array([alpha, beta], dtype=np.float64)
        d[p + "seeds"] = np.array([100 + ci, 200 + ci, 300 + ci], dtype=np.int64)
        d[p + "out"] = Cres.T.reshape(-1).astype(np.float32)


def gemm_cfg1(d):
    """The reference's CPU-runnable configuration (1024^3, 'n','n', lda = ldb =
    ldc = 1024, alpha 1.5, beta 0.5) on one splitmix64 stream, seed 0: A is
    its first 1024^2 values, B the next, C the next.  Stored: 64 x 64 sampled
    entries of the CBLAS result (rows and columns include the ends and the
    128-multiples) and its FP64 row and column sums."""
    m = n = k = 1024
    u = oracle.uniform(0, m * k + k * n + m * n)
    A, B, C = u[: m * k], u[m * k: m * k + k * n], u[m * k + k * n:]
    out = sblas.sgemm(1.5, colmajor(A, m, k, m), colmajor(B, k, n, k), beta=0.5,
                      c=np.asfortranarray(colmajor(C, m, n, m)))
    rng = np.random.default_rng(1024)
    fixed = [0, 1, 127, 128, 511, 512, 895, 896, 1022, 1023]
    rows = np.array(sorted(set(fixed) | set(rng.choice(m, 64, replace=False).tolist())))[:64]
    cols = np.array(sorted(set(fixed) | set(rng.choice(n, 64, replace=False).tolist())))[:64]
    d["cfg1_args"] = np.array([ord("n"), ord("n"), m, n, k, m, k, m], dtype=np.int64)
    d["cfg1_scal"] = np.array([1.5, 0.5])
    d["cfg1_seed"] = np.int64(0)
    d["cfg1_rows"] = rows.astype(np.int64)
    d["cfg1_cols"] = cols.astype(np.int64)
    d["cfg1_sampled"] = out[np.ix_(rows, cols)].astype(np.float32)
    d["cfg1_colsum"] = out.astype(np.float64).sum(axis=0)
    d["cfg1_rowsum"] = out.astype(np.float64).sum(axis=1)


def gemv_cases(d):
    cases = [("n", 11, 14, 1.0, 0.0, 1, 2, 3), ("t", 11, 14, 1.5, 1.5, 2, 2, 3),
             ("n", 1023, 1010, 1.0, 1.5, 1, 1, 1), ("t", 1023, 1010, 1.0, 0.0, 1, 1, 1)]
    for ci, (tr, m, n, alpha, beta, lda_mul, incx, incy) in enumerate(cases):
        lda = m * lda_mul
        xl, yl = (n, m) if tr == "n" else (m, n)
        A = oracle.uniform(400 + ci, lda * n)
        x = oracle.uniform(500 + ci, 1 + (xl - 1) * incx)
        y = oracle.uniform(600 + ci, 1 + (yl - 1) * incy)
        Am = colmajor(A, m, n, lda)
        yo = sblas.sgemv(alpha, Am, x, beta=beta, y=y.copy(), incx=incx, incy=incy, trans=int(tr == "t"))
        p = f"gemv{ci}_"
        d[p + "args"] = np.array([ord(tr), m, n, lda, incx, incy], dtype=np.int64)
        d[p + "scal"] = np.array([alpha, beta])
        d[p + "seeds"] = np.array([400 + ci, 500 + ci, 600 + ci], dtype=np.int64)
        d[p + "out"] = yo.astype(np.float32)


def blas1_cases(d):
    for ci, (n, incx, incy) in enumerate([(11, 1, 1), (1002, 4, 3), (1002, 1, 1), (65537, 1, 1)]):
        x = oracle.uniform(700 + ci, 1 + (n - 1) * incx)
        y = oracle.uniform(800 + ci, 1 + (n - 1) * incy)
        p = f"blas1_{ci}_"
        d[p + "args"] = np.array([n, incx, incy], dtype=np.int64)
        d[p + "seeds"] = np.array([700 + ci, 800 + ci], dtype=np.int64)
        d[p + "dot"] = np.float64(sblas.sdot(x, y, n=n, incx=incx, incy=incy))
        d[p + "asum"] = np.float64(sblas.sasum(x, n=n, incx=incx))
        d[p + "nrm2"] = np.float64(sblas.snrm2(x, n=n, incx=incx))
        d[p + "iamax"] = np.int64(sblas.isamax(x, n=n, incx=incx))
        ax = np.abs(x[: n * incx : incx].astype(np.float64))
        best = 0
        for i in range(1, n):  # system_reference_blas.hpp:91-101 (first strict minimum)
            if ax[i] < ax[best]:
                best = i
        d[p + "iamin"] = np.int64(best)


def iaminmax_patterns(d):
    """blas1_iaminmax_common.hpp:31-63 (Limit, Incrementing, Decrementing)."""
    for name in ("limit", "incrementing", "decrementing"):
        for n in (11, 65, 1000):
            if name == "limit":
                x = np.full(n, np.finfo(np.float32).max, dtype=np.float32)
            elif name == "incrementing":
                x = np.arange(n, dtype=np.float32)
            else:
                x = np.arange(n, 0, -1, dtype=np.float32) - 1
            xmin = x.copy()
            xmin[xmin == 0] = 1  # blas1_iamin_test.cpp:45-50
            d[f"pat_{name}_{n}_x"] = x
            d[f"pat_{name}_{n}_iamax"] = np.int64(sblas.isamax(x))
            d[f"pat_{name}_{n}_xmin"] = xmin
            ax = np.abs(xmin.astype(np.float64))
            d[f"pat_{name}_{n}_iamin"] = np.int64(int(np.argmin(ax)))  # argmin = first minimum


def reduction_cases(d):
    # doc/Reduction.md:522-546 worked example: column-major 1..9.
    A = np.arange(1, 10, dtype=np.float32)
    d["red_doc_A"] = A
    d["red_doc_inner"] = np.array([6, 15, 24], dtype=np.float32)
    d["red_doc_outer"] = np.array([12, 15, 18], dtype=np.float32)
    # Hand loop of reduction_test.cpp:155-180 on seeded inputs.
    ops = ["add", "product", "max", "min", "abs_add", "mean"]
    for ci, (rows, cols, ldm) in enumerate([(7, 65, 2), (513, 9, 1), (1, 8195, 3)]):
        ld = rows * ldm
        A = oracle.uniform(900 + ci, ld * cols, lo=-2.0, hi=2.0)
        M = colmajor(A, rows, cols, ld).astype(np.float64)
        d[f"red{ci}_args"] = np.array([rows, cols, ld], dtype=np.int64)
        d[f"red{ci}_seed"] = np.int64(900 + ci)
        for op in ops:
            for dim, axis in (("inner", 0), ("outer", 1)):
                if op == "add":
                    v = M.sum(axis=axis)
                elif op == "product":
                    v = np.prod(M, axis=axis)
                elif op == "max":
                    v = M.max(axis=axis)
                elif op == "min":
                    v = M.min(axis=axis)
                elif op == "abs_add":
                    v = np.abs(M).sum(axis=axis)
                else:
                    v = M.mean(axis=axis)
                d[f"red{ci}_{op}_{dim}"] = v


def main():
    d: dict[str, np.ndarray] = {}
    # Pins the input generator: the first values of stream 0 (inputs of the
    # other cases are regenerated from their seeds by the tests).
    d["uniform_seed0_head"] = oracle.uniform(0, 16)
    gemm_cases(d)
    gemm_cfg1(d)
    gemv_cases(d)
    blas1_cases(d)
    iaminmax_patterns(d)
    reduction_cases(d)
    np.savez_compressed(OUT, **d)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(d)} arrays)")


if __name__ == "__main__":
    main()
