"""FP32 / FP64 GEMM (plain MFMA and SIMT), batched, strided-batched and
interleaved on the GPU vs the FP64 oracle and the CBLAS golden fixtures.
Sweeps follow test/unittest/blas3/blas3_gemm_test.cpp:29-141 and
blas3_gemm_batched_test.cpp:29-140."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import U32, U64, cfg1_inputs, check_cfg1, colmajor, to_dev

pytestmark = pytest.mark.gpu


def bound(K, alpha, absprod, beta, cabs, u):
    return 1.01 * (K + 2) * u * abs(alpha) * absprod + 2 * u * abs(beta) * cabs + 1e-30


def run_case(handle, torch, ta, tb, m, n, k, alpha, beta, lda_mul=1, ldb_mul=1, ldc_mul=1,
             offset=0, dtype=np.float32, seed=0, simt=False, poison_c=False):
    import portblas_amd as pb

    ar, ac = (m, k) if ta == "n" else (k, m)
    br, bc = (k, n) if tb == "n" else (n, k)
    lda, ldb, ldc = max(ar * lda_mul, 1), max(br * ldb_mul, 1), max(m * ldc_mul, 1)
    A = oracle.uniform(seed * 3 + 1, offset + lda * ac).astype(dtype)
    B = oracle.uniform(seed * 3 + 2, offset + ldb * bc).astype(dtype)
    C = oracle.uniform(seed * 3 + 3, offset + ldc * n).astype(dtype)
    if poison_c:
        C[offset:] = np.nan
    dA, dB, dC = to_dev(torch, A), to_dev(torch, B), to_dev(torch, C)
    if simt:
        pb._gemm_ex(handle, pb.gemm_precision_t.fp32_simt, ta, tb, m, n, k, alpha, dA[offset:], lda,
                    dB[offset:], ldb, beta, dC[offset:], ldc)
    else:
        pb._gemm(handle, ta, tb, m, n, k, alpha, dA[offset:], lda, dB[offset:], ldb, beta, dC[offset:], ldc)
    got = dC.cpu().numpy().astype(np.float64)[offset:]
    Cin = np.nan_to_num(C[offset:]) if poison_c else C[offset:]
    ref = oracle.gemm_f64(ta, tb, m, n, k, alpha, A[offset:], lda, B[offset:], ldb, beta, Cin, ldc)
    absprod = colmajor(oracle.gemm_absprod_f64(ta, tb, m, n, k, A[offset:], lda, B[offset:], ldb), m, n, m)
    u = U32 if dtype == np.float32 else U64
    tol = bound(k, alpha, absprod, beta, np.abs(colmajor(Cin, m, n, ldc)), u)
    gm, rm = colmajor(got, m, n, ldc), colmajor(ref, m, n, ldc)
    assert np.all(np.isfinite(gm)), "non-finite output"
    err = np.abs(gm - rm)
    assert np.all(err <= tol), f"max err/tol {np.max(err / tol):.3g}"
    if ldc > m and not poison_c:  # padding rows untouched
        np.testing.assert_array_equal(got.reshape(n, ldc)[:, m:], C[offset:].reshape(n, ldc)[:, m:])
    return got


@pytest.mark.parametrize("ci", range(5))
def test_gemm_golden(handle, torch_cuda, golden, ci):
    import portblas_amd as pb

    torch = torch_cuda
    p = f"gemm{ci}_"
    ta, tb, m, n, k, lda, ldb, ldc = (int(v) for v in golden[p + "args"])
    ta, tb = chr(ta), chr(tb)
    alpha, beta = (float(v) for v in golden[p + "scal"])
    sa, sb, sc = (int(v) for v in golden[p + "seeds"])
    A = oracle.uniform(sa, lda * (k if ta == "n" else m))
    B = oracle.uniform(sb, ldb * (n if tb == "n" else k))
    C = oracle.uniform(sc, ldc * n)
    dC = to_dev(torch, C)
    pb._gemm(handle, ta, tb, m, n, k, alpha, to_dev(torch, A), lda, to_dev(torch, B), ldb, beta, dC, ldc)
    got = colmajor(dC.cpu().numpy().astype(np.float64), m, n, ldc)
    cb = colmajor(golden[p + "out"].astype(np.float64), m, n, ldc)
    absprod = colmajor(oracle.gemm_absprod_f64(ta, tb, m, n, k, A, lda, B, ldb), m, n, m)
    tol = 2 * bound(k, alpha, absprod, beta, np.abs(colmajor(C, m, n, ldc)), U32)
    assert np.all(np.abs(got - cb) <= tol)


SHAPES = [(11, 16, 17), (32, 11, 63), (16, 32, 253), (253, 130, 511), (128, 128, 128),
          (129, 257, 64), (1, 1, 1), (64, 64, 4096), (300, 9, 1000)]


@pytest.mark.parametrize("ta", "nt")
@pytest.mark.parametrize("tb", "nt")
@pytest.mark.parametrize("m,n,k", SHAPES)
def test_gemm_sweep(handle, torch_cuda, ta, tb, m, n, k):
    run_case(handle, torch_cuda, ta, tb, m, n, k, 1.5, 0.5, seed=m + n + k)


@pytest.mark.parametrize("alpha,beta", [(1.5, 1.5), (1.0, 0.0), (1.0, 1.0), (-2.0, 0.25)])
@pytest.mark.parametrize("lda_mul,ldb_mul,ldc_mul", [(1, 1, 1), (2, 3, 4)])
@pytest.mark.parametrize("offset", [0, 1, 10])
def test_gemm_scalars_ld_offset(handle, torch_cuda, alpha, beta, lda_mul, ldb_mul, ldc_mul, offset):
    for ta, tb in (("n", "n"), ("t", "n"), ("n", "t")):
        run_case(handle, torch_cuda, ta, tb, 253, 130, 253, alpha, beta, lda_mul, ldb_mul, ldc_mul,
                 offset, seed=7)


def test_gemm_beta_zero_never_reads_c(handle, torch_cuda):
    for ta in "nt":
        run_case(handle, torch_cuda, ta, "n", 130, 70, 33, 1.0, 0.0, ldc_mul=2, poison_c=True, seed=3)
        run_case(handle, torch_cuda, ta, "n", 256, 256, 64, 1.0, 0.0, poison_c=True, seed=4)


def test_gemm_cfg1_1024(handle, torch_cuda, golden):
    """The reference's CPU-runnable configuration: 1024^3, 'n','n', ld 1024,
    alpha 1.5, beta 0.5, inputs U[-1,1] from seed 0, checked against the
    committed CBLAS sgemm entries and checksums and against the FP64 oracle;
    then the same shape on other seeds through run_case."""
    import portblas_amd as pb

    torch = torch_cuda
    ta, tb, m, n, k, lda, ldb, ldc, alpha, beta, A, B, C = cfg1_inputs(golden)
    dC = to_dev(torch, C)
    pb._gemm(handle, ta, tb, m, n, k, alpha, to_dev(torch, A), lda, to_dev(torch, B), ldb, beta, dC, ldc)
    got = dC.cpu().numpy()
    check_cfg1(golden, got, 2.0)
    ref = oracle.gemm_f64(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    absprod = oracle.gemm_absprod_f64(ta, tb, m, n, k, A, lda, B, ldb)
    assert np.all(np.abs(got - ref) <= bound(k, alpha, absprod, beta, np.abs(C), U32))
    run_case(handle, torch, "n", "n", 1024, 1024, 1024, 1.5, 0.5, seed=0)


@pytest.mark.parametrize("ta", "nt")
@pytest.mark.parametrize("tb", "nt")
def test_dgemm(handle, torch_cuda, ta, tb):
    for (m, n, k) in [(33, 17, 40), (257, 129, 300), (128, 128, 512)]:
        run_case(handle, torch_cuda, ta, tb, m, n, k, 1.25, -0.5, dtype=np.float64, seed=m)


@pytest.mark.parametrize("ta", "nt")
@pytest.mark.parametrize("tb", "nt")
def test_gemm_simt(handle, torch_cuda, ta, tb):
    for (m, n, k) in [(11, 16, 17), (253, 130, 253), (256, 256, 96)]:
        run_case(handle, torch_cuda, ta, tb, m, n, k, 1.5, 0.5, seed=m + 1, simt=True)


def test_gemm_alpha_zero_rules(handle, torch_cuda):
    """gemm_interface.hpp:112-139: alpha == 0 scales C by beta, before any
    transpose check; _scal_matrix multiplies (NaN * 0 stays NaN, beta == 1
    leaves C), the contiguous strided-batch case uses _scal (beta == 0 zeroes)."""
    import portblas_amd as pb

    torch = torch_cuda
    m = n = 8
    C0 = torch.arange(m * n, dtype=torch.float32, device="cuda")
    C = C0.clone()
    pb._gemm(handle, "x", "y", m, n, 4, 0.0, C, m, C, m, 2.0, C, m)  # invalid trans, no error
    assert torch.equal(C, 2 * C0)
    C = C0.clone()
    C[5] = float("nan")
    pb._gemm(handle, "n", "n", m, n, 4, 0.0, C, m, C, m, 0.0, C, m)
    assert torch.isnan(C[5]) and torch.count_nonzero(torch.nan_to_num(C)) == 0
    C = C0.clone()
    pb._gemm(handle, "n", "n", m, n, 4, 0.0, C, m, C, m, 1.0, C, m)
    assert torch.equal(C, C0)
    # strided batch with stride == ldc * N and ldc == M: one _scal over all batches
    Cb = torch.full((3 * m * n,), float("nan"), device="cuda")
    pb._gemm_batched(handle, "n", "n", m, n, 4, 0.0, Cb, m, Cb, 4, 0.0, Cb, m, 3)
    assert torch.count_nonzero(Cb) == 0 and not torch.isnan(Cb).any()


def test_gemm_invalid_arguments(handle, torch_cuda):
    import portblas_amd as pb

    torch = torch_cuda
    X = torch.zeros(64, device="cuda")
    with pytest.raises(ValueError, match="_TransA"):
        pb._gemm(handle, "x", "n", 4, 4, 4, 1.0, X, 4, X, 4, 0.0, X, 4)
    with pytest.raises(ValueError, match="_TransB"):
        pb._gemm(handle, "n", "q", 4, 4, 4, 1.0, X, 4, X, 4, 0.0, X, 4)
    with pytest.raises(ValueError, match="_stridec"):
        pb._gemm_strided_batched(handle, "n", "n", 4, 4, 4, 1.0, X, 4, 16, X, 4, 16, 0.0, X, 4, 15, 2)
    with pytest.raises(ValueError, match="_stridea"):
        pb._gemm_strided_batched(handle, "n", "n", 4, 4, 4, 1.0, X, 4, -1, X, 4, 16, 0.0, X, 4, 16, 2)
    # 'C' is accepted as a transpose (real types) and upper case works.
    pb._gemm(handle, "C", "T", 4, 4, 4, 1.0, X, 4, X, 4, 0.0, X, 4)


def _batched_ref(ta, tb, m, n, k, alpha, A, lda, sa, B, ldb, sb, beta, C, ldc, sc, batch):
    out = C.astype(np.float64).copy()
    tols = []
    for b in range(batch):
        Ab, Bb = A[b * sa:], B[b * sb:]
        r = oracle.gemm_f64(ta, tb, m, n, k, alpha, Ab, lda, Bb, ldb, beta, C[b * sc: b * sc + ldc * n], ldc)
        out[b * sc: b * sc + ldc * n] = r
        ap = colmajor(oracle.gemm_absprod_f64(ta, tb, m, n, k, Ab, lda, Bb, ldb), m, n, m)
        tols.append(bound(k, alpha, ap, beta, np.abs(colmajor(C[b * sc:], m, n, ldc)), U32))
    return out, tols


@pytest.mark.parametrize("ta", "nt")
@pytest.mark.parametrize("tb", "nt")
@pytest.mark.parametrize("mnk", [63, 128, 129])
@pytest.mark.parametrize("batch", [1, 5])
def test_gemm_batched_strided(handle, torch_cuda, ta, tb, mnk, batch):
    import portblas_amd as pb

    torch = torch_cuda
    m = n = k = mnk
    lda, ldb, ldc = 2 * m, 3 * k, 4 * m  # ld multipliers of the reference test
    sa, sb, sc = lda * k, ldb * n, ldc * n  # default strides (gemm_interface.hpp:213-219)
    A = oracle.uniform(11, sa * batch)
    B = oracle.uniform(12, sb * batch)
    C = oracle.uniform(13, sc * batch)
    dC = to_dev(torch, C)
    pb._gemm_batched(handle, ta, tb, m, n, k, 3.0, to_dev(torch, A), lda, to_dev(torch, B), ldb, 7.0,
                     dC, ldc, batch)
    got = dC.cpu().numpy().astype(np.float64)
    ref, tols = _batched_ref(ta, tb, m, n, k, 3.0, A, lda, sa, B, ldb, sb, 7.0, C, ldc, sc, batch)
    for b in range(batch):
        e = np.abs(colmajor(got[b * sc:], m, n, ldc) - colmajor(ref[b * sc:], m, n, ldc))
        assert np.all(e <= tols[b])


def test_gemm_strided_batched_broadcast_and_padding(handle, torch_cuda):
    import portblas_amd as pb

    torch = torch_cuda
    m, n, k, batch = 63, 63, 128, 5
    lda, ldb, ldc = 2 * m, k, m
    sa, sb, sc = 0, ldb * n, 2 * ldc * n  # A broadcast, C stride padded
    A = oracle.uniform(21, lda * k)
    B = oracle.uniform(22, sb * batch)
    C = oracle.uniform(23, sc * batch)
    dC = to_dev(torch, C)
    pb._gemm_strided_batched(handle, "n", "n", m, n, k, 3.0, to_dev(torch, A), lda, sa,
                             to_dev(torch, B), ldb, sb, 1.0, dC, ldc, sc, batch)
    got = dC.cpu().numpy().astype(np.float64)
    ref, tols = _batched_ref("n", "n", m, n, k, 3.0, A, lda, sa, B, ldb, sb, 1.0, C, ldc, sc, batch)
    for b in range(batch):
        e = np.abs(colmajor(got[b * sc:], m, n, ldc) - colmajor(ref[b * sc:], m, n, ldc))
        assert np.all(e <= tols[b])
        # the gap between matrices is untouched
        np.testing.assert_array_equal(got[b * sc + ldc * n: (b + 1) * sc], C[b * sc + ldc * n: (b + 1) * sc])


def _to_interleaved(x, ld_rows, ld_cols, batch):
    """strided -> interleaved (test/unittest/blas3/blas3_gemm_common.hpp:53-69)."""
    return x.reshape(batch, ld_cols, ld_rows).transpose(1, 2, 0).reshape(-1).copy()


def _from_interleaved(x, ld_rows, ld_cols, batch):
    return x.reshape(ld_cols, ld_rows, batch).transpose(2, 0, 1).reshape(-1).copy()


@pytest.mark.parametrize("ta", "nt")
@pytest.mark.parametrize("tb", "nt")
def test_gemm_batched_interleaved(handle, torch_cuda, ta, tb):
    import portblas_amd as pb

    torch = torch_cuda
    m, n, k, batch = 63, 65, 33, 5
    ar, ac = (m, k) if ta == "n" else (k, m)
    br, bc = (k, n) if tb == "n" else (n, k)
    lda, ldb, ldc = ar, br, m
    A = oracle.uniform(31, lda * ac * batch)
    B = oracle.uniform(32, ldb * bc * batch)
    C = oracle.uniform(33, ldc * n * batch)
    dC = to_dev(torch, _to_interleaved(C, ldc, n, batch))
    pb._gemm_batched(handle, ta, tb, m, n, k, 1.5, to_dev(torch, _to_interleaved(A, lda, ac, batch)), lda,
                     to_dev(torch, _to_interleaved(B, ldb, bc, batch)), ldb, 0.5, dC, ldc, batch,
                     pb.gemm_batch_type_t.interleaved)
    got = _from_interleaved(dC.cpu().numpy().astype(np.float64), ldc, n, batch)
    sa, sb, sc = lda * ac, ldb * bc, ldc * n
    ref, tols = _batched_ref(ta, tb, m, n, k, 1.5, A, lda, sa, B, ldb, sb, 0.5, C, ldc, sc, batch)
    for b in range(batch):
        e = np.abs(colmajor(got[b * sc:], m, n, ldc) - colmajor(ref[b * sc:], m, n, ldc))
        assert np.all(e <= tols[b])


def test_strided_batched_4096_x_64(handle, torch_cuda):
    """The benchmark's strided set (4096 x 64^3, stride = matrix size) on 64
    sampled batches."""
    import portblas_amd as pb

    torch = torch_cuda
    batch, s = 4096, 64
    g = torch.Generator(device="cuda").manual_seed(5)
    A = torch.rand(batch * s * s, device="cuda", generator=g) * 2 - 1
    B = torch.rand(batch * s * s, device="cuda", generator=g) * 2 - 1
    C = torch.zeros(batch * s * s, device="cuda")
    pb._gemm_strided_batched(handle, "n", "n", s, s, s, 1.0, A, s, s * s, B, s, s * s, 0.0, C, s, s * s, batch)
    An, Bn, Cn = (t.cpu().numpy() for t in (A, B, C))
    for b in np.random.default_rng(1).choice(batch, 64, replace=False):
        sl = slice(b * s * s, (b + 1) * s * s)
        ref = oracle.gemm_f64("n", "n", s, s, s, 1.0, An[sl], s, Bn[sl], s, 0.0, Cn[sl], s)
        ap = oracle.gemm_absprod_f64("n", "n", s, s, s, An[sl], s, Bn[sl], s)
        assert np.all(np.abs(Cn[sl] - ref) <= bound(s, 1.0, ap, 0.0, 0.0, U32))


@pytest.mark.parametrize("ta,tb", [("n", "n"), ("n", "t"), ("t", "n"), ("t", "t")])
@pytest.mark.parametrize("batch,pad", [(256, 0), (300, 4), (1031, 0), (4099, 0)])
def test_small_64_cubed_beta_zero(handle, torch_cuda, ta, tb, batch, pad):
    """FP32 64^3 batches with beta = 0 run on the persistent kernel that keeps
    the next problems' operands in flight: one to 17 problems per workgroup,
    every transpose, tight and padded leading dimensions.  C starts as NaN:
    beta = 0 must not read it, and every element is written."""
    import portblas_amd as pb

    torch = torch_cuda
    s = 64
    ld = s + pad
    sa = sb = ld * s
    sc = ld * s
    A = oracle.uniform(31, sa * batch).astype(np.float32)
    B = oracle.uniform(32, sb * batch).astype(np.float32)
    dC = torch.full((sc * batch,), float("nan"), device="cuda")
    pb._gemm_strided_batched(handle, ta, tb, s, s, s, -1.5, to_dev(torch, A), ld, sa, to_dev(torch, B), ld, sb,
                             0.0, dC, ld, sc, batch)
    got = dC.cpu().numpy().astype(np.float64).reshape(batch, s, ld)[:, :, :s]
    assert np.all(np.isfinite(got))
    zero = np.zeros(sc, np.float32)
    for b in [0, 1, 255, batch // 2, batch - 2, batch - 1]:
        ref = oracle.gemm_f64(ta, tb, s, s, s, -1.5, A[b * sa:], ld, B[b * sb:], ld, 0.0, zero, ld)
        ap = colmajor(oracle.gemm_absprod_f64(ta, tb, s, s, s, A[b * sa:], ld, B[b * sb:], ld), s, s, s)
        e = np.abs(got[b].T - colmajor(ref, s, s, ld))
        assert np.all(e <= bound(s, -1.5, ap, 0.0, 0.0, U32)), (b, float(np.max(e)))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ta,tb", [("n", "n"), ("n", "t"), ("t", "n"), ("t", "t")])
@pytest.mark.parametrize("m,n,k,lda_pad", [(64, 64, 64, 0), (64, 64, 192, 0), (64, 64, 128, 4), (37, 50, 200, 3),
                                             (8, 61, 5, 1)])
def test_many_small_problems(handle, torch_cuda, dtype, ta, tb, m, n, k, lda_pad):
    """Batches of >= one problem per CU with M, N <= 64 take the
    one-problem-per-workgroup kernels: full 64 x 64 tiles with K in one to
    three chunks of 64 (tight and padded, 16-byte aligned leading
    dimensions), ragged M/N, K over several chunks,
    unaligned leading dimensions (scalar staging), alpha/beta; 300 problems,
    checked on a sample against the FP64 oracle."""
    import portblas_amd as pb

    torch = torch_cuda
    batch = 300
    ar, ac = (m, k) if ta == "n" else (k, m)
    br, bc = (k, n) if tb == "n" else (n, k)
    lda, ldb, ldc = ar + lda_pad, br + lda_pad, m + lda_pad
    sa, sb, sc = lda * ac, ldb * bc, ldc * n
    A = oracle.uniform(21, sa * batch).astype(dtype)
    B = oracle.uniform(22, sb * batch).astype(dtype)
    C = oracle.uniform(23, sc * batch).astype(dtype)
    dC = to_dev(torch, C)
    fn = pb._gemm_strided_batched
    fn(handle, ta, tb, m, n, k, -1.5, to_dev(torch, A), lda, sa, to_dev(torch, B), ldb, sb, 0.25, dC, ldc, sc,
       batch)
    got = dC.cpu().numpy().astype(np.float64)
    u = U32 if dtype == np.float32 else U64
    for b in list(range(3)) + [150, batch - 1]:
        Ab, Bb, Cb = A[b * sa:], B[b * sb:], C[b * sc: b * sc + ldc * n]
        ref = oracle.gemm_f64(ta, tb, m, n, k, -1.5, Ab, lda, Bb, ldb, 0.25, Cb, ldc)
        ap = colmajor(oracle.gemm_absprod_f64(ta, tb, m, n, k, Ab, lda, Bb, ldb), m, n, m)
        tol = bound(k, -1.5, ap, 0.25, np.abs(colmajor(Cb, m, n, ldc)), u)
        e = np.abs(colmajor(got[b * sc:], m, n, ldc) - colmajor(ref, m, n, ldc))
        assert np.all(e <= tol), (b, float(np.max(e / tol)))
        if ldc > m:  # padding rows untouched
            np.testing.assert_array_equal(got[b * sc: b * sc + ldc * n].reshape(n, ldc)[:, m:],
                                          Cb.reshape(n, ldc)[:, m:])
