"""Reduced-precision GEMMs (pb_gemm_ex with BF16x3 / BF16x4 / BF16 / FP8-e4m3,
the pack-once API and the BF16-input GEMM) on the GPU against an FP64
reference, with forward-error bounds derived from each format's rounding:

  BF16x3  x = hi + lo + r, |x - hi| <= 2^-8 |x|, |r| <= 2^-16 |x|; the three
          kept products miss hi*r_b + r_a*hi + lo*lo (<= 3.1 * 2^-16 |a||b|)
  BF16x4  also keeps lo*lo                          (<= 2.1 * 2^-16 |a||b|)
  BF16    |ab - hi_a hi_b| <= (2 * 2^-8 + 2^-16) |a||b|
  FP8     e4m3 after scaling by 2^s, the largest power of two with
          amax 2^s <= 448: relative 2^-4 plus an absolute 2^-10 2^-s below
          the normal range
  FP16x3  x_s = x 2^s (amax 2^s in [2^14, 2^15)) = hi + lo + r in binary16,
          |x_s - hi| <= 2^-11 |x_s|, |r| <= 2^-22 |x_s| while lo is normal,
          else an absolute 2^-25; the kept products miss hi*r_b + r_a*hi +
          lo*lo (<= 3.01 * 2^-22 |a||b|) plus 2^-25 2^-s per element
plus FP32 accumulation, (terms * K + 2) * 2^-24 |a||b|.  Shapes cover the
256 x 256 fused-term kernel (large M, N, ragged edges, K not a multiple of
the K step, batches) and the tile engine (small shapes)."""
from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U = 2.0 ** -24


def cm(flat, rows, cols, ld):
    """Column-major rows x cols view of a flat torch buffer."""
    return flat[: cols * ld].view(cols, ld).t()[:rows, :]


def make(torch, n, seed, dist="uniform"):
    g = torch.Generator(device="cuda").manual_seed(seed)
    if dist == "uniform":
        return torch.rand(n, generator=g, device="cuda") * 2 - 1
    if dist == "normal":
        return torch.randn(n, generator=g, device="cuda")
    mag = torch.exp2(torch.rand(n, generator=g, device="cuda") * 16 - 8)
    sign = torch.where(torch.rand(n, generator=g, device="cuda") < 0.5, -1.0, 1.0)
    return (mag * sign).float()


def check_lowp(handle, torch, prec, ta, tb, m, n, k, alpha=1.0, beta=0.0, batch=1, seed=0,
               dist="uniform", pad=0):
    import portblas_amd as pb

    P = pb.gemm_precision_t
    ar, ac = (m, k) if ta == "n" else (k, m)
    br, bc = (k, n) if tb == "n" else (n, k)
    lda, ldb, ldc = max(ar, 1) + pad, max(br, 1) + pad, max(m, 1)
    sa, sb, sc = lda * ac, ldb * bc, ldc * n
    A = make(torch, sa * batch, seed * 3 + 1, dist)
    B = make(torch, sb * batch, seed * 3 + 2, dist)
    C = make(torch, sc * batch, seed * 3 + 3)
    C0 = C.clone()
    pb._gemm_ex(handle, prec, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, sa, sb, sc, batch)
    torch.cuda.synchronize()
    terms = {P.bf16x3: 3, P.bf16x4: 4, P.bf16: 1, P.fp8_e4m3: 1, P.fp16x3: 3}[prec]
    rep = {P.bf16x3: 3.1 * 2.0 ** -16, P.bf16x4: 2.1 * 2.0 ** -16, P.bf16: 2.01 * 2.0 ** -8,
           P.fp8_e4m3: 0.13, P.fp16x3: 3.01 * 2.0 ** -22}[prec]
    amax_a = A.abs().max().double().item()
    amax_b = B.abs().max().double().item()
    worst = 0.0
    for b in range(batch):
        oa = cm(A[b * sa:], ar, ac, lda).double()
        ob = cm(B[b * sb:], br, bc, ldb).double()
        oa = oa if ta == "n" else oa.t()
        ob = ob if tb == "n" else ob.t()
        c0 = cm(C0[b * sc:], m, n, ldc).double()
        ref = alpha * (oa @ ob) + beta * c0
        absprod = oa.abs() @ ob.abs()
        tol = (rep + 1.01 * (terms * k + 2) * U) * abs(alpha) * absprod + 2 * U * abs(beta) * c0.abs()
        if prec in (P.fp8_e4m3, P.fp16x3):
            if prec == P.fp8_e4m3:
                ea, eb = 2.0 ** -10 / fp8_scale(amax_a), 2.0 ** -10 / fp8_scale(amax_b)
            else:
                ea, eb = 2.0 ** -25 / f16_scale(amax_a), 2.0 ** -25 / f16_scale(amax_b)
            sub = 1.07 * (ea * ob.abs().sum(0, keepdim=True) + eb * oa.abs().sum(1, keepdim=True)) + k * ea * eb
            tol = tol + abs(alpha) * sub
        got = cm(C[b * sc:], m, n, ldc).double()
        assert torch.isfinite(got).all(), "non-finite output"
        err = (got - ref).abs()
        r = (err / (tol + 1e-30)).max().item()
        worst = max(worst, r)
    assert worst <= 1.0, f"max err/tol {worst:.3g}"
    return worst


def f16_scale(amax):
    """The FP16x3 pre-pass scale 2^s: amax 2^s in [2^14, 2^15), s in [-63, 74]."""
    if not amax > 0:
        return 1.0
    s = 15 - math.frexp(amax)[1]
    return 2.0 ** min(max(s, -63), 74)


def fp8_scale(amax):
    """The FP8 pre-pass scale: the largest power of two 2^s with
    amax 2^s <= 448 (e4m3's largest finite value), s in [-63, 63]; 1 for an
    amax of 0 or infinity."""
    if not 0 < amax < math.inf:
        return 1.0
    s = math.floor(math.log2(448.0 / amax))
    while amax * 2.0 ** (s + 1) <= 448.0:
        s += 1
    while amax * 2.0 ** s > 448.0:
        s -= 1
    return 2.0 ** min(max(s, -63), 63)


PRECS = ["bf16x3", "bf16x4", "bf16", "fp8_e4m3", "fp16x3"]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("ta,tb", [("n", "n"), ("n", "t"), ("t", "n"), ("t", "t")])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (77, 130, 200), (33, 17, 1000), (300, 260, 64)])
def test_lowp_small(handle, torch_cuda, prec, ta, tb, m, n, k):
    import portblas_amd as pb

    check_lowp(handle, torch_cuda, pb.gemm_precision_t[prec], ta, tb, m, n, k, 1.5, 0.5)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("ta,tb,m,n,k,batch", [
    ("n", "n", 4096, 4096, 256, 1),   # whole 256 x 256 tiles
    ("t", "n", 4000, 4100, 1000, 1),  # ragged M, N edges; K tail inside the packed pitch
    ("n", "t", 4100, 4000, 72, 1),
    ("t", "t", 2048, 2048, 96, 4),    # batched
])
def test_lowp_large_tiles(handle, torch_cuda, prec, ta, tb, m, n, k, batch):
    import portblas_amd as pb

    check_lowp(handle, torch_cuda, pb.gemm_precision_t[prec], ta, tb, m, n, k, -1.25, 0.75, batch=batch,
               seed=3)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("ta,tb", [("n", "t"), ("t", "n")])
def test_lowp_padded_leading_dimensions(handle, torch_cuda, prec, ta, tb):
    """Operands with lda / ldb above their row count (the pre-passes walk them
    line by line rather than as one dense block)."""
    import portblas_amd as pb

    check_lowp(handle, torch_cuda, pb.gemm_precision_t[prec], ta, tb, 300, 260, 200, 1.5, 0.5, pad=5, seed=8)
    check_lowp(handle, torch_cuda, pb.gemm_precision_t[prec], ta, tb, 2048, 1024, 136, 1.0, 0.0, batch=2, pad=3,
               seed=9)


@pytest.mark.parametrize("dist", ["normal", "loguniform"])
def test_lowp_wide_range_inputs(handle, torch_cuda, dist):
    import portblas_amd as pb

    for prec in ("bf16x3", "bf16x4", "fp8_e4m3", "fp16x3"):
        check_lowp(handle, torch_cuda, pb.gemm_precision_t[prec], "n", "n", 4096, 4096, 512, dist=dist, seed=5)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("ta,tb,m,n,k,batch", [
    ("n", "n", 4096, 4096, 1000, 1),  # K ends inside the last packed K step
    ("t", "n", 4000, 4100, 2040, 1),
    ("n", "t", 4096, 4352, 3200, 2),  # K a whole number of ring revolutions, batched
])
def test_lowp_exact_on_integer_inputs(handle, torch_cuda, prec, ta, tb, m, n, k, batch):
    """Integers in [-2, 2] are exact in every format (BF16 hi with lo = 0;
    FP8 after the per-tensor scale 2^7, which maps them to 0, 128, 256),
    their products and FP32 sums are exact, so every K step must land
    exactly once: the result is the integer product up to the final
    dequantising multiply (a step dropped, repeated or read from the wrong
    ring stage moves it by whole integers)."""
    import portblas_amd as pb

    torch = torch_cuda
    g = torch.Generator(device="cuda").manual_seed(m + n + k)
    ar, ac = (m, k) if ta == "n" else (k, m)
    br, bc = (k, n) if tb == "n" else (n, k)
    A = torch.randint(-2, 3, (ar * ac * batch,), generator=g, device="cuda").float()
    B = torch.randint(-2, 3, (br * bc * batch,), generator=g, device="cuda").float()
    C = torch.full((m * n * batch,), float("nan"), device="cuda")
    pb._gemm_ex(handle, pb.gemm_precision_t[prec], ta, tb, m, n, k, 1.0, A, ar, B, br, 0.0, C, m,
                ar * ac, br * bc, m * n, batch)
    torch.cuda.synchronize()
    for b in range(batch):
        oa = cm(A[b * ar * ac:], ar, ac, ar).double()
        ob = cm(B[b * br * bc:], br, bc, br).double()
        oa = oa if ta == "n" else oa.t()
        ob = ob if tb == "n" else ob.t()
        ref = oa @ ob
        got = cm(C[b * m * n:], m, n, m).double()
        assert torch.isfinite(got).all()
        assert ((got - ref).abs() <= 2.0 ** -20 * ref.abs()).all(), (b, (got - ref).abs().max().item())


@pytest.mark.parametrize("pos", ["first", "last"])
def test_fp8_amax_over_every_workgroup(handle, torch_cuda, pos):
    """The per-tensor amax folds one partial per pre-pass workgroup: a needle
    (1000, against U[-1,1] elsewhere) at the first or last element of a
    16.8M-element operand, which the first or last of the 2048 workgroups
    sees, must set the scale (a missed partial would saturate the needle at
    448 2^-s(rest) and break the error bound)."""
    import portblas_amd as pb

    torch = torch_cuda
    m = n = 64
    k = 1 << 18
    idx = 0 if pos == "first" else -1
    A = make(torch, m * k, 31)
    B = make(torch, k * n, 32)
    A[idx] = 1000.0
    B[idx] = -1000.0
    C = torch.zeros(m * n, device="cuda")
    pb._gemm_ex(handle, pb.gemm_precision_t.fp8_e4m3, "n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    torch.cuda.synchronize()
    oa, ob = cm(A, m, k, m).double(), cm(B, k, n, k).double()
    ref = oa @ ob
    absprod = oa.abs() @ ob.abs()
    ea = eb = 2.0 ** -10 / fp8_scale(1000.0)
    tol = (0.13 + 1.01 * (k + 2) * U) * absprod + 1.07 * (ea * ob.abs().sum(0, keepdim=True)
                                                       + eb * oa.abs().sum(1, keepdim=True)) + k * ea * eb
    err = (cm(C, m, n, m).double() - ref).abs()
    assert torch.isfinite(err).all()
    assert (err <= tol).all(), (err / tol).max().item()


def test_bf16x4_not_less_accurate_than_bf16x3(handle, torch_cuda):
    """The extra lo*lo term can only shrink the representation error: on
    normal inputs the mean componentwise error must drop."""
    import portblas_amd as pb

    torch = torch_cuda
    m = n = 4096
    k = 2048
    A = make(torch, m * k, 11, "normal")
    B = make(torch, k * n, 12, "normal")
    ref = cm(A, m, k, m).double() @ cm(B, k, n, k).double()
    absprod = cm(A, m, k, m).double().abs() @ cm(B, k, n, k).double().abs()
    errs = {}
    for prec in ("bf16x3", "bf16x4"):
        C = torch.zeros(m * n, device="cuda")
        pb._gemm_ex(handle, pb.gemm_precision_t[prec], "n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
        errs[prec] = ((cm(C, m, n, m).double() - ref).abs() / absprod).mean().item()
    assert errs["bf16x4"] < 0.9 * errs["bf16x3"], errs


def _unslab(x, rows, ldp, slab=16):
    """Packed K-slab-major operand -> rows x ldp (element (r, k) at
    x[(k // slab) * rows * slab + r * slab + k % slab])."""
    return x.view(ldp // slab, rows, slab).permute(1, 0, 2).reshape(rows, ldp)


@pytest.mark.parametrize("m,n,k", [(4096, 4096, 512), (300, 200, 100)])
def test_pack_once_matches_gemm_ex(handle, torch_cuda, m, n, k):
    """pb_pack_bf16x3 + pb_gemm_bf16x3_packed: the split is exact to 2^-16,
    the pad is zero, and the product matches pb_gemm_ex(BF16X3) (bit for bit
    where both run the 256 x 256 kernel)."""
    import portblas_amd as pb

    torch = torch_cuda
    L = pb.load()
    A = make(torch, m * k, 21)
    B = make(torch, k * n, 22)
    C1 = torch.zeros(m * n, device="cuda")
    C2 = torch.zeros(m * n, device="cuda")
    pb._gemm_ex(handle, pb.gemm_precision_t.bf16x3, "n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C1, m)
    ldp = ((k + 63) // 64) * 64
    ah = torch.empty(m * ldp, dtype=torch.int16, device="cuda")
    al = torch.empty_like(ah)
    bh = torch.empty(n * ldp, dtype=torch.int16, device="cuda")
    bl = torch.empty_like(bh)
    pb.check(L.pb_pack_bf16x3(handle.handle, 0, b"n", m, k, A.data_ptr(), m, ah.data_ptr(), al.data_ptr(), ldp))
    pb.check(L.pb_pack_bf16x3(handle.handle, 1, b"n", n, k, B.data_ptr(), k, bh.data_ptr(), bl.data_ptr(), ldp))
    pb.check(L.pb_gemm_bf16x3_packed(handle.handle, m, n, k, 1.0, ah.data_ptr(), al.data_ptr(), ldp,
                                     bh.data_ptr(), bl.data_ptr(), ldp, 0.0, C2.data_ptr(), m))
    torch.cuda.synchronize()
    if m * n >= 4096 * 4096:
        assert torch.equal(C1, C2)
    else:
        oa, ob = cm(A, m, k, m).double(), cm(B, k, n, k).double()
        tol = (3.1 * 2.0 ** -16 + 1.01 * (3 * k + 2) * U) * (oa.abs() @ ob.abs())
        assert ((cm(C2, m, n, m).double() - oa @ ob).abs() <= tol + 1e-30).all()
    # the split itself: hi + lo reproduces x to 2^-16 relative; the pad is zero
    hi16, lo16 = _unslab(ah, m, ldp), _unslab(al, m, ldp)
    hi = (hi16[:, :k].int() << 16).view(torch.float32)
    lo = (lo16[:, :k].int() << 16).view(torch.float32)
    x = cm(A, m, k, m)
    assert ((hi.double() + lo.double() - x.double()).abs() <= 2.0 ** -16 * x.double().abs() + 1e-38).all()
    assert (hi16[:, k:] == 0).all() and (lo16[:, k:] == 0).all()


@pytest.mark.parametrize("ta,tb", [("n", "n"), ("t", "t")])
@pytest.mark.parametrize("m,n,k", [(96, 80, 70), (4096, 4096, 128)])
def test_gemm_bf16_inputs(handle, torch_cuda, ta, tb, m, n, k):
    """BF16 inputs are exact in FP32: only accumulation error remains."""
    import portblas_amd as pb

    torch = torch_cuda
    ar, ac = (m, k) if ta == "n" else (k, m)
    br, bc = (k, n) if tb == "n" else (n, k)
    A = make(torch, ar * ac, 31).to(torch.bfloat16)
    B = make(torch, br * bc, 32).to(torch.bfloat16)
    C = make(torch, m * n, 33)
    C0 = C.clone()
    pb._gemm_bf16(handle, ta, tb, m, n, k, 2.0, A, ar, B, br, -1.0, C, m)
    oa = cm(A, ar, ac, ar).double()
    ob = cm(B, br, bc, br).double()
    oa = oa if ta == "n" else oa.t()
    ob = ob if tb == "n" else ob.t()
    ref = 2.0 * (oa @ ob) - cm(C0, m, n, m).double()
    tol = 1.01 * (k + 2) * U * 2.0 * (oa.abs() @ ob.abs()) + 2 * U * cm(C0, m, n, m).double().abs()
    err = (cm(C, m, n, m).double() - ref).abs()
    assert (err <= tol + 1e-30).all(), (err / tol).max().item()


@pytest.mark.parametrize("dist", ["normal", "loguniform"])
def test_fp16x3_max_error_not_above_fp32(handle, torch_cuda, dist):
    """The FP16 split keeps 22 significand bits per operand, so at K = 4096
    its largest componentwise error against FP64 stays at or below the plain
    FP32 MFMA GEMM's on the same inputs with the same accumulation structure
    (64 x 64 tiles, K not split: the FP16x3 plan at this shape, as both run
    unsplit in the 8192^3 headline).  A K split shortens FP32's accumulation
    chains and can take its error below FP16x3's (1.14e-7 against 1.22e-7
    here with split-K 4); BF16x3, with 16 bits, does not match FP32 at all."""
    import portblas_amd as pb

    torch = torch_cuda
    L = pb.load()
    m = n = 2048
    k = 4096
    A = make(torch, m * k, 41, dist)
    B = make(torch, k * n, 42, dist)
    oa, ob = cm(A, m, k, m).double(), cm(B, k, n, k).double()
    ref = oa @ ob
    absprod = oa.abs() @ ob.abs()
    worst = {}
    for prec in ("fp32_mfma", "fp16x3"):
        C = torch.zeros(m * n, device="cuda")
        if prec == "fp32_mfma":  # plain config 8: 64 x 64 tiles, no split
            pb.check(L.pb_gemm_plain_config(handle.handle, 1, 8, b"n", b"n", m, n, k, 1.0, A.data_ptr(), m,
                                            B.data_ptr(), k, 0.0, C.data_ptr(), m, 1))
        else:
            pb._gemm_ex(handle, pb.gemm_precision_t[prec], "n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
        worst[prec] = ((cm(C, m, n, m).double() - ref).abs() / absprod).max().item()
    assert worst["fp16x3"] <= worst["fp32_mfma"], worst


@pytest.mark.parametrize("pos", ["first", "last"])
def test_fp16x3_scale_over_every_workgroup_and_finite_only(handle, torch_cuda, pos):
    """The FP16x3 scale comes from the amax over the finite elements, folded
    over every pre-pass workgroup: a needle of 1000 (against U[-1,1]) at the
    first or last element of a 16.8M-element operand must set it (a missed
    partial would scale the needle past 65504 into infinity), and an
    infinity elsewhere in the operand must not (it would flush the whole
    operand's scale to 2^-114).  Only the infinity's row of C is non-finite."""
    import portblas_amd as pb

    torch = torch_cuda
    m = n = 64
    k = 1 << 18
    idx = 0 if pos == "first" else -1
    A = make(torch, m * k, 51)
    B = make(torch, k * n, 52)
    A[idx] = 1000.0
    B[idx] = -1000.0
    A[m * (k // 2) + 5] = float("inf")  # row 5, column k / 2
    C = torch.zeros(m * n, device="cuda")
    pb._gemm_ex(handle, pb.gemm_precision_t.fp16x3, "n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    torch.cuda.synchronize()
    got = cm(C, m, n, m).double()
    assert not torch.isfinite(got[5]).any()
    keep = [r for r in range(m) if r != 5]
    oa, ob = cm(A, m, k, m).double()[keep], cm(B, k, n, k).double()
    ref = oa @ ob
    absprod = oa.abs() @ ob.abs()
    ea, eb = 2.0 ** -25 / f16_scale(1000.0), 2.0 ** -25 / f16_scale(1000.0)
    tol = (3.01 * 2.0 ** -22 + 1.01 * (3 * k + 2) * U) * absprod + 1.07 * (
        ea * ob.abs().sum(0, keepdim=True) + eb * oa.abs().sum(1, keepdim=True)) + k * ea * eb
    err = (got[keep] - ref).abs()
    assert torch.isfinite(err).all()
    assert (err <= tol).all(), (err / tol).max().item()


def test_fp16x3_guessed_scales_match_two_pass(torch_cuda):
    """pb_gemm_ex(FP16X3) packs with the scale exponents of the handle's
    previous FP16x3 call and packs an operand again when its amax gives
    another exponent.  On a fresh handle, over a sequence that starts with
    both guesses wrong, then right, then A's wrong, then B's, then both, every
    product equals the two-pass pack-once path (amax first) bit for bit."""
    import portblas_amd as pb

    torch = torch_cuda
    L = pb.load()
    h = pb.SB_Handle()
    m = n = 4096
    k = 256
    A0 = make(torch, m * k, 71)
    B0 = make(torch, k * n, 72)
    ldp = k
    ah = torch.empty(m * ldp, dtype=torch.int16, device="cuda")
    al, bh, bl = torch.empty_like(ah), torch.empty_like(ah), torch.empty_like(ah)
    fac = torch.zeros(2, device="cuda")
    seq = [(1.0, 1.0), (1.0, 1.0), (2.0 ** 20, 1.0), (2.0 ** 20, 2.0 ** -30), (1.0, 1.0), (1.0, 1.0)]
    for i, (fa, fb) in enumerate(seq):
        A, B = A0 * fa, B0 * fb
        C1 = torch.zeros(m * n, device="cuda")
        C2 = torch.zeros(m * n, device="cuda")
        pb._gemm_ex(h, pb.gemm_precision_t.fp16x3, "n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C1, m)
        pb.check(L.pb_pack_fp16x3(h.handle, 0, b"n", m, k, A.data_ptr(), m, ah.data_ptr(), al.data_ptr(), ldp,
                                  fac.data_ptr()))
        pb.check(L.pb_pack_fp16x3(h.handle, 1, b"n", n, k, B.data_ptr(), k, bh.data_ptr(), bl.data_ptr(), ldp,
                                  fac.data_ptr() + 4))
        pb.check(L.pb_gemm_fp16x3_packed(h.handle, m, n, k, 1.0, ah.data_ptr(), al.data_ptr(), fac.data_ptr(),
                                         ldp, bh.data_ptr(), bl.data_ptr(), fac.data_ptr() + 4, ldp, 0.0,
                                         C2.data_ptr(), m))
        torch.cuda.synchronize()
        assert torch.isfinite(C1).all(), i
        assert torch.equal(C1, C2), (i, (C1 - C2).abs().max().item())


@pytest.mark.parametrize("m,n,k", [(4096, 4096, 512), (300, 200, 520)])
def test_fp8_power_of_two_scale_exact_on_integers(handle, torch_cuda, m, n, k):
    """The FP8 scale is the largest power of two 2^s with amax 2^s <= 448:
    integers in [-3, 3] (amax 3, s = 7) become 0, +-128, +-256, +-384, all
    exact in e4m3, so the product is the exact integer product.  A scale of
    448 / amax would map 1 to 149.33 and round it."""
    import portblas_amd as pb

    torch = torch_cuda
    g = torch.Generator(device="cuda").manual_seed(91)
    A = torch.randint(-3, 4, (m * k,), generator=g, device="cuda").float()
    B = torch.randint(-3, 4, (k * n,), generator=g, device="cuda").float()
    A[0], B[0] = 3.0, -3.0
    C = torch.zeros(m * n, device="cuda")
    pb._gemm_ex(handle, pb.gemm_precision_t.fp8_e4m3, "n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    torch.cuda.synchronize()
    ref = cm(A, m, k, m).double() @ cm(B, k, n, k).double()
    assert torch.equal(cm(C, m, n, m).double(), ref)


def test_fp8_guessed_scales_match_fresh_handle(torch_cuda):
    """pb_gemm_ex(FP8) packs with the power-of-two scales of the handle's
    previous FP8 call and packs an operand again when its amax gives another
    one.  Over a sequence of hits and misses on one handle, every product
    equals the product on a fresh handle (whose first call always takes the
    scales from the amax it measures), bit for bit."""
    import portblas_amd as pb

    torch = torch_cuda
    h = pb.SB_Handle()
    m = n = 4096
    k = 256
    A0 = make(torch, m * k, 81)
    B0 = make(torch, k * n, 82)
    seq = [(1.0, 1.0), (1.0, 1.0), (2.0 ** 20, 1.0), (2.0 ** 20, 2.0 ** -30), (3.0, 1.0), (1.0, 1.0)]
    for i, (fa, fb) in enumerate(seq):
        A, B = A0 * fa, B0 * fb
        C1 = torch.zeros(m * n, device="cuda")
        C2 = torch.zeros(m * n, device="cuda")
        pb._gemm_ex(h, pb.gemm_precision_t.fp8_e4m3, "n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C1, m)
        pb._gemm_ex(pb.SB_Handle(), pb.gemm_precision_t.fp8_e4m3, "n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C2, m)
        torch.cuda.synchronize()
        assert torch.isfinite(C1).all(), i
        assert torch.equal(C1, C2), (i, (C1 - C2).abs().max().item())


@pytest.mark.parametrize("m,n,k", [(4096, 4096, 512), (300, 200, 100)])
def test_pack_once_fp16x3_matches_gemm_ex(handle, torch_cuda, m, n, k):
    """pb_pack_fp16x3 + pb_gemm_fp16x3_packed: the factor is 2^-s with
    amax 2^s in [2^14, 2^15), hi + lo reproduces x 2^s to 2^-22 (relative)
    plus 2^-25 (absolute), the pad is zero, and the product matches
    pb_gemm_ex(FP16X3) bit for bit where both run the 256 x 256 kernel."""
    import portblas_amd as pb

    torch = torch_cuda
    L = pb.load()
    A = make(torch, m * k, 61, "loguniform")
    B = make(torch, k * n, 62)
    C1 = torch.zeros(m * n, device="cuda")
    C2 = torch.zeros(m * n, device="cuda")
    pb._gemm_ex(handle, pb.gemm_precision_t.fp16x3, "n", "n", m, n, k, 1.0, A, m, B, k, 0.0, C1, m)
    ldp = ((k + 63) // 64) * 64
    ah = torch.empty(m * ldp, dtype=torch.int16, device="cuda")
    al = torch.empty_like(ah)
    bh = torch.empty(n * ldp, dtype=torch.int16, device="cuda")
    bl = torch.empty_like(bh)
    fac = torch.zeros(2, device="cuda")
    pb.check(L.pb_pack_fp16x3(handle.handle, 0, b"n", m, k, A.data_ptr(), m, ah.data_ptr(), al.data_ptr(), ldp,
                              fac.data_ptr()))
    pb.check(L.pb_pack_fp16x3(handle.handle, 1, b"n", n, k, B.data_ptr(), k, bh.data_ptr(), bl.data_ptr(), ldp,
                              fac.data_ptr() + 4))
    pb.check(L.pb_gemm_fp16x3_packed(handle.handle, m, n, k, 1.0, ah.data_ptr(), al.data_ptr(), fac.data_ptr(),
                                     ldp, bh.data_ptr(), bl.data_ptr(), fac.data_ptr() + 4, ldp, 0.0,
                                     C2.data_ptr(), m))
    torch.cuda.synchronize()
    sa, sb = f16_scale(A.abs().max().item()), f16_scale(B.abs().max().item())
    assert fac.tolist() == [1.0 / sa, 1.0 / sb]
    if m * n >= 4096 * 4096:
        assert torch.equal(C1, C2)
    else:
        oa, ob = cm(A, m, k, m).double(), cm(B, k, n, k).double()
        tol = (3.01 * 2.0 ** -22 + 1.01 * (3 * k + 2) * U) * (oa.abs() @ ob.abs()) + 1.07 * (
            2.0 ** -25 / sa * ob.abs().sum(0, keepdim=True) + 2.0 ** -25 / sb * oa.abs().sum(1, keepdim=True))
        assert ((cm(C2, m, n, m).double() - oa @ ob).abs() <= tol + 1e-30).all()
    hi16, lo16 = _unslab(ah, m, ldp), _unslab(al, m, ldp)
    hi = hi16[:, :k].view(torch.float16).double()
    lo = lo16[:, :k].view(torch.float16).double()
    xs = cm(A, m, k, m).double() * sa
    assert ((hi + lo - xs).abs() <= 2.0 ** -22 * xs.abs() + 2.0 ** -25).all()
    assert (hi16[:, k:] == 0).all() and (lo16[:, k:] == 0).all()
