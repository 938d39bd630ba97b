"""Mixed-precision path on the GPU: the FP32-storage update GEMM in its
three precisions against the numpy model of its products (oracle/mixed.py,
pinned against torch's bfloat16 conversion), the triangle conversions, the
FP32-factor solve kernels, and the whole mixed solve with iterative
refinement against an FP64 solve.

Tolerances: FP32 accumulation over k terms is within 1e-5 of the model
relative to |A|^T |B|; integer inputs small enough to be exact in bf16 and
in FP32 sums are bit-exact; the refined solution reaches normwise backward
error <= 1e-14, the FP64 bar."""
import os
import tempfile

import numpy as np
import pytest

import oracle
from oracle.mixed import gemm_tn_model

PERF_MODEL_DIR = os.path.join(tempfile.gettempdir(), "starpu_amd_perfmodels")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def K():
    from starpu_amd import kernels
    if kernels.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return kernels


def dev(K, a, dtype=np.float64, ld=None):
    return K.DeviceArray.from_host(np.asfortranarray(np.asarray(a, dtype=dtype)), ld=ld)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("m,n,k", [(1024, 1024, 1024), (256, 128, 64), (100, 70, 33), (64, 64, 4096)])
def test_sgemm_tn_matches_model(K, mode, m, n, k):
    print(f"sgemm_tn mode={mode} m={m} n={n} k={k}", flush=True)
    rng = np.random.default_rng(m + n + k + mode)
    A = rng.uniform(-1, 1, (k, m)).astype(np.float32)
    B = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    C0 = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    dA, dB, dC = dev(K, A, np.float32), dev(K, B, np.float32), dev(K, C0, np.float32)
    K.sgemm_tn(mode, 0, m, n, k, -1.0, dA.ptr, dA.ld, dB.ptr, dB.ld, 1.0, dC.ptr, dC.ld)
    got = dC.download().astype(np.float64)
    exp = C0.astype(np.float64) - gemm_tn_model(mode, A, B)
    scale = np.abs(A).T.astype(np.float64) @ np.abs(B).astype(np.float64) + 1.0
    assert np.max(np.abs(got - exp) / scale) < 1e-5


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_sgemm_tn_integer_exact_and_upper_mask(K, mode):
    rng = np.random.default_rng(40 + mode)
    n, k = 512, 256
    A = rng.integers(-16, 17, (k, n)).astype(np.float32)     # exact in bf16; sums < 2^24
    C0 = rng.integers(-64, 65, (n, n)).astype(np.float32)
    dA, dC = dev(K, A, np.float32), dev(K, C0, np.float32)
    K.sgemm_tn(mode, 1, n, n, k, -1.0, dA.ptr, dA.ld, dA.ptr, dA.ld, 1.0, dC.ptr, dC.ld)
    got = dC.download()
    exp = C0 - A.T @ A
    iu = np.triu_indices(n)
    il = np.tril_indices(n, -1)
    assert np.array_equal(got[iu], exp[iu])
    assert np.array_equal(got[il], C0[il])                    # strict lower part untouched


@pytest.mark.parametrize("upper,m,n,k", [(0, 1024, 1024, 1024), (1, 1024, 1024, 1024), (0, 512, 384, 96),
                                         (1, 256, 256, 64)])
def test_bf16x3_split_planes_match_fp32_operands(K, upper, m, n, k):
    """The BF16x3 product on FP32 operands (split inside the kernel) and the
    product on planes split ahead of time, staged straight to LDS or through
    registers, apply the same MFMAs to each accumulator in the same order:
    equal bits, also in the upper-triangle form."""
    print(f"bf16x3 split planes upper={upper} m={m} n={n} k={k}", flush=True)
    L = K.lib()
    rng = np.random.default_rng(m + n + k + upper)
    A = rng.standard_normal((k, m)).astype(np.float32)
    B = A if upper else rng.standard_normal((k, n)).astype(np.float32)
    C0 = rng.standard_normal((m, n)).astype(np.float32)
    dA, dB = dev(K, A, np.float32), dev(K, B, np.float32)
    planes = []
    for d, cols in ((dA, m), (dB, n)):
        h, lo = K.DeviceArray(k, cols, np.uint16), K.DeviceArray(k, cols, np.uint16)
        K.check(L.sa_split_bf16x2(None, k, cols, d.ptr, d.ld, h.ptr, lo.ptr, k))
        planes.append((h, lo))
    dC1 = dev(K, C0, np.float32)
    K.sgemm_tn(2, upper, m, n, k, -1.0, dA.ptr, dA.ld, dB.ptr, dB.ld, 1.0, dC1.ptr, dC1.ld)
    (ah, al), (bh, bl) = planes
    assert L.sa_sgemm_tn_split_ok(m, n, k, ah.ptr, al.ptr, k, bh.ptr, bl.ptr, k) == 1
    got1 = dC1.download()
    # both stagings of the split planes: straight to LDS (the default) and
    # through registers
    for direct in (1, 0):
        prev = L.sa_set_sgemm_direct(direct)
        try:
            dC2 = dev(K, C0, np.float32)
            K.check(L.sa_sgemm_tn_split(None, upper, m, n, k, -1.0, ah.ptr, al.ptr, k, bh.ptr, bl.ptr, k, 1.0,
                                        dC2.ptr, dC2.ld))
            assert np.array_equal(got1, dC2.download()), direct
        finally:
            L.sa_set_sgemm_direct(prev)
    assert L.sa_set_sgemm_direct(2) == -22
    if upper:
        il = np.tril_indices(m, -1)
        assert np.array_equal(got1[il], C0[il])


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("m,n", [(1024, 1024), (1024, 384), (256, 128), (100, 70)])
def test_triangular_lhs_product_equals_full_product(K, mode, m, n):
    """sa_sgemm_tn_utri (U upper triangular, zeros stored below) stops each
    128-row block's k-range at its last row: the same values as the full
    product sa_sgemm_tn, whose extra terms are exact zeros (m = 100: the
    unstaged kernels, which take the full range)."""
    L = K.lib()
    rng = np.random.default_rng(m + n + mode)
    U = np.triu(rng.uniform(-1, 1, (m, m))).astype(np.float32)
    B = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    C0 = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    dU, dB = dev(K, U, np.float32), dev(K, B, np.float32)
    dC1, dC2 = dev(K, C0, np.float32), dev(K, C0, np.float32)
    K.sgemm_tn(mode, 0, m, n, m, -1.0, dU.ptr, dU.ld, dB.ptr, dB.ld, 1.0, dC1.ptr, dC1.ld)
    K.check(L.sa_sgemm_tn_utri(None, mode, m, n, -1.0, dU.ptr, dU.ld, dB.ptr, dB.ld, 1.0, dC2.ptr, dC2.ld))
    assert np.array_equal(dC1.download(), dC2.download())


@pytest.mark.parametrize("m,n,lda,ldb", [(1024, 1024, 32768, 1024), (1, 1, 1, 1), (1001, 7, 1003, 1001),
                                         (8, 3, 8, 12), (333, 70000, 333, 336)])
def test_matrix_copies(K, m, n, lda, ldb):
    """sa_dlacpy / sa_slacpy (16-byte vectors when the columns are so
    aligned, element by element otherwise; more than 65535 columns in
    several launches) copy exactly A(0:m, 0:n) and nothing else."""
    L = K.lib()
    rng = np.random.default_rng(m + n + lda)
    for dt, fn in ((np.float64, L.sa_dlacpy), (np.float32, L.sa_slacpy)):
        A = rng.standard_normal((lda, n)).astype(dt)
        B0 = rng.standard_normal((ldb, n)).astype(dt)
        dA, dB = dev(K, A, dt), dev(K, B0, dt)
        K.check(fn(None, m, n, dA.ptr, dA.ld, dB.ptr, dB.ld))
        got = dB.download()
        assert np.array_equal(got[:m], A[:m]), dt
        assert np.array_equal(got[m:], B0[m:]), dt


@pytest.mark.parametrize("n", [1, 63, 300, 1024, 2050])
def test_conversion_with_frobenius_norm(K, n):
    """The fused FP64-lower to FP32-upper conversion writes the same bits as
    the plain one and ||A||_F^2 of the symmetric matrix (diagonal once, the
    strict lower triangle twice) to FP64 rounding."""
    print(f"dlag2s fro n={n}", flush=True)
    L = K.lib()
    rng = np.random.default_rng(n + 11)
    M = rng.standard_normal((n, n))
    A = M + M.T
    dA = dev(K, A)
    dU1 = dev(K, np.full((n, n), 7.0, np.float32), np.float32)
    dU2 = dev(K, np.full((n, n), 7.0, np.float32), np.float32)
    K.check(L.sa_dlag2s_lower_to_upper(None, n, dA.ptr, dA.ld, dU1.ptr, dU1.ld))
    wb = L.sa_dlag2s_lower_to_upper_fro_worksize(n)
    ws = K.DeviceArray(wb, 1, np.uint8)
    fro = K.DeviceArray(1, 1)
    K.check(L.sa_dlag2s_lower_to_upper_fro(None, n, dA.ptr, dA.ld, dU2.ptr, dU2.ld, fro.ptr, ws.ptr, wb))
    assert np.array_equal(dU1.download(), dU2.download())
    got = float(fro.download()[0, 0])
    exp = float(np.sum(A * A))
    assert abs(got - exp) <= 1e-13 * exp


# whole 64x64 blocks move as 16-byte vectors when both leading dimensions
# allow it (even for FP64, multiples of 4 for FP32), edge blocks and odd
# leading dimensions element by element
@pytest.mark.parametrize("n,lda,ldu", [(300, 300, 300), (257, 257, 257), (1024, 1024, 1024), (1024, 1027, 1030),
                                       (130, 136, 132), (64, 64, 64), (1, 1, 1)])
def test_triangle_conversions(K, n, lda, ldu):
    print(f"triangle conversions n={n} lda={lda} ldu={ldu}", flush=True)
    rng = np.random.default_rng(5 + n)
    A = rng.standard_normal((n, n))
    dA = dev(K, A, ld=lda)
    dU = dev(K, np.full((n, n), 7.0), np.float32, ld=ldu)
    K.check(K.lib().sa_dlag2s_lower_to_upper(None, n, dA.ptr, dA.ld, dU.ptr, dU.ld))
    U = dU.download()
    iu = np.triu_indices(n)
    assert np.array_equal(U[iu], A.T.astype(np.float32)[iu])  # U(j,i) = float(A(i,j)), i >= j
    assert np.all(U[np.tril_indices(n, -1)] == 7.0)
    dL = dev(K, np.full((n, n), -3.0), ld=lda)
    K.check(K.lib().sa_slag2d_upper_to_lower(None, n, dU.ptr, dU.ld, dL.ptr, dL.ld))
    L = dL.download()
    il = np.tril_indices(n)
    assert np.array_equal(L[il], U.T.astype(np.float64)[il])
    assert np.all(L[np.triu_indices(n, 1)] == -3.0)


@pytest.mark.parametrize("n,ld", [(1, 1), (300, 300), (1000, 1003), (1024, 1024)])
def test_tile_outputs_one_launch(K, n, ld):
    """The mixed tile POTRF's FP32 outputs in one launch: the same bits as
    the triangle conversion and the plain conversion it replaces."""
    print(f"tile outputs n={n} ld={ld}", flush=True)
    rng = np.random.default_rng(n + 3)
    Lh, Wh = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    dL, dW = dev(K, Lh, ld=ld), dev(K, Wh, ld=ld)
    U1 = dev(K, np.full((n, n), 5.0), np.float32, ld=ld)
    U2 = dev(K, np.full((n, n), 5.0), np.float32, ld=ld)
    W1 = dev(K, np.zeros((n, n)), np.float32, ld=ld)
    W2 = dev(K, np.zeros((n, n)), np.float32, ld=ld)
    L = K.lib()
    K.check(L.sa_dlag2s_lower_to_upper(None, n, dL.ptr, ld, U1.ptr, ld))
    K.check(L.sa_dlag2s(None, n, n, dW.ptr, ld, W1.ptr, ld))
    K.check(L.sa_tile_outputs_f32(None, n, dL.ptr, ld, U2.ptr, ld, dW.ptr, ld, W2.ptr, ld))
    assert np.array_equal(U1.download(), U2.download())
    assert np.array_equal(W1.download(), W2.download())
    assert np.array_equal(W2.download(), Wh.astype(np.float32))


@pytest.mark.parametrize("m,n,k", [(1024, 16, 1024), (1000, 5, 3000), (4096, 16, 1024), (2052, 7, 520),
                                   (64, 32, 4096), (600, 33, 1004), (5000, 16, 1024), (4098, 3, 1024),
                                   (4100, 33, 40)])
def test_fp32_factor_solve_kernels(K, m, n, k):
    """The thin products of the refinement solves (FP32 factor, FP64
    right-hand sides) against numpy in FP64: one workgroup per 16 rows when
    k % 8 == 0 and m < 4096, per 64 rows from 4096 rows (a partial last
    workgroup at m = 5000 and 4100; NN with m % 4 != 0 stays on 16-row
    workgroups, m = 4098; waves without k steps at k = 40), the split-k
    forms otherwise (k = 1004), right-hand sides in chunks of 16 (n = 32,
    33)."""
    rng = np.random.default_rng(m + k)
    L = K.lib()
    A = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    At = rng.uniform(-1, 1, (k, m)).astype(np.float32)
    B = rng.standard_normal((k, n))
    C0 = rng.standard_normal((m, n))
    dA, dAt, dB = dev(K, A, np.float32), dev(K, At, np.float32), dev(K, B)
    ws = K.DeviceArray(1 << 20, 8)
    for kind, Am, dAm in (("nn", A, dA), ("tn", At, dAt)):
        dC = dev(K, C0)
        fn = L.sa_gemm_nn_sd if kind == "nn" else L.sa_gemm_tn_sd
        K.check(fn(None, m, n, k, -1.0, dAm.ptr, dAm.ld, dB.ptr, dB.ld, 1.0, dC.ptr, dC.ld, ws.ptr, ws.nbytes))
        opA = Am.astype(np.float64) if kind == "nn" else Am.astype(np.float64).T
        exp = C0 - opA @ B
        assert np.linalg.norm(dC.download() - exp) / np.linalg.norm(exp) < 1e-13
    U = np.triu(rng.uniform(-1, 1, (m, m))).astype(np.float32)
    dU = dev(K, U, np.float32)
    for trans in (0, 1):
        dX = dev(K, C0)
        K.check(L.sa_trmm_lu_sd(None, trans, m, n, 1.0, dU.ptr, dU.ld, dX.ptr, dX.ld, ws.ptr, ws.nbytes))
        U64 = U.astype(np.float64)
        exp = (U64.T if trans else U64) @ C0
        assert np.linalg.norm(dX.download() - exp) / np.linalg.norm(exp) < 1e-13


@pytest.mark.parametrize("n,r", [(1, 1), (17, 3), (255, 16), (256, 16), (300, 17), (1000, 5), (2049, 16),
                                 (4500, 16), (8192, 16)])
def test_symmetric_residual_reads_lower_triangle_only(K, n, r):
    """R = B - A X from A's lower triangle (sa_dsymm_lower_resid) against
    numpy on the symmetric matrix, with NaN in the strict upper triangle so
    that any read of it shows; uneven sizes (partial row blocks, chunks and
    tiles), right-hand sides in chunks of 16 (r = 17), several runs of
    column chunks per row block (n >= 2049)."""
    print(f"symmetric residual n={n} r={r}", flush=True)
    rng = np.random.default_rng(n + r)
    L = K.lib()
    S = rng.standard_normal((n, n))
    S = np.tril(S) + np.tril(S, -1).T
    Ast = S.copy()
    Ast[np.triu_indices(n, 1)] = np.nan
    X = rng.standard_normal((n, r))
    Bm = rng.standard_normal((n, r))
    dA, dX, dB = dev(K, Ast), dev(K, X), dev(K, Bm)
    dR = K.DeviceArray(n, r)
    ws = L.sa_dsymm_lower_resid_worksize(n, r)
    W = K.DeviceArray(max(1, ws // 8 + 1), 1)
    for rep in range(2):            # the second call reuses the workspace as left
        K.check(L.sa_dsymm_lower_resid(None, n, r, dA.ptr, dA.ld, dX.ptr, dX.ld, dB.ptr, dB.ld, dR.ptr, dR.ld,
                                       W.ptr, W.nbytes))
        got = dR.download()
        exp = Bm - S @ X
        scale = np.abs(S).sum(axis=1, keepdims=True) * np.abs(X).max() + np.abs(Bm)
        assert np.isfinite(got).all(), rep
        assert (np.abs(got - exp) <= 1e-13 * scale).all(), (rep, float(np.abs(got - exp).max()))
        if rep == 0:
            first = got.copy()
        else:
            assert np.array_equal(first, got)      # fixed summation order
    assert L.sa_dsymm_lower_resid(None, n, r, dA.ptr, dA.ld, dX.ptr, dX.ld, dB.ptr, dB.ld, dR.ptr, dR.ld,
                                  W.ptr, ws - 1) == -22


@pytest.fixture(scope="module")
def rt(K):
    from starpu_amd import runtime as R
    R.init(STARPU_NHIP=1, STARPU_NCPU=0, STARPU_SCHED="dmda", STARPU_PERF_MODEL_DIR=PERF_MODEL_DIR)
    R.set_cpu_kernels()
    yield R
    R.shutdown()


def spd_device(K, n, shift, seed):
    """A = B B^T / n + shift I on the device (both triangles)."""
    L = K.lib()
    A = K.DeviceArray(n, n)
    Bm = K.DeviceArray(n, n)
    K.check(L.sa_fill_random(None, n, n, Bm.ptr, n, seed, 1))
    K.dsyrk_ln(n, n, 1.0 / n, Bm.ptr, n, 0.0, A.ptr, n)
    K.check(L.sa_symmetrize_lower(None, n, A.ptr, n))
    K.check(L.sa_add_diagonal(None, n, A.ptr, n, shift))
    K.sync()
    return A


@pytest.mark.parametrize("prec,shift", [("fp32", 1e-3), ("bf16x3", 1e-3), ("bf16", 1.0)])
@pytest.mark.parametrize("n,nblocks", [(2048, 8), (1536, 3), (1000, 4)])
def test_mixed_solve_reaches_fp64_backward_error(K, rt, prec, shift, n, nblocks):
    R = rt
    print(f"mixed solve prec={prec} shift={shift} n={n} nblocks={nblocks}", flush=True)
    nrhs = 16
    dA = spd_device(K, n, shift, seed=n + nblocks)
    rng = np.random.default_rng(n)
    b = rng.standard_normal((n, nrhs))
    dB = dev(K, b)
    dX = K.DeviceArray(n, nrhs)
    st = R.mixed_solve(dA.ptr, n, n, nblocks, dB.ptr, n, nrhs, dX.ptr, n, prec=prec, tol=1e-14, maxiter=40)
    assert st.info == 0
    assert st.converged == 1, (st.final_berr, st.iterations, list(st.berr[:st.iterations]))
    assert st.final_berr <= 1e-14
    A = dA.download()
    X = dX.download()
    Xref = np.linalg.solve(A, b)
    # forward error is bounded by cond(A) times the backward error
    cond = np.linalg.cond(A)
    assert np.linalg.norm(X - Xref) / np.linalg.norm(Xref) < 10 * cond * 1e-14
    # the backward error reported matches one recomputed on the host
    r = b - A @ X
    berr = max(np.linalg.norm(r[:, j]) / (np.linalg.norm(A) * np.linalg.norm(X[:, j]) + np.linalg.norm(b[:, j]))
               for j in range(nrhs))
    assert berr <= 2e-14
    # the input matrix is not modified
    assert np.array_equal(A, dA.download())


def test_mixed_solve_reads_lower_triangle_only(K, rt):
    """The whole mixed solve (conversion, factorisation, refinement) on a
    matrix whose strict upper triangle holds NaN reaches the FP64 backward
    error of the symmetric matrix its lower triangle stands for."""
    R = rt
    n, nblocks, nrhs = 1536, 3, 16
    dS = spd_device(K, n, 1e-3, seed=77)
    S = dS.download()
    Ast = S.copy()
    Ast[np.triu_indices(n, 1)] = np.nan
    dA = dev(K, Ast)
    b = np.random.default_rng(77).standard_normal((n, nrhs))
    dB, dX = dev(K, b), K.DeviceArray(n, nrhs)
    st = R.mixed_solve(dA.ptr, n, n, nblocks, dB.ptr, n, nrhs, dX.ptr, n, prec="bf16x3", tol=1e-14, maxiter=40)
    assert st.info == 0 and st.converged == 1 and st.final_berr <= 1e-14
    X = dX.download()
    assert np.isfinite(X).all()
    r = b - S @ X
    berr = max(np.linalg.norm(r[:, j]) / (np.linalg.norm(S) * np.linalg.norm(X[:, j]) + np.linalg.norm(b[:, j]))
               for j in range(nrhs))
    assert berr <= 2e-14


@pytest.mark.parametrize("n,nblocks", [(2048, 8), (1536, 3)])
def test_bf16x3_presplit_panels_bit_identical(K, rt, n, nblocks, monkeypatch):
    # SA_MIXED_PRESPLIT=1 on tiles of multiples of 128: the TRSMs write the
    # panels' bf16 planes and the updates read them; the result must equal,
    # bit for bit, the default run that splits inside every product
    R = rt
    print(f"bf16x3 presplit n={n} nblocks={nblocks}", flush=True)
    nrhs = 4
    dA = spd_device(K, n, 1e-3, seed=n + 3)
    b = np.random.default_rng(n + 1).standard_normal((n, nrhs))
    dB = dev(K, b)
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("SA_MIXED_PRESPLIT", flag)
        dX = K.DeviceArray(n, nrhs)
        st = R.mixed_solve(dA.ptr, n, n, nblocks, dB.ptr, n, nrhs, dX.ptr, n, prec="bf16x3", tol=1e-14, maxiter=40)
        assert st.info == 0 and st.converged == 1
        out.append((dX.download(), st.iterations))
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])


def test_split_bf16x2_planes(K):
    # H = bf16(x) (round to nearest even), L = bf16(x - H): H + L carries
    # about 16 significant bits of x, and both planes are exact bf16 values
    L = K.lib()
    m, n = 300, 7
    x = np.random.default_rng(5).standard_normal((m, n)).astype(np.float32)
    dx = dev(K, x, np.float32)
    dh = K.DeviceArray(m, n, np.uint16)
    dl = K.DeviceArray(m, n, np.uint16)
    K.check(L.sa_split_bf16x2(None, m, n, dx.ptr, dx.ld, dh.ptr, dl.ptr, m))
    h, lo = dh.download(), dl.download()
    hf = (h.astype(np.uint32) << 16).view(np.float32)
    lf = (lo.astype(np.uint32) << 16).view(np.float32)
    # H is x rounded to 8 significant bits: |x - H| <= 2^-8 |x|
    assert np.all(np.abs(x - hf) <= np.abs(x) * 2.0 ** -8)
    # L is the rounded remainder: |x - H - L| <= 2^-8 |x - H|
    rem = (x - hf).astype(np.float64)
    assert np.all(np.abs(rem - lf) <= np.abs(rem) * 2.0 ** -8 + 1e-45)


def test_mixed_solve_reports_non_spd_pivot(K, rt):
    R = rt
    n, nb = 512, 4
    A = np.eye(n) * 2.0
    A[300, 300] = -1.0
    dA = dev(K, A)
    dB = dev(K, np.ones((n, 2)))
    dX = K.DeviceArray(n, 2)
    st = R.mixed_solve(dA.ptr, n, n, nb, dB.ptr, n, 2, dX.ptr, n, prec="fp32")
    assert st.info == 301
