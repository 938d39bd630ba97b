"""GPU parity of the hand-written gfx950 tile kernels against the CPU oracle
and the LAPACK golden vectors, called through the C ABI of
include/starpu_amd_kernels.h.

Tolerances: FP64 results within 1e-13 relative (Frobenius) of the oracle;
integer-valued GEMM/SYRK inputs must match bit for bit (every partial sum is
an exactly representable integer)."""
import os

import numpy as np
from scipy.linalg import solve_triangular
import pytest

import oracle

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "lapack_vectors.npz"))


@pytest.fixture(scope="module")
def K():
    from starpu_amd import kernels
    if kernels.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return kernels


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def dev(K, a, ld=None):
    return K.DeviceArray.from_host(np.asfortranarray(a), ld=ld)


@pytest.fixture(params=[64, 128], ids=["wg64", "wg128"])
def shape(K, request):
    """Forces the GEMM workgroup tile (64x64 or 128x128) for one test."""
    prev = K.set_gemm_shape(request.param)
    assert K.gemm_tile_shape(False, 1024, 1024) == request.param
    yield request.param
    K.set_gemm_shape(prev)


@pytest.mark.parametrize("m,n,k", [(128, 128, 32), (256, 384, 64), (1024, 1024, 1024),
                                   (1, 1, 1), (7, 5, 3), (130, 129, 70), (300, 200, 17),
                                   (64, 256, 48), (129, 127, 0), (1000, 1000, 1000), (257, 1023, 129)])
@pytest.mark.parametrize("alpha,beta", [(-1.0, 1.0), (0.5, 0.0), (1.0, -2.0)])
def test_dgemm_nt_matches_oracle(K, shape, m, n, k, alpha, beta):
    print(f"dgemm_nt wg={shape} m={m} n={n} k={k} alpha={alpha} beta={beta}", flush=True)
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (n, k))
    C0 = rng.uniform(-1, 1, (m, n))
    if beta == 0.0:
        C0[0, 0] = np.nan  # beta == 0 must not propagate NaNs from C
    dA, dB, dC = dev(K, A), dev(K, B), dev(K, C0)
    K.dgemm_nt(m, n, k, alpha, dA.ptr, dA.ld, dB.ptr, dB.ld, beta, dC.ptr, dC.ld)
    got = dC.download()
    exp = oracle.dgemm_nt(alpha, A, B, beta, np.nan_to_num(C0) if beta == 0.0 else C0)
    assert np.all(np.isfinite(got))
    assert rel(got, exp) < 1e-13


def test_dgemm_nt_integer_inputs_bit_exact(K, shape):
    rng = np.random.default_rng(5)
    m, n, k = 256, 384, 512
    A = rng.integers(-16, 17, (m, k)).astype(float)
    B = rng.integers(-16, 17, (n, k)).astype(float)
    C0 = rng.integers(-64, 65, (m, n)).astype(float)
    dA, dB, dC = dev(K, A), dev(K, B), dev(K, C0)
    K.dgemm_nt(m, n, k, -1.0, dA.ptr, m, dB.ptr, n, 1.0, dC.ptr, m)
    assert np.array_equal(dC.download(), C0 - A @ B.T)


def test_dgemm_nt_submatrix_odd_ld_and_unaligned(K, shape):
    # tiles inside a bigger matrix with an odd leading dimension: exercises
    # the unaligned (scalar-load) path; the rest of the matrix must not move.
    rng = np.random.default_rng(6)
    big = rng.uniform(-1, 1, (301, 301))
    d = dev(K, big)
    m, n, k = 100, 90, 80
    # A at (1,0), B at (150, 10), C at (3, 200)
    K.dgemm_nt(m, n, k, -1.0, d.at(1, 0), d.ld, d.at(150, 10), d.ld, 1.0, d.at(3, 200), d.ld)
    got = d.download()
    exp = big.copy()
    exp[3:3 + m, 200:200 + n] -= big[1:1 + m, 0:k] @ big[150:150 + n, 10:10 + k].T
    assert rel(got, exp) < 1e-14
    mask = np.ones_like(big, dtype=bool)
    mask[3:3 + m, 200:200 + n] = False
    assert np.array_equal(got[mask], big[mask])


@pytest.mark.parametrize("n,k", [(128, 64), (256, 256), (1024, 1024), (5, 3), (130, 70), (1, 9),
                                 (1000, 1000), (385, 200)])
def test_dsyrk_ln_lower_only(K, shape, n, k):
    print(f"dsyrk_ln wg={shape} n={n} k={k}", flush=True)
    rng = np.random.default_rng(n + k)
    A = rng.uniform(-1, 1, (n, k))
    C0 = rng.uniform(-1, 1, (n, n))
    dA, dC = dev(K, A), dev(K, C0)
    K.dsyrk_ln(n, k, -1.0, dA.ptr, dA.ld, 1.0, dC.ptr, dC.ld)
    got = dC.download()
    exp = oracle.dsyrk_ln(-1.0, A, 1.0, C0)
    assert rel(np.tril(got), np.tril(exp)) < 1e-13
    iu = np.triu_indices(n, 1)
    assert np.array_equal(got[iu], C0[iu])


@pytest.mark.parametrize("dims", [(5, 3), (70, 40), (65, 33)])
def test_dsyrk_ln_golden(K, shape, dims):
    n, k = dims
    p = f"syrk_{n}_{k}_"
    dA, dC = dev(K, GOLD[p + "A"]), dev(K, GOLD[p + "C"])
    K.dsyrk_ln(n, k, -1.0, dA.ptr, dA.ld, 1.0, dC.ptr, dC.ld)
    assert rel(np.tril(dC.download()), np.tril(GOLD[p + "R"])) < 1e-14


@pytest.mark.parametrize("op", ["gemm", "syrk", "trsm_gemm"])
@pytest.mark.parametrize("concurrency", [1, 12])
def test_tile_updates_at_the_headline_shape(K, op, concurrency):
    """The three tile updates of the N=32768, tile-1024 factorisation exactly
    as its codelets issue them: 1024^3 products on views inside a
    32768-row matrix (ld = 32768).  With 12 launches in flight, as the
    benchmark keeps them, the 128x128 workgroup kernels run; alone, the
    64x64 ones.  GEMM: C -= A B^T (cholesky_kernels.c:75-76); SYRK: lower
    C -= A A^T (:169-170); TRSM: A_mk = T W^T with beta = 0, T the copy of
    A_mk and W = L_kk^-1 (the inverse form of :254), over the nonzero
    k-range of the triangular W (sa_dgemm_nt_ltri) and over all of it."""
    nb, ld = 1024, 32768
    prev_c = K.set_launch_concurrency(concurrency)
    prev_s = K.set_gemm_shape(0)
    try:
        want = 128 if concurrency == 12 else 64
        assert K.gemm_tile_shape(op == "syrk", nb, nb) == want
        print(f"headline {op} concurrency={concurrency} wg={want}", flush=True)
        rng = np.random.default_rng(concurrency * 3 + len(op))
        big = np.zeros((ld, 2 * nb))
        ra, rb, rc = 4096, 17408, 30720   # row offsets of the three tiles (multiples of nb)
        big[ra:ra + nb, 0:nb] = rng.uniform(-1, 1, (nb, nb))
        big[rb:rb + nb, 0:nb] = rng.uniform(-1, 1, (nb, nb))
        big[rc:rc + nb, nb:2 * nb] = rng.uniform(-1, 1, (nb, nb))
        if op == "trsm_gemm":
            # W = L^-1 of a well-conditioned lower tile, the TRTRI task's output
            Lt = np.tril(rng.uniform(-1, 1, (nb, nb)), -1) / nb + np.diag(rng.uniform(1, 2, nb))
            big[rb:rb + nb, 0:nb] = solve_triangular(Lt, np.eye(nb), lower=True)
        A = big[ra:ra + nb, 0:nb].copy()
        Bm = big[rb:rb + nb, 0:nb].copy()
        C0 = big[rc:rc + nb, nb:2 * nb].copy()
        d = dev(K, big)
        if op == "gemm":
            K.dgemm_nt(nb, nb, nb, -1.0, d.at(ra, 0), ld, d.at(rb, 0), ld, 1.0, d.at(rc, nb), ld)
            exp = oracle.dgemm_nt(-1.0, A, Bm, 1.0, C0)
        elif op == "syrk":
            K.dsyrk_ln(nb, nb, -1.0, d.at(ra, 0), ld, 1.0, d.at(rc, nb), ld)
            exp = oracle.dsyrk_ln(-1.0, A, 1.0, C0)
        else:
            # the codelet's call: the product over the nonzero k-range of the
            # triangular W, which must give the full product's bits
            T = dev(K, C0)
            K.dgemm_nt(nb, nb, nb, 1.0, T.ptr, nb, d.at(rb, 0), ld, 0.0, d.at(rc, nb), ld)
            full = d.download()[rc:rc + nb, nb:2 * nb].copy()
            K.check(K.lib().sa_dgemm_nt_ltri(None, nb, nb, 1.0, T.ptr, nb, d.at(rb, 0), ld, 0.0, d.at(rc, nb), ld))
            exp = oracle.dgemm_nt(1.0, C0, Bm, 0.0, np.zeros_like(C0))
        got = d.download()
        if op == "trsm_gemm":
            assert np.array_equal(got[rc:rc + nb, nb:2 * nb], full)
        tile = got[rc:rc + nb, nb:2 * nb]
        if op == "syrk":
            assert rel(np.tril(tile), np.tril(exp)) < 1e-13
            iu = np.triu_indices(nb, 1)
            assert np.array_equal(tile[iu], C0[iu])
        else:
            assert rel(tile, exp) < 1e-13
        # nothing outside the output tile moves
        rest = got.copy()
        rest[rc:rc + nb, nb:2 * nb] = big[rc:rc + nb, nb:2 * nb]
        assert np.array_equal(rest, big)
    finally:
        K.set_gemm_shape(prev_s)
        K.set_launch_concurrency(prev_c)


@pytest.mark.parametrize("concurrency", [1, 12])
def test_headline_shape_integer_inputs_bit_exact(K, concurrency):
    # small integers: every partial sum is exact, so any dropped or
    # misplaced k-step of the pipelined kernels changes bits
    nb = 1024
    prev = K.set_launch_concurrency(concurrency)
    try:
        rng = np.random.default_rng(40 + concurrency)
        A = rng.integers(-8, 9, (nb, nb)).astype(float)
        Bm = rng.integers(-8, 9, (nb, nb)).astype(float)
        C0 = rng.integers(-64, 65, (nb, nb)).astype(float)
        dA, dB, dC, dS = dev(K, A), dev(K, Bm), dev(K, C0), dev(K, C0)
        K.dgemm_nt(nb, nb, nb, -1.0, dA.ptr, nb, dB.ptr, nb, 1.0, dC.ptr, nb)
        K.dsyrk_ln(nb, nb, -1.0, dA.ptr, nb, 1.0, dS.ptr, nb)
        assert np.array_equal(dC.download(), C0 - A @ Bm.T)
        S = dS.download()
        assert np.array_equal(np.tril(S), np.tril(C0 - A @ A.T))
        assert np.array_equal(np.triu(S, 1), np.triu(C0, 1))
    finally:
        K.set_launch_concurrency(prev)


def spd(rng, n):
    B = rng.uniform(-1, 1, (n, n))
    return B @ B.T + n * np.eye(n)


@pytest.mark.parametrize("m,n", [(1024, 1024), (256, 256), (4, 3), (65, 64), (100, 130), (300, 65), (64, 1)])
@pytest.mark.parametrize("alpha", [1.0, -0.5])
def test_dtrsm_rltn_matches_oracle(K, m, n, alpha):
    print(f"dtrsm m={m} n={n} alpha={alpha}", flush=True)
    rng = np.random.default_rng(m * 11 + n)
    L, _ = oracle.dpotrf_l(spd(rng, n))
    L = np.tril(L)
    Bm = rng.uniform(-1, 1, (m, n))
    dL, dB = dev(K, L), dev(K, Bm)
    K.dtrsm_rltn(m, n, alpha, dL.ptr, dL.ld, dB.ptr, dB.ld)
    got = dB.download()
    exp = oracle.dtrsm_rltn(alpha, L, Bm)
    assert rel(got, exp) < 1e-13


@pytest.mark.parametrize("n", [1, 2, 17, 64, 65, 100])
def test_dpotrf_l_golden(K, n):
    A = GOLD[f"potrf_{n}_A"]
    dA = dev(K, A)
    info = K.DeviceInt(-7)
    K.dpotrf_l(n, dA.ptr, dA.ld, info.ptr)
    got = dA.download()
    assert info.get() == 0
    assert rel(np.tril(got), GOLD[f"potrf_{n}_L"]) < 1e-13
    iu = np.triu_indices(n, 1)
    assert np.array_equal(got[iu], A[iu])   # upper triangle not referenced


@pytest.mark.parametrize("n", [256, 1000, 1024])
def test_dpotrf_l_matches_oracle(K, n):
    print(f"dpotrf n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = spd(rng, n)
    dA = dev(K, A)
    info = K.DeviceInt(-7)
    K.dpotrf_l(n, dA.ptr, dA.ld, info.ptr)
    L = np.tril(dA.download())
    ref, rinfo = oracle.dpotrf_l(A)
    assert info.get() == 0 == rinfo
    assert oracle.rel_diff_lower(L, np.tril(ref)) < 1e-13
    assert oracle.backward_error_llt(A, L) < 1e-14


@pytest.mark.parametrize("bad", [3, 70, 200])
def test_dpotrf_l_not_spd_reports_column(K, bad):
    rng = np.random.default_rng(bad)
    A = spd(rng, 256)
    A[bad, bad] = -1e6
    dA = dev(K, A)
    info = K.DeviceInt(-7)
    K.dpotrf_l(256, dA.ptr, dA.ld, info.ptr)
    _, rinfo = oracle.dpotrf_l(A)
    assert info.get() == rinfo == bad + 1


@pytest.mark.parametrize("n", [1, 17, 64, 65, 100, 300, 1000, 1024, 2048])
def test_dpotrf_tile_factor_and_inverse_match_oracle(K, n):
    # the single-launch tile factorisation with the inverse: L against the
    # oracle's dpotrf, W = L^{-T} against a triangular solve of the identity, Linv = W^T exactly, the zero triangles written, the strict
    # upper triangle of A untouched, and the same bits on runs with other
    # numbers of bulk workgroups (the default, 63, 9).
    print(f"dpotrf_tile n={n}", flush=True)
    rng = np.random.default_rng(n + 99)
    A = spd(rng, n)
    ref, rinfo = oracle.dpotrf_l(A)
    ref = np.tril(ref)
    Wref = solve_triangular(ref, np.eye(n), lower=True).T      # L^{-T} by LAPACK's substitution
    runs = []
    for wg in (0, 63, 9):
        dA = dev(K, A)
        dX = dev(K, np.full((n, n), np.nan))
        dW = dev(K, np.full((n, n), np.nan))
        info = K.DeviceInt(-7)
        K.dpotrf_tile(n, dA.ptr, dA.ld, dX.ptr, dX.ld, dW.ptr, dW.ld, info.ptr, workgroups=wg)
        runs.append((dA.download(), dX.download(), dW.download(), info.get()))
    got, X, W, inf = runs[0]
    assert inf == 0 == rinfo
    assert oracle.rel_diff_lower(np.tril(got), ref) < 1e-13
    iu = np.triu_indices(n, 1)
    assert np.array_equal(got[iu], A[iu])
    assert np.array_equal(X, W.T)
    assert not np.any(np.triu(X, 1)) and not np.any(np.tril(W, -1))
    assert rel(W, Wref) < 1e-13
    for other in runs[1:]:
        assert other[3] == 0
        for a, b in zip(runs[0][:3], other[:3]):
            assert np.array_equal(a, b)
    # the inverse of the resident factor alone: the same bits
    dL = dev(K, got)
    dX2, dW2 = dev(K, np.full((n, n), np.nan)), dev(K, np.full((n, n), np.nan))
    K.dtrtri_tile(n, dL.ptr, dL.ld, dX2.ptr, dX2.ld, dW2.ptr, dW2.ld)
    assert np.array_equal(dX2.download(), X) and np.array_equal(dW2.download(), W)
    assert np.array_equal(dL.download(), got)
    # W alone (no Linv), as the mixed-precision POTRF asks for it: same bits
    dA3, dW3, info3 = dev(K, A), dev(K, np.full((n, n), np.nan)), K.DeviceInt(-7)
    K.dpotrf_tile(n, dA3.ptr, dA3.ld, None, 0, dW3.ptr, dW3.ld, info3.ptr)
    assert info3.get() == 0
    assert np.array_equal(dW3.download(), W) and np.array_equal(dA3.download(), got)


def test_tile_limit_and_large_tile_inverse(K):
    # tiles past the single-launch limit: sa_dpotrf_tile refuses them, and
    # sa_dtrtri_tile inverts them by substitution
    assert K.lib().sa_dpotrf_tile_maxn() == 2048
    n = 2100
    rng = np.random.default_rng(2100)
    A = spd(rng, n)
    L, _ = oracle.dpotrf_l(A)
    L = np.tril(L)
    dL = dev(K, L)
    dX, dW = dev(K, np.full((n, n), np.nan)), dev(K, np.full((n, n), np.nan))
    K.dtrtri_tile(n, dL.ptr, dL.ld, dX.ptr, dX.ld, dW.ptr, dW.ld)
    X, W = dX.download(), dW.download()
    assert np.array_equal(X, W.T)
    assert rel(W, solve_triangular(L, np.eye(n), lower=True).T) < 1e-13
    info = K.DeviceInt(-7)
    ws = K.DeviceArray(K.lib().sa_dpotrf_tile_worksize(n), 1, dtype=np.uint8)
    rc = K.lib().sa_dpotrf_tile(None, n, dL.ptr, n, None, 0, None, 0, info.ptr, ws.ptr,
                                K.lib().sa_dpotrf_tile_worksize(n))
    assert rc == -22


@pytest.mark.parametrize("bad", [5, 130, 1000])
def test_dpotrf_tile_not_spd_reports_column(K, bad):
    n = 1024
    rng = np.random.default_rng(bad)
    A = spd(rng, n)
    A[bad, bad] = -1e6
    dA, dX, dW = dev(K, A), dev(K, np.zeros((n, n))), dev(K, np.zeros((n, n)))
    info = K.DeviceInt(-7)
    K.dpotrf_tile(n, dA.ptr, dA.ld, dX.ptr, dX.ld, dW.ptr, dW.ld, info.ptr)
    _, rinfo = oracle.dpotrf_l(A)
    assert info.get() == rinfo == bad + 1


def test_dpotrf_l_above_tile_limit_matches_oracle(K):
    # past 2048 the factorisation runs one launch per 64-column step
    n = 2100
    rng = np.random.default_rng(n)
    A = spd(rng, n)
    dA = dev(K, A)
    info = K.DeviceInt(-7)
    K.dpotrf_l(n, dA.ptr, dA.ld, info.ptr)
    ref, rinfo = oracle.dpotrf_l(A)
    assert info.get() == 0 == rinfo
    assert oracle.rel_diff_lower(np.tril(dA.download()), np.tril(ref)) < 1e-13


def test_invalid_arguments_are_rejected(K):
    with pytest.raises(RuntimeError, match="-22"):
        K.dgemm_nt(10, 10, 10, 1.0, 0, 5, 0, 10, 0.0, 0, 10)   # lda < m
    with pytest.raises(RuntimeError, match="-22"):
        K.dpotrf_l(10, 0, 10, 0)                              # no info pointer
    with pytest.raises(RuntimeError, match="-22"):
        K.dpotrf_tile(4096, 0, 4096, 0, 4096, 0, 4096, 0)    # past the tile limit


@pytest.mark.parametrize("n", [64, 128, 200, 1024, 30])
def test_dtrtri_diag64_inverts_diagonal_blocks(K, n):
    rng = np.random.default_rng(n + 5)
    L, _ = oracle.dpotrf_l(spd(rng, n))
    L = np.tril(L)
    dL = dev(K, L)
    nblk = (n + 63) // 64
    dW = K.DeviceArray(64, 64 * nblk)
    K.check(K.lib().sa_dtrtri_diag64(None, n, dL.ptr, dL.ld, dW.ptr), "trtri")
    W = dW.download()
    for b in range(nblk):
        j0, w = 64 * b, min(64, n - 64 * b)
        blk = W[:, 64 * b:64 * b + 64]
        Ljj = L[j0:j0 + w, j0:j0 + w]
        # inv(L_jj) * L_jj = I (a triangular inverse, so the check is exact up to rounding)
        assert np.max(np.abs(blk[:w, :w] @ Ljj - np.eye(w))) < 1e-13
        assert np.array_equal(np.triu(blk[:w, :w], 1), np.zeros((w, w)))
        if w < 64:   # identity padding past n
            assert np.array_equal(blk[w:, w:], np.eye(64 - w)) and not blk[w:, :w].any()


def test_workspace_is_required(K):
    with pytest.raises(RuntimeError, match="-22"):
        K.check(K.lib().sa_dtrsm_rltn(None, 10, 10, 1.0, 0, 10, 0, 10, None, 0), "trsm")
    with pytest.raises(RuntimeError, match="-22"):
        K.check(K.lib().sa_dpotrf_l(None, 10, 0, 10, 0, None, 0), "potrf")


# ---- solve-phase kernels ----------------------------------------------------
SOLVE = np.load(os.path.join(os.path.dirname(__file__), "golden", "solve_vectors.npz"))


@pytest.mark.parametrize("key", ["5_3", "64_16", "100_7", "130_9"])
def test_dtrsm_left_golden(K, key):
    L, B = SOLVE[f"trsm_llnn_{key}_L"], SOLVE[f"trsm_llnn_{key}_B"]
    alpha = float(SOLVE[f"trsm_llnn_{key}_alpha"])
    m, n = B.shape
    dL, dB = dev(K, L), dev(K, B)
    K.dtrsm_llnn(m, n, alpha, dL.ptr, dL.ld, dB.ptr, dB.ld)
    assert rel(dB.download(), SOLVE[f"trsm_llnn_{key}_R"]) < 1e-13
    dB = dev(K, B)
    K.dtrsm_lltn(m, n, alpha, dL.ptr, dL.ld, dB.ptr, dB.ld)
    assert rel(dB.download(), SOLVE[f"trsm_lltn_{key}_R"]) < 1e-13


@pytest.mark.parametrize("m,n", [(1024, 16), (1024, 1), (1000, 13), (64, 8), (65, 9), (1, 1),
                                 (2500, 16), (2048, 33), (3, 40)])
@pytest.mark.parametrize("alpha", [1.0, -2.0])
def test_dtrsm_left_matches_oracle(K, m, n, alpha):
    # m > 1024 takes the chunked path (solve 1024 rows, thin GEMM update)
    print(f"dtrsm_left m={m} n={n} alpha={alpha}", flush=True)
    rng = np.random.default_rng(m * 3 + n)
    L = np.tril(oracle.dpotrf_l(spd(rng, m))[0])
    Bm = rng.uniform(-1, 1, (m, n))
    dL = dev(K, L, ld=m + 3)
    dB = dev(K, Bm, ld=m + 1)
    K.dtrsm_llnn(m, n, alpha, dL.ptr, dL.ld, dB.ptr, dB.ld)
    assert rel(dB.download(), oracle.dtrsm_llnn(alpha, L, Bm)) < 1e-13
    dB = dev(K, Bm, ld=m + 1)
    K.dtrsm_lltn(m, n, alpha, dL.ptr, dL.ld, dB.ptr, dB.ld)
    assert rel(dB.download(), oracle.dtrsm_lltn(alpha, L, Bm)) < 1e-13


@pytest.mark.parametrize("key", ["7_3_5", "64_16_64", "70_9_33"])
def test_dgemm_nn_tn_golden(K, key):
    B, C0 = SOLVE[f"gemm_{key}_B"], SOLVE[f"gemm_{key}_C"]
    A, At = SOLVE[f"gemm_nn_{key}_A"], SOLVE[f"gemm_tn_{key}_A"]
    m, n = C0.shape
    k = B.shape[0]
    dA, dAt, dB, dC = dev(K, A), dev(K, At), dev(K, B), dev(K, C0)
    K.dgemm_nn(m, n, k, -1.0, dA.ptr, dA.ld, dB.ptr, dB.ld, 1.0, dC.ptr, dC.ld)
    assert rel(dC.download(), SOLVE[f"gemm_nn_{key}_R"]) < 1e-14
    dC = dev(K, C0)
    K.dgemm_tn(m, n, k, -1.0, dAt.ptr, dAt.ld, dB.ptr, dB.ld, 1.0, dC.ptr, dC.ld)
    assert rel(dC.download(), SOLVE[f"gemm_tn_{key}_R"]) < 1e-14


@pytest.mark.parametrize("m,n,k", [(1024, 16, 1024), (1000, 7, 999), (63, 33, 130), (1, 1, 1), (5, 9, 0)])
@pytest.mark.parametrize("beta", [1.0, 0.0])
def test_dgemm_thin_integer_inputs_bit_exact(K, m, n, k, beta):
    rng = np.random.default_rng(m + n + k)
    A = rng.integers(-8, 9, (m, k)).astype(float)
    At = rng.integers(-8, 9, (k, m)).astype(float)
    B = rng.integers(-8, 9, (k, n)).astype(float)
    C0 = rng.integers(-8, 9, (m, n)).astype(float)
    if beta == 0.0:
        C0[0, 0] = np.nan    # beta == 0 must not read C
    dA, dAt, dB = dev(K, A), dev(K, At), dev(K, B)
    exp_base = 0.0 if beta == 0.0 else C0
    dC = dev(K, C0)
    K.dgemm_nn(m, n, k, -1.0, dA.ptr, dA.ld, dB.ptr, dB.ld, beta, dC.ptr, dC.ld)
    assert np.array_equal(dC.download(), exp_base - A @ B)
    dC = dev(K, C0)
    K.dgemm_tn(m, n, k, -1.0, dAt.ptr, dAt.ld, dB.ptr, dB.ld, beta, dC.ptr, dC.ld)
    assert np.array_equal(dC.download(), exp_base - At.T @ B)


@pytest.mark.parametrize("m,n,k", [(4096, 16, 256), (3000, 5, 96), (2052, 20, 1000), (2050, 3, 64),
                                   (2048, 17, 36)])
@pytest.mark.parametrize("beta", [1.0, 0.0])
def test_dgemm_thin_tall_operands_integer_exact(K, m, n, k, beta):
    # m >= 2048 takes the MFMA form when k % 8 == 0 (and m % 4 == 0 for NN):
    # (2050, 3, 64) runs NN on the VALU form and TN on the MFMA form,
    # (2048, 17, 36) both on the VALU form.  Exact on small integers, with
    # the split-k workspace and without it (one k slice).
    print(f"dgemm thin tall m={m} n={n} k={k} beta={beta}", flush=True)
    rng = np.random.default_rng(m * 5 + n + k)
    A = rng.integers(-8, 9, (m, k)).astype(float)
    At = rng.integers(-8, 9, (k, m)).astype(float)
    B = rng.integers(-8, 9, (k, n)).astype(float)
    C0 = rng.integers(-8, 9, (m, n)).astype(float)
    if beta == 0.0:
        C0[0, 0] = np.nan    # beta == 0 must not read C
    exp_base = 0.0 if beta == 0.0 else C0
    dA, dAt, dB = dev(K, A), dev(K, At), dev(K, B)
    for work in (-1, None):
        dC = dev(K, C0)
        K.dgemm_nn(m, n, k, -1.0, dA.ptr, dA.ld, dB.ptr, dB.ld, beta, dC.ptr, dC.ld, work=work)
        assert np.array_equal(dC.download(), exp_base - A @ B)
        dC = dev(K, C0)
        K.dgemm_tn(m, n, k, -1.0, dAt.ptr, dAt.ld, dB.ptr, dB.ld, beta, dC.ptr, dC.ld, work=work)
        assert np.array_equal(dC.download(), exp_base - At.T @ B)


@pytest.mark.parametrize("n", [1024, 1000, 64, 65, 1, 300])
def test_dtrtri_lt_inverts_transpose(K, n):
    rng = np.random.default_rng(n + 77)
    L = np.tril(oracle.dpotrf_l(spd(rng, n))[0])
    dL = dev(K, L)
    dW = dev(K, np.full((n, n), np.nan), ld=n + 2)     # every entry is written
    K.dtrtri_lt(n, dL.ptr, dL.ld, dW.ptr, dW.ld)
    W = dW.download()
    # W = L^{-T}: W^T L = I, upper triangular with zeros stored below
    assert np.max(np.abs(W.T @ L - np.eye(n))) < 1e-12
    assert np.array_equal(np.tril(W, -1), np.zeros((n, n)))


@pytest.mark.parametrize("m,n", [(1024, 16), (1000, 3), (64, 1), (130, 17)])
@pytest.mark.parametrize("trans", [0, 1])
def test_dtrmm_lu_integer_exact(K, m, n, trans):
    rng = np.random.default_rng(m + n + trans)
    U = np.triu(rng.integers(-4, 5, (m, m)).astype(float))
    Bm = rng.integers(-4, 5, (m, n)).astype(float)
    dU, dB = dev(K, U), dev(K, Bm, ld=m + 5)
    K.dtrmm_lu(trans, m, n, 2.0, dU.ptr, dU.ld, dB.ptr, dB.ld)
    exp = 2.0 * ((U.T if trans else U) @ Bm)
    assert np.array_equal(dB.download(), exp)


@pytest.mark.parametrize("m,n,k", [(1024, 16, 1024), (200, 5, 3000), (32768, 2, 64)])
def test_dgemm_thin_split_and_single_slice_agree(K, m, n, k):
    # with and without the split-k workspace: exact on integers, and both
    # equal to the oracle on random inputs
    rng = np.random.default_rng(m ^ k)
    A = rng.integers(-8, 9, (m, k)).astype(float)
    B = rng.integers(-8, 9, (k, n)).astype(float)
    dA, dB = dev(K, A), dev(K, B)
    for work in (-1, None):
        dC = dev(K, np.zeros((m, n)))
        K.dgemm_nn(m, n, k, 1.0, dA.ptr, dA.ld, dB.ptr, dB.ld, 0.0, dC.ptr, dC.ld, work=work)
        assert np.array_equal(dC.download(), A @ B)
    At = rng.integers(-8, 9, (k, m)).astype(float)
    dAt = dev(K, At)
    for work in (-1, None):
        dC = dev(K, np.zeros((m, n)))
        K.dgemm_tn(m, n, k, 1.0, dAt.ptr, dAt.ld, dB.ptr, dB.ld, 0.0, dC.ptr, dC.ld, work=work)
        assert np.array_equal(dC.download(), At.T @ B)
