"""Pins the CPU oracle (oracle/chol_oracle.c) before anything is checked
against it: LAPACK/BLAS golden vectors, closed forms, integer-exact cases and
the reference's own Hilbert + N*I known-answer test
(examples/cholesky/cholesky_implicit.c:474-484,568-583)."""
import os

import numpy as np
import pytest

import oracle

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "lapack_vectors.npz"))


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def lower(a):
    return np.tril(a)


@pytest.mark.parametrize("shape", [(7, 5, 3), (64, 48, 33), (67, 65, 35)])
def test_gemm_nt_matches_lapack(shape):
    m, n, k = shape
    p = f"gemm_{m}_{n}_{k}_"
    R = oracle.dgemm_nt(-1.0, GOLD[p + "A"], GOLD[p + "B"], 1.0, GOLD[p + "C"])
    assert rel(R, GOLD[p + "R"]) < 1e-14


@pytest.mark.parametrize("shape", [(5, 3), (70, 40), (65, 33)])
def test_syrk_lower_matches_lapack_and_keeps_upper(shape):
    n, k = shape
    p = f"syrk_{n}_{k}_"
    C0 = GOLD[p + "C"]
    R = oracle.dsyrk_ln(-1.0, GOLD[p + "A"], 1.0, C0)
    assert rel(lower(R), lower(GOLD[p + "R"])) < 1e-14
    iu = np.triu_indices(n, 1)
    assert np.array_equal(R[iu], C0[iu])  # strict upper triangle untouched


@pytest.mark.parametrize("shape", [(4, 3), (65, 64), (30, 70)])
def test_trsm_rltn_matches_lapack(shape):
    m, n = shape
    p = f"trsm_{m}_{n}_"
    R = oracle.dtrsm_rltn(1.0, GOLD[p + "L"], GOLD[p + "B"])
    assert rel(R, GOLD[p + "R"]) < 1e-13


@pytest.mark.parametrize("n", [1, 2, 17, 64, 65, 100])
def test_potrf_and_potrs_match_lapack(n):
    L, info = oracle.dpotrf_l(GOLD[f"potrf_{n}_A"])
    assert info == 0
    assert rel(lower(L), GOLD[f"potrf_{n}_L"]) < 1e-13
    x = oracle.dpotrs_l(lower(L), GOLD[f"potrs_{n}_b"])
    assert rel(x, GOLD[f"potrs_{n}_x"]) < 1e-12


def test_potrf_not_spd_info_matches_lapack():
    _, info = oracle.dpotrf_l(GOLD["notspd_A"])
    assert info == int(GOLD["notspd_info"]) and info > 0


def test_potrf_closed_form_2x2():
    a, b, c = 4.0, 2.0, 5.0
    L, info = oracle.dpotrf_l(np.array([[a, b], [b, c]]))
    exp = np.array([[2.0, 0.0], [1.0, 2.0]])  # sqrt(a), b/sqrt(a), sqrt(c - b^2/a)
    assert info == 0 and np.array_equal(lower(L), exp)


def test_potrf_diagonal_and_identity():
    d = np.array([1.0, 4.0, 9.0, 16.0, 0.25])
    L, info = oracle.dpotrf_l(np.diag(d))
    assert info == 0 and np.array_equal(lower(L), np.diag(np.sqrt(d)))
    L, info = oracle.dpotrf_l(np.eye(33))
    assert info == 0 and np.array_equal(lower(L), np.eye(33))


def test_gemm_integer_exact_and_transposed_operand_is_caught():
    rng = np.random.default_rng(1)
    A = rng.integers(-8, 9, (9, 6)).astype(float)
    B = rng.integers(-8, 9, (7, 6)).astype(float)
    C0 = rng.integers(-8, 9, (9, 7)).astype(float)
    R = oracle.dgemm_nt(-1.0, A, B, 1.0, C0)
    assert np.array_equal(R, C0 - A @ B.T)
    # The brute-force triple loop catches an N/N instead of N/T mix-up.
    brute = C0.copy()
    for i in range(9):
        for j in range(7):
            brute[i, j] -= sum(A[i, p] * B[j, p] for p in range(6))
    assert np.array_equal(R, brute)


def test_trsm_is_right_side_transposed():
    # X L^T = B  <=>  X = B L^{-T}; check the defining identity directly.
    rng = np.random.default_rng(2)
    L = np.tril(rng.uniform(0.5, 1.0, (6, 6))) + 3 * np.eye(6)
    B = rng.uniform(-1, 1, (5, 6))
    X = oracle.dtrsm_rltn(1.0, L, B)
    assert rel(X @ L.T, B) < 1e-14
    X2 = oracle.dtrsm_rltn(2.5, L, B)
    assert rel(X2, 2.5 * X) < 1e-14


def hilbert_plus_n(n):
    i = np.arange(n, dtype=float)
    return 1.0 / (1.0 + i[:, None] + i[None, :]) + n * np.eye(n)


@pytest.mark.parametrize("n,nblocks", [(64, 1), (64, 4), (100, 8), (10, 3), (257, 5)])
def test_tiled_cholesky_hilbert_kat(n, nblocks):
    # cholesky_implicit.c:474-484 builds 1/(1+i+j) + N*delta_ij and checks
    # that L L^T reproduces it (float, 1e-4); in FP64 it holds to ~1e-15.
    A = hilbert_plus_n(n)
    L, info = oracle.tiled_cholesky(A, nblocks)
    assert info == 0
    L = lower(L)
    assert np.max(np.abs(L @ L.T - A) / np.abs(A)) < 1e-13
    assert oracle.backward_error_llt(A, L) < 1e-15
    # tiling does not change the factor beyond rounding
    L1, _ = oracle.dpotrf_l(A)
    assert oracle.rel_diff_lower(L, lower(L1)) < 1e-14


def test_rel_diff_lower_against_numpy():
    # a known, non-zero answer: ||tril(E)||_F / ||tril(Lref)||_F, with junk
    # in both strict upper triangles that must not count
    rng = np.random.default_rng(31)
    n = 37
    Lref = np.tril(rng.uniform(-1, 1, (n, n))) + 3 * np.eye(n)
    E = np.tril(rng.uniform(-1e-3, 1e-3, (n, n)))
    L = Lref + E
    junk_l, junk_r = L.copy(), Lref.copy()
    iu = np.triu_indices(n, 1)
    junk_l[iu] = rng.uniform(-1e6, 1e6, len(iu[0]))
    junk_r[iu] = rng.uniform(-1e6, 1e6, len(iu[0]))
    want = np.linalg.norm(E) / np.linalg.norm(Lref)
    got = oracle.rel_diff_lower(junk_l, junk_r)
    assert 1e-5 < want and abs(got - want) <= 1e-13 * want
    # identical lower triangles give exactly zero whatever the upper parts
    assert oracle.rel_diff_lower(junk_l, L) == 0.0
    # a single changed entry: |delta| / ||tril(Lref)||_F, and the operands'
    # roles are not symmetric (the denominator is the reference's)
    L2 = Lref.copy()
    L2[20, 4] += 0.5
    assert abs(oracle.rel_diff_lower(L2, Lref) - 0.5 / np.linalg.norm(Lref)) <= 1e-15
    assert abs(oracle.rel_diff_lower(Lref, L2) - 0.5 / np.linalg.norm(L2)) <= 1e-15
    # leading dimension larger than n: a submatrix view
    big = np.zeros((n + 5, n), order="F")
    big[:n] = L
    assert abs(oracle.rel_diff_lower(big[:n], Lref) - want) <= 1e-13 * want


def test_backward_error_llt_against_numpy():
    # A = L0 L0^T + S with an integer L0 (so L0 L0^T is exact) and a known
    # symmetric S: the error is ||tril(S)||_F / ||tril(A)||_F; the strict upper
    # triangles of A and L hold junk that must not count
    rng = np.random.default_rng(32)
    n = 41
    L0 = np.tril(rng.integers(-4, 5, (n, n))).astype(float) + 6 * np.eye(n)
    S = rng.uniform(-1e-2, 1e-2, (n, n))
    S = (S + S.T) / 2
    A = L0 @ L0.T + S
    want = np.linalg.norm(np.tril(S)) / np.linalg.norm(np.tril(A))
    Aj, Lj = A.copy(), L0.copy()
    iu = np.triu_indices(n, 1)
    Aj[iu] = 1e9
    Lj[iu] = -1e9
    got = oracle.backward_error_llt(Aj, Lj)
    assert 1e-5 < want and abs(got - want) <= 1e-12 * want
    # the exact factor of A = L0 L0^T gives zero
    assert oracle.backward_error_llt(L0 @ L0.T, Lj) == 0.0
    # one wrong factor entry changes column j of L L^T and row j: the error
    # is the norm of that rank-2 change restricted to the lower triangle
    L1 = L0.copy()
    L1[30, 7] += 1.0
    D = L1 @ L1.T - L0 @ L0.T
    want1 = np.linalg.norm(np.tril(D)) / np.linalg.norm(np.tril(L0 @ L0.T))
    assert abs(oracle.backward_error_llt(L0 @ L0.T, L1) - want1) <= 1e-14 * want1


def test_tiled_cholesky_reports_global_pivot_column():
    A = hilbert_plus_n(12)
    A[7, 7] = -1.0
    _, info = oracle.tiled_cholesky(A, 3)   # tiles of 4: column 8 is in tile 1
    _, info1 = oracle.dpotrf_l(A)
    assert info == info1 == 8


def test_residual_and_backward_error_helpers():
    A = hilbert_plus_n(20)
    L, _ = oracle.dpotrf_l(A)
    b = np.linspace(-1, 1, 40).reshape(20, 2, order="F")
    x = oracle.dpotrs_l(lower(L), b)
    r = oracle.residual_sym_l(np.tril(A), x, b)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-14
    assert np.allclose(r, b - A @ x, atol=1e-14)


# ---- solve phase -----------------------------------------------------------
SOLVE = np.load(os.path.join(os.path.dirname(__file__), "golden", "solve_vectors.npz"))


@pytest.mark.parametrize("key", ["5_3", "64_16", "100_7", "130_9"])
def test_trsm_left_lower_both_transposes_match_lapack(key):
    L, B = SOLVE[f"trsm_llnn_{key}_L"], SOLVE[f"trsm_llnn_{key}_B"]
    alpha = float(SOLVE[f"trsm_llnn_{key}_alpha"])
    X = oracle.dtrsm_llnn(alpha, L, B)
    assert rel(X, SOLVE[f"trsm_llnn_{key}_R"]) < 1e-14
    Y = oracle.dtrsm_lltn(alpha, L, B)
    assert rel(Y, SOLVE[f"trsm_lltn_{key}_R"]) < 1e-14
    # the defining equations: L X = alpha B and L^T Y = alpha B; a swapped
    # transpose or a dropped alpha fails them
    assert rel(L @ X, alpha * B) < 1e-13
    assert rel(L.T @ Y, alpha * B) < 1e-13


def test_trsm_left_lower_ignores_upper_triangle():
    L = SOLVE["trsm_llnn_64_16_L"].copy(order="F")
    B = SOLVE["trsm_llnn_64_16_B"]
    L[np.triu_indices(64, 1)] = np.nan
    assert np.all(np.isfinite(oracle.dtrsm_llnn(1.0, L, B)))
    assert np.all(np.isfinite(oracle.dtrsm_lltn(1.0, L, B)))


@pytest.mark.parametrize("key", ["7_3_5", "64_16_64", "70_9_33"])
def test_gemm_nn_tn_match_lapack(key):
    B, C0 = SOLVE[f"gemm_{key}_B"], SOLVE[f"gemm_{key}_C"]
    assert rel(oracle.dgemm_nn(-1.0, SOLVE[f"gemm_nn_{key}_A"], B, 1.0, C0), SOLVE[f"gemm_nn_{key}_R"]) < 1e-14
    assert rel(oracle.dgemm_tn(-1.0, SOLVE[f"gemm_tn_{key}_A"], B, 1.0, C0), SOLVE[f"gemm_tn_{key}_R"]) < 1e-14


def test_gemm_nn_tn_integer_exact():
    rng = np.random.default_rng(3)
    A = rng.integers(-9, 10, (40, 30)).astype(float)
    B = rng.integers(-9, 10, (30, 6)).astype(float)
    C0 = rng.integers(-9, 10, (40, 6)).astype(float)
    assert np.array_equal(oracle.dgemm_nn(2.0, A, B, -1.0, C0), 2.0 * A @ B - C0)
    assert np.array_equal(oracle.dgemm_tn(2.0, A, C0, 0.0, np.full((30, 6), np.nan)), 2.0 * A.T @ C0)


@pytest.mark.parametrize("n,nrhs", [(96, 16), (130, 5)])
@pytest.mark.parametrize("nblocks", [1, 3, 7])
def test_tiled_potrs_matches_lapack_cho_solve(n, nrhs, nblocks):
    A, b = SOLVE[f"potrs_{n}_{nrhs}_A"], SOLVE[f"potrs_{n}_{nrhs}_b"]
    L, info = oracle.tiled_cholesky(A, nblocks)
    assert info == 0
    X = oracle.tiled_potrs(np.tril(L), b, nblocks)
    assert rel(X, SOLVE[f"potrs_{n}_{nrhs}_x"]) < 1e-13
    # tiling does not change the arithmetic beyond reassociation
    assert rel(X, oracle.dpotrs_l(np.tril(L), b)) < 1e-14
    # normwise backward error of the solve
    r = oracle.residual_sym_l(A, X, b)
    assert np.linalg.norm(r) / (np.linalg.norm(A) * np.linalg.norm(X) + np.linalg.norm(b)) < 1e-15


# ---- FP32 factor + FP64 refinement (the dsposv scheme) ------------------------
def berr_of(A, x, b):
    r = b - A @ x
    return np.linalg.norm(r) / (np.linalg.norm(A) * np.linalg.norm(x) + np.linalg.norm(b))


@pytest.mark.parametrize("n,nrhs", [(96, 16), (130, 5)])
def test_sposv_ir_reaches_the_lapack_solution(n, nrhs):
    A, b = SOLVE[f"potrs_{n}_{nrhs}_A"], SOLVE[f"potrs_{n}_{nrhs}_b"]
    X, info, it, be = oracle.sposv_ir(A, b, maxiter=10, tol=1e-14)
    assert info == 0
    assert rel(X, SOLVE[f"potrs_{n}_{nrhs}_x"]) < 1e-13
    assert (be <= 1e-14).all() and (it >= 2).all() and (it <= 4).all()
    for c in range(nrhs):
        assert berr_of(A, X[:, c], b[:, c]) < 2e-14


def test_sposv_ir_first_step_is_an_fp32_solve_and_refinement_contracts():
    rng = np.random.default_rng(11)
    n = 200
    G = rng.standard_normal((n, n))
    A = G @ G.T / n + 0.1 * np.eye(n)          # kappa ~ 40
    b = rng.standard_normal((n, 1))
    x = np.linalg.solve(A, b)
    errs = []
    for k in (1, 2, 3):
        Xk, info, it, be = oracle.sposv_ir(A, b, maxiter=k, tol=0.0)
        assert info == 0 and it[0] == k
        assert abs(be[0] - berr_of(A, Xk, b)) <= 1e-3 * be[0] + 1e-17
        errs.append(rel(Xk, x))
    # one FP32 solve: error ~ kappa * 2^-24, far from FP64 and far from wrong
    assert 1e-9 < errs[0] < 1e-3
    # each refinement step gains at least three digits until FP64 rounding
    assert errs[1] < 1e-3 * errs[0] and errs[2] < max(1e-3 * errs[1], 1e-14)
    X, info, it, be = oracle.sposv_ir(A, b)
    assert rel(X, x) < 1e-13 and be[0] <= 1e-14 and it[0] <= 4


def test_sposv_ir_exact_on_diagonal_systems():
    d = np.array([4.0, 9.0, 16.0, 25.0, 64.0])
    A = np.diag(d)
    b = (d * np.array([1.0, -2.0, 3.0, 0.5, -0.25])).reshape(-1, 1)
    X, info, it, be = oracle.sposv_ir(A, b)
    assert info == 0 and it[0] == 1 and be[0] == 0.0
    assert np.array_equal(X[:, 0], [1.0, -2.0, 3.0, 0.5, -0.25])


def test_sposv_ir_reports_an_fp32_breakdown_of_an_fp64_spd_matrix():
    A = np.eye(6)
    A[1:3, 1:3] = [[1.0, 1.0], [1.0, 1.0 + 1e-10]]     # SPD, but 1 + 1e-10 rounds to 1 in FP32
    assert oracle.dpotrf_l(A)[1] == 0
    X, info, it, be = oracle.sposv_ir(A, np.ones((6, 2)))
    assert info == 3 and (it == 0).all() and np.isinf(be).all() and not X.any()


def test_sposv_ir_stops_at_maxiter():
    A = SOLVE["potrs_96_16_A"]
    X, info, it, be = oracle.sposv_ir(A, SOLVE["potrs_96_16_b"][:, :2], maxiter=3, tol=0.0)
    assert info == 0 and (it == 3).all() and (be > 0).all()


# ---- bfloat16 model of the mixed-precision updates ---------------------------
def test_bf16_round_known_values_and_ties_to_even():
    from oracle.mixed import bf16_round
    ulp = 2.0 ** -7                       # bf16 spacing in [1, 2)
    x = np.array([1.0, 1.0 + ulp, 1.0 + ulp / 2, 1.0 + 3 * ulp / 2, 1.0 + ulp / 2 + 2 ** -20,
                  -2.5, 256.0, 257.0, 0.0], dtype=np.float32)
    got = bf16_round(x)
    exp = np.array([1.0, 1.0 + ulp, 1.0, 1.0 + 2 * ulp, 1.0 + ulp, -2.5, 256.0, 256.0, 0.0], dtype=np.float32)
    assert np.array_equal(got, exp)       # 257 -> 256: tie to the even mantissa


def test_bf16_round_matches_torch_conversion():
    # an independent implementation of the same rounding
    import torch
    from oracle.mixed import bf16_round
    rng = np.random.default_rng(11)
    x = (rng.standard_normal(100000) * np.exp(rng.uniform(-20, 20, 100000))).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(bf16_round(x), ref)


def test_bf16_split_products_error_bounds():
    from oracle.mixed import gemm_tn_model, split_bf16
    rng = np.random.default_rng(12)
    A = rng.uniform(-1, 1, (256, 40)).astype(np.float32)
    B = rng.uniform(-1, 1, (256, 30)).astype(np.float32)
    exact = A.astype(np.float64).T @ B.astype(np.float64)
    scale = np.abs(A).T.astype(np.float64) @ np.abs(B).astype(np.float64)
    e1 = np.max(np.abs(gemm_tn_model(1, A, B) - exact) / scale)
    e3 = np.max(np.abs(gemm_tn_model(2, A, B) - exact) / scale)
    assert 2 ** -14 < e1 < 2 ** -7       # one bf16 product: about 8 bits (sums cancel some)
    assert e3 < 2 ** -15 and e3 < e1 / 100  # the split recovers about 16 bits
    hi, lo = split_bf16(A)
    assert np.max(np.abs(A - hi - lo) / np.maximum(np.abs(A), 1e-30)) < 2 ** -15
