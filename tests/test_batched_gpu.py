"""The batched persistent kernels on the GPU: many independent SPD systems
of uneven sizes factored and solved at once, per system against the CPU
oracle.  FP64 (sa_dpotrs_batched): factor ||L - L_ref|| / ||L_ref|| <= 1e-12,
solution within 1e-12 relative, backward error at the oracle's level.  FP32
factor + FP64 refinement (sa_dposv_batched_ir): backward error <= 1e-14 as
reported and as recomputed on the host, solution within the forward error
the residuals allow ((||r|| + ||r_ref|| + residual rounding) / lambda_min)
of the FP64 oracle solve, step count within one of the oracle's
restatement, A and b left intact, an FP32 breakdown sent to the FP64
kernel.  Non-SPD systems
reported per system, hipGraph replays identical to the eager launch, both
panel-update tiles of the FP64 kernel, and the autotuner's choice and
cache."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    from starpu_amd import batched, kernels
    if kernels.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return batched


def check_system(batch, i):
    A, b, x, L = batch.system(i)
    Lref, info = oracle.dpotrf_l(A)
    assert info == 0
    assert oracle.rel_diff_lower(L, np.tril(Lref)) <= 1e-12, i
    xref = oracle.dpotrs_l(np.tril(Lref), b)
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) <= 1e-12, i
    r = b - A @ x
    assert np.linalg.norm(r) / (np.linalg.norm(A) * np.linalg.norm(x) + np.linalg.norm(b)) < 1e-15, i


@pytest.mark.parametrize("grid", [1, 7, 64])
def test_uneven_batch_matches_oracle(B, grid):
    rng = np.random.default_rng(grid)
    sizes = [1, 7, 32, 63, 64, 65, 130, 256, 333, 512] + list(rng.integers(1, 20, 14) * 32)
    batch = B.Batch(sizes, nrhs=1, shift=0.1, seed=grid)
    batch.launch(grid)
    assert np.array_equal(batch.infos(), np.zeros(len(sizes), dtype=np.int32))
    for i in range(len(sizes)):
        check_system(batch, i)
    batch.free()


def test_several_right_hand_sides_and_largest_n(B):
    batch = B.Batch([1024, 96, 1000], nrhs=3, shift=0.1, seed=4)
    batch.launch(32)
    for i in range(3):
        check_system(batch, i)
    batch.free()


def test_non_spd_system_reports_its_pivot(B):
    from starpu_amd import kernels as K
    batch = B.Batch([128, 200, 64], nrhs=1, shift=0.1, seed=5)
    # make A_1(150, 150) very negative in the pristine copy
    v = -1e6
    addr = batch.A0.ptr + (int(batch.offA[1]) + 150 + 150 * 200) * 8
    K.check(K.lib().sa_memcpy_h2d(addr, np.array([v]).ctypes.data, 8))
    batch.restore()
    batch.launch(4)
    info = batch.infos()
    A1 = batch.system(1)[0]
    _, rinfo = oracle.dpotrf_l(A1)
    assert info[1] == rinfo == 151
    assert info[0] == 0 and info[2] == 0
    check_system(batch, 0)
    check_system(batch, 2)
    batch.free()


def test_graph_replays_match_eager(B):
    import ctypes as C
    from starpu_amd import kernels as K
    sizes = [int(s) for s in np.random.default_rng(3).integers(1, 17, 30) * 32]
    batch = B.Batch(sizes, nrhs=1, shift=0.1, seed=6)
    batch.launch(64)
    K.sync()
    eager_x = batch.B.download().copy()
    eager_l = batch.A.download().copy()
    s = C.c_void_p()
    K.check(K.lib().sa_stream_create(C.byref(s)))
    exe = batch.capture(64, "largest_first", s)
    L = B._lib()
    for rep in range(3):
        batch.restore()
        K.sync()
        K.check(L.sa_graph_launch(exe, s))
        K.check(K.lib().sa_stream_sync(s))
        assert np.array_equal(batch.B.download(), eager_x), rep
        assert np.array_equal(batch.A.download(), eager_l), rep
    K.check(L.sa_graph_destroy(exe))
    K.check(K.lib().sa_stream_destroy(s))
    batch.free()


def test_autotuner_picks_fastest_and_caches(B, tmp_path):
    from starpu_amd import autotune
    batch = B.Batch([64, 128, 256, 512, 96], nrhs=1, shift=0.1, seed=7)
    cands = [(16, "largest_first"), (64, "natural")]
    path = str(tmp_path / "tune.json")
    r = autotune.tune(batch, candidates=cands, reps=2, cache_path=path)
    # each launch shape with both panel tiles
    assert not r["cached"] and len(r["table"]) == 4
    assert sorted((t["grid"], t["rows"]) for t in r["table"]) == [(16, 2), (16, 4), (64, 2), (64, 4)]
    assert r["ms"] == min(t["ms"] for t in r["table"])
    r2 = autotune.tune(batch, candidates=cands, reps=2, cache_path=path)
    assert r2["cached"] and (r2["grid"], r2["order"], r2["rows"]) == (r["grid"], r["order"], r["rows"])
    batch.restore()             # tuning left the matrices factored in place
    batch.launch(r["grid"], r["order"], rows=r["rows"])
    for i in range(5):
        check_system(batch, i)
    B.set_rows(0)
    batch.free()


def test_four_row_panel_tile_matches_two_row_tile(B):
    """The 64-row-per-wave panel update gives the factors and solutions of
    the 32-row one bit for bit (each element's products are accumulated in
    the same order) and so the oracle's."""
    rng = np.random.default_rng(11)
    sizes = [1, 33, 64, 65, 129, 250, 320, 700, 1024] + list(rng.integers(1, 33, 12) * 32)
    batch = B.Batch(sizes, nrhs=2, shift=0.1, seed=11)
    try:
        batch.launch(48, rows=2)
        x2, l2 = batch.B.download().copy(), batch.A.download().copy()
        batch.restore()
        batch.launch(48, rows=4)
        assert np.array_equal(batch.infos(), np.zeros(len(sizes), dtype=np.int32))
        assert np.array_equal(batch.A.download(), l2)
        assert np.array_equal(batch.B.download(), x2)
        for i in range(len(sizes)):
            check_system(batch, i)
        with pytest.raises(ValueError):
            B.set_rows(3)
    finally:
        B.set_rows(0)
        batch.free()


def check_ir_system(batch, i, tol=1e-14):
    A, b, _, _ = batch.system(i)
    x, it, be = batch.ir_result(i)
    xo, info, ito, beo = oracle.sposv_ir(A, b, maxiter=10, tol=tol)
    assert info == 0, i
    assert (be <= tol).all(), (i, be)
    assert (np.abs(it - ito) <= 1).all(), (i, it, ito)
    Lref, _ = oracle.dpotrf_l(A)
    xref = oracle.dpotrs_l(np.tril(Lref), b)
    r = oracle.residual_sym_l(A, x, b)
    rref = oracle.residual_sym_l(A, xref, b)
    for c in range(b.shape[1]):
        host = np.linalg.norm(r[:, c]) / (np.linalg.norm(A) * np.linalg.norm(x[:, c]) + np.linalg.norm(b[:, c]))
        assert host <= 2 * tol, (i, c, host)
        # forward error allowed by the residuals: x - xref = A^{-1} (r - rref),
        # so ||x - xref|| <= (||r|| + ||rref|| + rounding of the two FP64
        # residuals) / lambda_min(A); refinement stops at berr <= tol
        lam_min = np.linalg.eigvalsh(A)[0]
        rnd = 2 * A.shape[0] * np.finfo(float).eps * np.linalg.norm(A) * np.linalg.norm(x[:, c])
        bound = (np.linalg.norm(r[:, c]) + np.linalg.norm(rref[:, c]) + rnd) / lam_min
        fwd = np.linalg.norm(x[:, c] - xref[:, c])
        assert fwd <= bound, (i, c, fwd, bound)
        assert fwd / np.linalg.norm(xref[:, c]) <= 1e-11, (i, c)


@pytest.mark.parametrize("grid", [1, 7, 64])
def test_ir_batch_matches_oracle(B, grid):
    rng = np.random.default_rng(100 + grid)
    sizes = [1, 7, 32, 63, 64, 65, 130, 256, 333, 512] + list(rng.integers(1, 20, 14) * 32)
    batch = B.Batch(sizes, nrhs=1, shift=0.1, seed=grid, ir=True)
    batch.launch_ir(grid)
    assert np.array_equal(batch.infos(), np.zeros(len(sizes), dtype=np.int32))
    for i in range(len(sizes)):
        check_ir_system(batch, i)
    # the refinement kernel only reads A and b
    assert np.array_equal(batch.A.download(), batch.A0.download())
    assert np.array_equal(batch.B.download(), batch.B0.download())
    batch.free()


def test_ir_several_right_hand_sides_and_chunked_residual(B):
    # n > 1024 takes the residual through two row chunks
    batch = B.Batch([1100, 1024, 96], nrhs=3, shift=0.1, seed=8, ir=True)
    batch.launch_ir(32)
    assert (batch.infos() == 0).all()
    for i in range(3):
        check_ir_system(batch, i)
    batch.free()


def test_ir_fp32_breakdown_falls_back_to_fp64(B):
    from starpu_amd import kernels as K
    batch = B.Batch([128, 64, 200], nrhs=1, shift=0.1, seed=9, ir=True)
    # system 1: identity with an SPD 2x2 block that is singular in FP32
    A1 = np.eye(64, order="F")
    A1[10:12, 10:12] = [[1.0, 1.0], [1.0, 1.0 + 1e-10]]
    K.check(K.lib().sa_memcpy_h2d(batch.A0.ptr + int(batch.offA[1]) * 8, A1.ctypes.data, A1.nbytes))
    batch.restore()
    batch.launch_ir(8)
    K.sync()
    assert list(batch.infos()) == [0, 12, 0]
    _, it, be = batch.ir_result(1)
    assert it[0] == 0 and np.isinf(be[0])
    batch.restore()
    fell_back = batch.solve_ir(8)
    assert list(fell_back) == [1]
    assert (batch.infos() == 0).all()
    A, b, _, _ = batch.system(1)
    x, _, _ = batch.ir_result(1)
    Lref, _ = oracle.dpotrf_l(A)
    xref = oracle.dpotrs_l(np.tril(Lref), b)
    # kappa ~ 4e10: FP64 solutions agree to ~kappa * eps; the solve is backward stable
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) <= 1e-5
    r = oracle.residual_sym_l(A, x, b)
    assert np.linalg.norm(r) / (np.linalg.norm(A) * np.linalg.norm(x) + np.linalg.norm(b)) < 1e-15
    check_ir_system(batch, 0)
    check_ir_system(batch, 2)
    batch.free()


def test_ir_graph_replays_match_eager(B):
    import ctypes as C
    from starpu_amd import kernels as K
    sizes = [int(s) for s in np.random.default_rng(13).integers(1, 17, 30) * 32]
    batch = B.Batch(sizes, nrhs=1, shift=0.1, seed=10, ir=True)
    batch.launch_ir(64)
    K.sync()
    eager_x = batch.X.download().copy()
    eager_it = batch.iters.download().copy()
    s = C.c_void_p()
    K.check(K.lib().sa_stream_create(C.byref(s)))
    exe = batch.capture(64, "largest_first", s, ir=True)
    L = B._lib()
    for rep in range(3):
        K.check(K.lib().sa_memcpy_d2d(batch.X.ptr, batch.B0.ptr, batch.totB * 8))   # scribble over x
        K.sync()
        K.check(L.sa_graph_launch(exe, s))
        K.check(K.lib().sa_stream_sync(s))
        assert np.array_equal(batch.X.download(), eager_x), rep
        assert np.array_equal(batch.iters.download(), eager_it), rep
    K.check(L.sa_graph_destroy(exe))
    K.check(K.lib().sa_stream_destroy(s))
    batch.free()
