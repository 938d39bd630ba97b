"""The batched persistent kernel (sa_dpotrs_batched) on the GPU: many
independent SPD systems of uneven sizes factored and solved at once, per
system against the CPU oracle (factor ||L - L_ref|| / ||L_ref|| <= 1e-12,
solution within 1e-12 relative, backward error at the oracle's level),
non-SPD systems reported per system, hipGraph replays identical to the
eager launch, and the autotuner's choice and cache."""
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
    assert not r["cached"] and len(r["table"]) == 2
    assert r["ms"] == min(t["ms"] for t in r["table"])
    r2 = autotune.tune(batch, candidates=cands, reps=2, cache_path=path)
    assert r2["cached"] and (r2["grid"], r2["order"]) == (r["grid"], r["order"])
    batch.restore()             # tuning left the matrices factored in place
    batch.launch(r["grid"], r["order"])
    for i in range(5):
        check_system(batch, i)
    batch.free()
