"""End-to-end tiled Cholesky on the GPU through the StarPU-style runtime:
the reference's POTRF/TRSM/SYRK/GEMM DAG (examples/cholesky/cholesky_implicit.c)
submitted with starpu_task_insert, scheduled by dmda onto HIP workers whose
codelets call the hand-written gfx950 kernels.

Oracle bar (FP64): ||L - L_ref||_F / ||L_ref||_F <= 1e-12 against the CPU
oracle's tiled factorisation of the same matrix, and backward error
||A - L L^T|| / ||A|| no worse than 10x the oracle's."""
import os
import tempfile

import numpy as np
import pytest

import oracle

PERF_MODEL_DIR = os.path.join(tempfile.gettempdir(), "starpu_amd_perfmodels")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from starpu_amd import kernels as K
    from starpu_amd import runtime as R
    if K.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    R.init(STARPU_NHIP=1, STARPU_NCPU=0, STARPU_SCHED="dmda", STARPU_PERF_MODEL_DIR=PERF_MODEL_DIR)
    R.set_cpu_kernels()
    yield R, K
    R.shutdown()


def spd(n, seed, shift=None):
    rng = np.random.default_rng(seed)
    B = rng.uniform(-1, 1, (n, n))
    return np.asfortranarray(B @ B.T + (n if shift is None else shift) * np.eye(n))


def check(A0, L, nblocks):
    ref, info = oracle.tiled_cholesky(A0, nblocks)
    assert info == 0
    rel = oracle.rel_diff_lower(np.tril(L), np.tril(ref))
    be = oracle.backward_error_llt(A0, np.tril(L))
    be_ref = oracle.backward_error_llt(A0, np.tril(ref))
    assert rel <= 1e-12, rel
    assert be <= 10 * max(be_ref, 1e-17), (be, be_ref)
    return rel, be


# uneven tilings: the first n mod nblocks tiles get one more row and column
# (matrix_filters.c:26-103), so the kernels see non-square and unaligned tiles
UNEVEN = [(1000, 3), (2050, 8), (1100, 3), (777, 5), (1025, 2)]


@pytest.mark.parametrize("n,nblocks", [(512, 4), (1024, 8), (2048, 8), (768, 3), (1000, 4), (256, 1),
                                       (1100, 2)] + UNEVEN)
def test_device_resident_factor_matches_oracle(rt, n, nblocks):
    R, K = rt
    print(f"device-resident chol n={n} nblocks={nblocks}", flush=True)
    A0 = spd(n, n + nblocks)
    d = K.DeviceArray.from_host(A0)
    st = R.factor(d.ptr, d.ld, n, nblocks, home_node=R.hip_node(0))
    assert st.info == 0
    L = d.download()
    check(A0, L, nblocks)
    # the strict upper triangle is never written (lower-only semantics)
    iu = np.triu_indices(n, 1)
    assert np.array_equal(L[iu], A0[iu])


@pytest.mark.parametrize("n,nblocks", [(512, 4), (640, 5), (1000, 3), (1025, 2)])
def test_host_resident_factor_moves_tiles_through_msi(rt, n, nblocks):
    # A registered in main memory: tiles go to HBM on demand and the
    # unpartition gathers the factor back (cholesky_implicit.c:423,456).
    R, K = rt
    print(f"host-resident chol n={n} nblocks={nblocks}", flush=True)
    A0 = spd(n, 7 * n)
    A = A0.copy(order="F")
    st = R.factor(A.ctypes.data, n, n, nblocks, home_node=0)
    assert st.info == 0
    check(A0, A, nblocks)


def test_reference_hilbert_kat(rt):
    # cholesky_implicit.c:474-484,568-583: 1/(1+i+j) + N*delta_ij, L L^T == A
    R, K = rt
    n, nb = 1024, 8
    i = np.arange(n, dtype=float)
    A0 = np.asfortranarray(1.0 / (1.0 + i[:, None] + i[None, :]) + n * np.eye(n))
    d = K.DeviceArray.from_host(A0)
    R.factor(d.ptr, d.ld, n, nb, home_node=R.hip_node(0))
    L = np.tril(d.download())
    err = np.max(np.abs(L @ L.T - A0) / np.abs(A0))
    assert err < 1e-13


def test_not_spd_reports_global_column(rt):
    R, K = rt
    n, nb = 512, 4
    A0 = spd(n, 3)
    A0[300, 300] = -1e4
    d = K.DeviceArray.from_host(A0)
    st = R.factor(d.ptr, d.ld, n, nb, home_node=R.hip_node(0))
    _, info = oracle.tiled_cholesky(A0, nb)
    assert st.info == info == 301


def test_repeat_runs_are_identical(rt):
    # same DAG, same tiles: the result does not depend on the schedule
    R, K = rt
    n, nb = 1024, 8
    A0 = spd(n, 11)
    outs = []
    for _ in range(3):
        d = K.DeviceArray.from_host(A0)
        R.factor(d.ptr, d.ld, n, nb, home_node=R.hip_node(0))
        outs.append(d.download())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_driver_thread_count_does_not_change_results(rt):
    # one host driver thread for the 8 streams, then one per stream: the same
    # DAG gives the same bits as with the default four threads
    R, K = rt
    n, nb, nrhs = 2048, 8, 4
    A0 = spd(n, 23)
    b = np.asfortranarray(np.random.default_rng(23).standard_normal((n, nrhs)))

    def run():
        dA, dB = K.DeviceArray.from_host(A0), K.DeviceArray.from_host(b)
        st = R.factor_solve(dA.ptr, dA.ld, n, nb, dB.ptr, dB.ld, nrhs, home_node=R.hip_node(0))
        assert st.info == 0
        return dA.download(), dB.download()

    ref = run()
    try:
        for env in ({"STARPU_HIP_DRIVER_THREADS": 1}, {"STARPU_HIP_THREAD_PER_WORKER": 1}):
            print(f"driver threads {env}", flush=True)
            R.shutdown()
            R.init(**env)
            R.set_cpu_kernels()
            out = run()
            assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1])
            os.environ.pop(next(iter(env)))
    finally:
        os.environ.pop("STARPU_HIP_DRIVER_THREADS", None)
        os.environ.pop("STARPU_HIP_THREAD_PER_WORKER", None)
        R.shutdown()
        R.init()
        R.set_cpu_kernels()


@pytest.mark.parametrize("n,nblocks,nrhs", [(2048, 8, 16), (1000, 4, 3), (2048, 2, 16), (512, 1, 1),
                                            (1000, 3, 16), (2050, 8, 16), (1100, 2, 16), (777, 5, 7)])
def test_device_resident_factor_solve_matches_oracle(rt, n, nblocks, nrhs):
    # factor + solve as one DAG; X against the oracle's tiled solve with the
    # oracle's factor, and the normwise backward error of the solve
    R, K = rt
    print(f"device-resident factor_solve n={n} nblocks={nblocks} nrhs={nrhs}", flush=True)
    A0 = spd(n, n + nrhs)
    rng = np.random.default_rng(nrhs)
    b = np.asfortranarray(rng.standard_normal((n, nrhs)))
    dA = K.DeviceArray.from_host(A0)
    dB = K.DeviceArray.from_host(b)
    st = R.factor_solve(dA.ptr, dA.ld, n, nblocks, dB.ptr, dB.ld, nrhs, home_node=R.hip_node(0))
    assert st.info == 0
    check(A0, dA.download(), nblocks)
    X = dB.download()
    Lref, _ = oracle.tiled_cholesky(A0, nblocks)
    Xref = oracle.tiled_potrs(np.tril(Lref), b, nblocks)
    assert np.linalg.norm(X - Xref) / np.linalg.norm(Xref) < 1e-12
    r = oracle.residual_sym_l(A0, X, b)
    berr = np.linalg.norm(r) / (np.linalg.norm(A0) * np.linalg.norm(X) + np.linalg.norm(b))
    r_ref = oracle.residual_sym_l(A0, Xref, b)
    berr_ref = np.linalg.norm(r_ref) / (np.linalg.norm(A0) * np.linalg.norm(Xref) + np.linalg.norm(b))
    assert berr <= 10 * max(berr_ref, 1e-17), (berr, berr_ref)
    # a repeat solve on the resident factor gives the same X
    dB2 = K.DeviceArray.from_host(b)
    R.solve(dA.ptr, dA.ld, n, nblocks, dB2.ptr, dB2.ld, nrhs, home_node=R.hip_node(0))
    assert np.array_equal(dB2.download(), X)


@pytest.mark.parametrize("n,nblocks,nrhs", [(1000, 3, 16), (1025, 2, 5)])
def test_host_resident_factor_solve_uneven(rt, n, nblocks, nrhs):
    # A and b in main memory, uneven tiles: every transfer is a 2-D copy of
    # a view with the parent's ld; the factor and X come back on unpartition
    R, K = rt
    print(f"host-resident factor_solve n={n} nblocks={nblocks} nrhs={nrhs}", flush=True)
    A0 = spd(n, 3 * n + nrhs)
    b = np.asfortranarray(np.random.default_rng(n).standard_normal((n, nrhs)))
    A, X = A0.copy(order="F"), b.copy(order="F")
    st = R.factor_solve(A.ctypes.data, n, n, nblocks, X.ctypes.data, n, nrhs, home_node=0)
    assert st.info == 0
    check(A0, A, nblocks)
    Lref, _ = oracle.tiled_cholesky(A0, nblocks)
    Xref = oracle.tiled_potrs(np.tril(Lref), b, nblocks)
    assert np.linalg.norm(X - Xref) / np.linalg.norm(Xref) < 1e-12


def spectrum_spd(n, cond, seed):
    """Q diag(d) Q^T with d log-spaced from 1 down to 1/cond."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Q * np.logspace(0, -np.log10(cond), n)) @ Q.T
    return np.asfortranarray((A + A.T) / 2)


@pytest.mark.parametrize("cond", [1e8, 1e12])
def test_ill_conditioned_factor_backward_error(rt, cond):
    # The TRSM tasks multiply by the inverse of the diagonal tile instead of
    # substituting; on SPD matrices with a prescribed spectrum the backward
    # error must stay within 10x the oracle's substitution-based factor.
    R, K = rt
    n, nblocks = 2048, 8
    print(f"ill-conditioned chol cond={cond:.0e}", flush=True)
    A0 = spectrum_spd(n, cond, 5)
    d = K.DeviceArray.from_host(A0)
    st = R.factor(d.ptr, d.ld, n, nblocks, home_node=R.hip_node(0))
    assert st.info == 0
    L = np.tril(d.download())
    ref, info = oracle.tiled_cholesky(A0, nblocks)
    assert info == 0
    be = oracle.backward_error_llt(A0, L)
    be_ref = oracle.backward_error_llt(A0, np.tril(ref))
    print(f"  backward error {be:.2e} oracle {be_ref:.2e}", flush=True)
    assert be <= 10 * be_ref, (be, be_ref)


@pytest.fixture
def rt12(rt):
    """The runtime with twelve HIP streams (workers) per GPU, the benchmark's
    setting: 1024-wide tiles then run on the 128x128 GEMM kernels."""
    R, K = rt
    R.shutdown()
    os.environ["STARPU_NWORKER_PER_HIP"] = "12"
    try:
        R.init()
        R.set_cpu_kernels()
        yield R, K
    finally:
        os.environ.pop("STARPU_NWORKER_PER_HIP", None)
        R.shutdown()
        R.init()
        R.set_cpu_kernels()


@pytest.mark.parametrize("n,nblocks,nrhs", [(4096, 4, 16), (4000, 4, 16)])
def test_tile1024_dag_on_twelve_streams(rt12, n, nblocks, nrhs):
    # the headline's tile size and stream count on a 4x4 grid: POTRF, TRSM
    # (beta = 0 product), SYRK and GEMM all run the 128x128 workgroup kernels
    R, K = rt12
    assert R.worker_counts()[1] == 12
    print(f"tile-1024 DAG n={n} nblocks={nblocks} nrhs={nrhs} on 12 streams", flush=True)
    K.set_launch_concurrency(12)
    try:
        assert K.gemm_tile_shape(False, 1024, 1024) == 128
        assert K.gemm_tile_shape(True, 1000, 1000) == 128
    finally:
        K.set_launch_concurrency(1)
    A0 = spd(n, n + 1)
    b = np.asfortranarray(np.random.default_rng(n).standard_normal((n, nrhs)))
    dA, dB = K.DeviceArray.from_host(A0), K.DeviceArray.from_host(b)
    st = R.factor_solve(dA.ptr, dA.ld, n, nblocks, dB.ptr, dB.ld, nrhs, home_node=R.hip_node(0))
    assert st.info == 0
    check(A0, dA.download(), nblocks)
    X = dB.download()
    Lref, _ = oracle.tiled_cholesky(A0, nblocks)
    Xref = oracle.tiled_potrs(np.tril(Lref), b, nblocks)
    assert np.linalg.norm(X - Xref) / np.linalg.norm(Xref) < 1e-12


@pytest.fixture
def rt_chain(rt):
    """Thirteen HIP streams per GPU, the first one set aside, on all compute
    units, for the tasks routed to it (STARPU_HIP_RESERVE_CUS=-1): the
    factorisations then put their critical chain there."""
    R, K = rt
    R.shutdown()
    env = {"STARPU_NWORKER_PER_HIP": "13", "STARPU_HIP_RESERVE_CUS": "-1"}
    os.environ.update(env)
    try:
        R.init()
        R.set_cpu_kernels()
        yield R, K
    finally:
        for k in env:
            os.environ.pop(k, None)
        R.shutdown()
        R.init()
        R.set_cpu_kernels()


def chain_workers(R, nblocks, crit):
    """Check the last run's trace: POTRF(k), TRSM(k+1,k) and SYRK(k+1,k+1)
    on the set-aside worker, every other task elsewhere."""
    tr = R.trace()
    on = {}
    for r in tr:
        key = (r["name"], r["worker"] == crit)
        on[key] = on.get(key, 0) + 1
    assert on.get(("POTRF", True), 0) == nblocks and on.get(("POTRF", False), 0) == 0, on
    assert on.get(("TRSM", True), 0) == nblocks - 1, on
    assert on.get(("SYRK", True), 0) == nblocks - 1, on
    for name in ("GEMM", "SOLVE_FWD", "SOLVE_BWD", "REFINE_solve", "REFINE_update"):
        assert on.get((name, True), 0) == 0, on


@pytest.mark.parametrize("n,nblocks,nrhs", [(4096, 4, 16), (2050, 8, 16)])
def test_chain_on_set_aside_stream(rt_chain, n, nblocks, nrhs):
    # the chain's tasks routed to the set-aside worker give the same factor
    # and solution as the unrouted graph (the oracle's)
    R, K = rt_chain
    crit = R.rt().starpu_hip_reserved_worker(0)
    assert crit >= 0 and R.worker_counts()[1] == 13
    print(f"routed chain n={n} nblocks={nblocks} nrhs={nrhs} on worker {crit}", flush=True)
    A0 = spd(n, n + 3)
    b = np.asfortranarray(np.random.default_rng(n + 1).standard_normal((n, nrhs)))
    dA, dB = K.DeviceArray.from_host(A0), K.DeviceArray.from_host(b)
    st = R.factor_solve(dA.ptr, dA.ld, n, nblocks, dB.ptr, dB.ld, nrhs, home_node=R.hip_node(0))
    assert st.info == 0
    chain_workers(R, nblocks, crit)
    check(A0, dA.download(), nblocks)
    Lref, _ = oracle.tiled_cholesky(A0, nblocks)
    Xref = oracle.tiled_potrs(np.tril(Lref), b, nblocks)
    X = dB.download()
    assert np.linalg.norm(X - Xref) / np.linalg.norm(Xref) < 1e-12


def test_mixed_chain_on_set_aside_stream(rt_chain):
    R, K = rt_chain
    crit = R.rt().starpu_hip_reserved_worker(0)
    n, nblocks, nrhs = 2048, 8, 16
    print(f"routed mixed chain n={n} nblocks={nblocks} on worker {crit}", flush=True)
    rng = np.random.default_rng(7)
    G = rng.standard_normal((n, n))
    A0 = np.asfortranarray(G @ G.T / n + 1e-3 * np.eye(n))
    b = np.asfortranarray(rng.standard_normal((n, nrhs)))
    dA, dB = K.DeviceArray.from_host(A0), K.DeviceArray.from_host(b)
    dX = K.DeviceArray(n, nrhs)
    R.rt().starpu_profiling_status_set(1)
    R.rt().starpu_trace_reset()
    try:
        st = R.mixed_solve(dA.ptr, n, n, nblocks, dB.ptr, n, nrhs, dX.ptr, n, prec="bf16x3", tol=1e-14, maxiter=40)
    finally:
        R.rt().starpu_profiling_status_set(0)
    assert st.info == 0 and st.converged == 1 and st.final_berr <= 1e-14
    chain_workers(R, nblocks, crit)
    X = dX.download()
    r = b - A0 @ X
    berr = max(np.linalg.norm(r[:, j]) / (np.linalg.norm(A0) * np.linalg.norm(X[:, j]) + np.linalg.norm(b[:, j]))
               for j in range(nrhs))
    assert berr <= 2e-14


@pytest.fixture
def rt_small_hbm(rt):
    """The runtime with a 16 MiB budget on the GPU's memory node
    (STARPU_LIMIT_HIP_MEM, src/drivers/hip/driver_hip.c:456-480), two
    streams and one task in flight per stream, so that at most about 6 MiB
    of tiles are in use by running tasks at any time."""
    R, K = rt
    R.shutdown()
    env = {"STARPU_LIMIT_HIP_MEM": "16", "STARPU_NWORKER_PER_HIP": "2", "STARPU_HIP_PIPELINE": "1"}
    os.environ.update(env)
    try:
        R.init()
        R.set_cpu_kernels()
        yield R, K
    finally:
        for k in env:
            os.environ.pop(k, None)
        R.shutdown()
        R.init()
        R.set_cpu_kernels()


def test_host_resident_factor_under_hbm_limit_evicts(rt_small_hbm):
    # A (32 MiB, 18 MiB of lower tiles) and the 16 diagonal-tile inverses
    # (8 MiB) do not fit the 16 MiB budget: tiles are evicted least recently
    # used first (memalloc.c:1142-1221), owners written back to main memory,
    # and fetched again when a later task reads them; the factor is unchanged
    R, K = rt_small_hbm
    n, nblocks = 2048, 8
    node = R.hip_node(0)
    assert R.rt().starpu_memory_get_total(node) == 16 << 20
    ev0 = R.memory_node_stats(node)[0]
    A0 = spd(n, 77)
    A = A0.copy(order="F")
    print("host-resident chol n=2048 nblocks=8 under a 16 MiB HBM budget", flush=True)
    st = R.factor(A.ctypes.data, n, n, nblocks, home_node=0)
    assert st.info == 0
    ev1 = R.memory_node_stats(node)[0]
    print(f"  evictions {ev1 - ev0}, used {R.rt().starpu_memory_get_used(node)}", flush=True)
    assert ev1 - ev0 > 0
    assert R.rt().starpu_memory_get_used(node) <= 16 << 20
    check(A0, A, nblocks)
    iu = np.triu_indices(n, 1)
    assert np.array_equal(A[iu], A0[iu])


def read_bus_matrix(path):
    rows = [l for l in open(path) if not l.startswith("#") and l.strip()]
    return np.array([[float(v) for v in r.split("\t")] for r in rows])


def test_bus_model_is_saved_and_reloaded(rt, tmp_path):
    # perfmodel_bus.c: the host<->GPU bandwidth and latency are measured
    # once, written to $STARPU_PERF_MODEL_DIR/bus/<host>.bandwidth and
    # .latency (STARPU_MAXNODES x STARPU_MAXNODES, NaN for absent nodes), and
    # loaded instead of measured at the next initialisation;
    # STARPU_BUS_CALIBRATE=1 measures again
    R, K = rt
    L = R.rt()
    env = {"STARPU_PERF_MODEL_DIR": str(tmp_path), "STARPU_HOSTNAME": "bustest"}
    saved = {k: os.environ.get(k) for k in env}
    R.shutdown()
    try:
        R.init(**env)
        node = R.hip_node(0)
        bw_path, lat_path = tmp_path / "bus" / "bustest.bandwidth", tmp_path / "bus" / "bustest.latency"
        assert bw_path.exists() and lat_path.exists()
        bw, lat = read_bus_matrix(bw_path), read_bus_matrix(lat_path)
        assert bw.shape == lat.shape and bw.shape[0] == bw.shape[1] >= 2
        measured = L.starpu_transfer_bandwidth(0, node)
        print(f"bus: host->GPU {measured:.0f} MB/s, latency {L.starpu_transfer_latency(0, node):.1f} us", flush=True)
        assert 1e3 < measured < 1e6
        assert abs(bw[0, node] - measured) <= 1e-5 * measured
        assert bw[0, 0] == 0.0 and np.isnan(bw[node + 1:, :]).all() and np.isnan(bw[:, node + 1:]).all()
        R.shutdown()
        # a planted value is loaded, not measured over
        bw[0, node] = 12345.0
        with open(bw_path, "w") as f:
            f.write("# edited\n")
            for row in bw:
                f.write("\t".join(f"{v:e}" for v in row) + "\n")
        R.init()
        assert L.starpu_transfer_bandwidth(0, node) == 12345.0
        R.shutdown()
        os.environ["STARPU_BUS_CALIBRATE"] = "1"
        R.init()
        assert L.starpu_transfer_bandwidth(0, node) != 12345.0
        assert read_bus_matrix(bw_path)[0, node] != 12345.0
    finally:
        os.environ.pop("STARPU_BUS_CALIBRATE", None)
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        R.shutdown()
        R.init()
        R.set_cpu_kernels()


@pytest.mark.parametrize("n,nblocks,nrhs", [(2048, 8, 16), (1000, 3, 5)])
def test_dmda_prefetch_host_resident(rt, n, nblocks, nrhs):
    # dmda starts moving a ready task's inputs to its GPU at push time
    # (deque_modeling_policy_data_aware.c:354-355) on a transfer stream; the
    # factor and X are bit-identical with prefetch off (STARPU_PREFETCH=0)
    R, K = rt
    node = R.hip_node(0)
    A0 = spd(n, 5 * n)
    b = np.asfortranarray(np.random.default_rng(n + 1).standard_normal((n, nrhs)))
    outs = []
    try:
        for pf in ("1", "0"):
            R.shutdown()
            os.environ["STARPU_PREFETCH"] = pf
            R.init()
            R.set_cpu_kernels()
            p0 = R.memory_node_stats(node)[2]
            A, X = A0.copy(order="F"), b.copy(order="F")
            print(f"host-resident factor_solve n={n} prefetch={pf}", flush=True)
            st = R.factor_solve(A.ctypes.data, n, n, nblocks, X.ctypes.data, n, nrhs, home_node=0)
            assert st.info == 0
            outs.append((A, X, R.memory_node_stats(node)[2] - p0))
    finally:
        os.environ.pop("STARPU_PREFETCH", None)
        R.shutdown()
        R.init()
        R.set_cpu_kernels()
    (A1, X1, pf1), (A2, X2, pf2) = outs
    assert pf1 > 0 and pf2 == 0
    assert np.array_equal(np.tril(A1), np.tril(A2)) and np.array_equal(X1, X2)
    check(A0, A1, nblocks)
