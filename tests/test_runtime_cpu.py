"""The StarPU-style runtime on CPU workers only (no GPU needed).

Mirrors the reference's test strategy (tests/main, tests/datawizard,
tests/sched_policies/simple_cpu_gpu_sched.c, tests/perfmodels/feed.c,
tests/fault-tolerance/retry.c): ordering under implicit dependencies,
scheduler placement with fake cost models, history-model feeding and the
v45 file format, tags, fault-tolerant retry, the watchdog, and the full
tiled Cholesky DAG against the oracle."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from oracle.codelets import addresses

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def helpers(tmp_path_factory):
    from starpu_amd import runtime as R
    R.rt()  # loads the runtime with RTLD_GLOBAL so the helper resolves against it
    out = tmp_path_factory.mktemp("rt") / "librt_codelets.so"
    subprocess.check_call(["gcc", "-O1", "-fPIC", "-shared", "-std=gnu11", f"-I{ROOT}/include",
                           os.path.join(HERE, "helpers", "rt_codelets.c"), "-o", str(out),
                           f"-L{ROOT}/starpu_amd", "-lstarpu_amd", f"-Wl,-rpath,{ROOT}/starpu_amd", "-pthread"])
    H = C.CDLL(str(out))
    for name, res, args in [
        ("tc_log_reset", None, []), ("tc_log_get", C.c_int, [C.c_void_p, C.c_int]),
        ("tc_insert_write", C.c_int, [C.c_void_p, C.c_int]), ("tc_insert_read", C.c_int, [C.c_void_p, C.c_int]),
        ("tc_ft_run", C.c_int, [C.c_void_p, C.c_int]), ("tc_place", C.c_int, [C.c_int, C.c_int, C.c_void_p]),
        ("tc_tags", C.c_int, []), ("tc_task_deps", C.c_int, []), ("tc_stall", C.c_int, [C.c_long]),
        ("tc_sync_tasks", C.c_int, [C.c_void_p, C.c_int]),
        ("tc_place_mask", C.c_int, [C.c_int, C.c_int, C.c_uint, C.c_void_p]),
    ]:
        f = getattr(H, name)
        f.restype, f.argtypes = res, args
    return H


@pytest.fixture
def runtime(tmp_path):
    from starpu_amd import runtime as R
    started = []

    def start(**env):
        base = dict(STARPU_NHIP=0, STARPU_NCPU=4, STARPU_SCHED="dmda",
                    STARPU_PERF_MODEL_DIR=str(tmp_path / "pm"), STARPU_CALIBRATE=0,
                    STARPU_WATCHDOG_TIMEOUT=0)
        base.update(env)
        R.init(**base)
        started.append(1)
        return R

    yield start
    if started:
        R.set_cpu_kernels()
        R.shutdown()
    for k in ("STARPU_WATCHDOG_TIMEOUT", "STARPU_CALIBRATE"):
        os.environ.pop(k, None)


def vector_handle(R, arr):
    L = R.rt()
    L.starpu_vector_data_register.restype = None
    L.starpu_vector_data_register.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_void_p, C.c_size_t, C.c_size_t]
    h = C.c_void_p()
    L.starpu_vector_data_register(C.byref(h), 0, arr.ctypes.data, arr.size, arr.itemsize)
    return h


def unregister(R, h):
    L = R.rt()
    L.starpu_data_unregister.restype = None
    L.starpu_data_unregister.argtypes = [C.c_void_p]
    L.starpu_data_unregister(h)


def read_log(H):
    buf = np.zeros(65536, dtype=np.int32)
    n = H.tc_log_get(buf.ctypes.data, buf.size)
    return list(buf[:n])


@pytest.mark.parametrize("sched", ["dmda", "eager", "prio", "random"])
def test_implicit_dependencies_order_accesses(runtime, helpers, sched):
    # sequential consistency (implicit_data_deps.c:215-434): writers run in
    # submission order; readers between two writers run after the first
    # and before the second.
    R = runtime(STARPU_SCHED=sched, STARPU_NCPU=6)
    arr = np.zeros(4, dtype=np.int32)
    h = vector_handle(R, arr)
    helpers.tc_log_reset()
    plan = []
    tid = 1
    for rnd in range(6):
        assert helpers.tc_insert_write(h, tid) == 0
        plan.append(("w", tid))
        tid += 1
        for _ in range(4):
            assert helpers.tc_insert_read(h, tid) == 0
            plan.append(("r", tid))
            tid += 1
    R.rt().starpu_task_wait_for_all()
    log = read_log(helpers)
    pos = {v: i for i, v in enumerate(log)}
    writers = [t for k, t in plan if k == "w"]
    assert [pos[w] for w in writers] == sorted(pos[w] for w in writers)
    for i, w in enumerate(writers):
        readers = [t for k, t in plan if k == "r" and w < t < (writers[i + 1] if i + 1 < len(writers) else 10 ** 9)]
        for r in readers:
            assert pos[w] < pos[-r]
            if i + 1 < len(writers):
                assert pos[-r] < pos[writers[i + 1]]
    unregister(R, h)
    assert arr[0] == len(writers)   # written back to the home node


def test_synchronous_created_tasks(runtime, helpers):
    # starpu_task_create() + synchronous = 1 with detach at its default: the
    # submit returns after the task ran, and the task is freed by the waiter
    # only (the reference clears detach for synchronous tasks,
    # src/core/task.c:979-985); each task incremented element 0 once
    R = runtime(STARPU_NCPU=4)
    v = np.zeros(4, dtype=np.int32)
    h = vector_handle(R, v)
    assert helpers.tc_sync_tasks(h, 200) == 0
    unregister(R, h)
    assert v[0] == 200


def test_fault_tolerant_retry(runtime, helpers):
    # tests/fault-tolerance/retry.c: a codelet reports failure N times, the
    # task is re-executed until it succeeds, and the result is correct.
    R = runtime(STARPU_NCPU=2)
    L = R.rt()
    L.starpu_variable_data_register.restype = None
    L.starpu_variable_data_register.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_void_p, C.c_size_t]
    x = np.zeros(1)
    h = C.c_void_p()
    L.starpu_variable_data_register(C.byref(h), 0, x.ctypes.data, 8)
    runs = helpers.tc_ft_run(h, 10)
    unregister(R, h)
    assert runs == 11
    assert x[0] == 42.0


def test_dmda_follows_fake_cost_model(runtime, helpers):
    # tests/sched_policies/simple_cpu_gpu_sched.c: a cost function that
    # predicts 1 us on one worker and 1e7 us elsewhere pins every task there.
    R = runtime(STARPU_NCPU=4)
    out = np.zeros(1024, dtype=np.int32)
    for fast in (0, 2):
        n = helpers.tc_place(40, fast, out.ctypes.data)
        assert n == 40
        assert set(out[:n]) == {fast}


def test_task_workerids_restrict_placement(runtime, helpers):
    # STARPU_TASK_WORKERIDS (reference include/starpu_task.h:1410-1432): only
    # workers whose bit is set may run the task, whatever the cost model says
    R = runtime(STARPU_NCPU=4, STARPU_SCHED="dmda")
    out = np.zeros(1024, dtype=np.int32)
    n = helpers.tc_place_mask(60, 0, 0b1010, out.ctypes.data)   # worker 0 is the fast one
    assert n == 60
    assert set(out[:n]) <= {1, 3}


def test_tags_and_explicit_task_deps(runtime, helpers):
    R = runtime(STARPU_NCPU=4)
    helpers.tc_log_reset()
    assert helpers.tc_tags() == 0
    log = read_log(helpers)
    assert log.index(3) > log.index(1) and log.index(3) > log.index(2)
    helpers.tc_log_reset()
    assert helpers.tc_task_deps() == 2        # both callbacks ran
    log = read_log(helpers)
    assert log == [0, 1]                      # b declared to depend on a


def test_watchdog_warns_on_stalled_task(runtime, helpers, capfd):
    R = runtime(STARPU_NCPU=1, STARPU_WATCHDOG_TIMEOUT=100000)   # 0.1 s
    helpers.tc_stall(600000)                                      # 0.6 s
    err = capfd.readouterr().err
    assert "watchdog" in err


@pytest.mark.parametrize("sched", ["dmda", "eager", "prio"])
@pytest.mark.parametrize("n,nblocks", [(512, 4), (300, 7), (1000, 8), (96, 1)])
def test_tiled_cholesky_on_cpu_workers_matches_oracle(runtime, sched, n, nblocks):
    # the whole DAG of cholesky_implicit.c on CPU workers running the oracle
    # codelets: every tile sees the same operations in the same order as the
    # sequential oracle, so the factor is bit-identical.
    R = runtime(STARPU_SCHED=sched, STARPU_NCPU=4)
    R.set_cpu_kernels(**addresses())
    rng = np.random.default_rng(n * nblocks)
    B = rng.uniform(-1, 1, (n, n))
    A = np.asfortranarray(B @ B.T + n * np.eye(n))
    A0 = A.copy(order="F")
    st = R.factor(A.ctypes.data, n, n, nblocks)
    ref, info = oracle.tiled_cholesky(A0, nblocks)
    assert st.info == info == 0
    assert np.array_equal(np.tril(A), np.tril(ref))
    assert np.array_equal(np.triu(A, 1), np.triu(A0, 1))
    # POTRF (which also forms the tile inverses), TRSM, SYRK, GEMM
    nt = nblocks
    assert st.ntasks == nt + nt * (nt - 1) + nt * (nt - 1) * (nt - 2) // 6


@pytest.mark.parametrize("prio", ["none", "base", "kut", "timings", "parsec"])
def test_priority_attributions_give_the_same_factor(runtime, prio):
    # the reference's four priority attributions (and none) reorder the
    # tasks, never the operations on a tile: the factor is bit-identical
    R = runtime(STARPU_SCHED="dmda", STARPU_NCPU=4)
    R.set_cpu_kernels(**addresses())
    n, nblocks = 384, 6
    rng = np.random.default_rng(7)
    B = rng.uniform(-1, 1, (n, n))
    A = np.asfortranarray(B @ B.T + n * np.eye(n))
    A0 = A.copy(order="F")
    st = R.factor(A.ctypes.data, n, n, nblocks, prio=prio)
    ref, info = oracle.tiled_cholesky(A0, nblocks)
    assert st.info == info == 0
    assert np.array_equal(np.tril(A), np.tril(ref))


@pytest.mark.parametrize("sched", ["dmda", "eager"])
@pytest.mark.parametrize("n,nblocks,nrhs", [(512, 4, 16), (300, 7, 3), (96, 1, 1)])
def test_factor_solve_on_cpu_workers_matches_oracle(runtime, sched, n, nblocks, nrhs):
    # factorisation and solve submitted as one DAG: each right-hand-side
    # block sees the same tile operations in the same order as the oracle's
    # sequential tiled solve, so X is bit-identical to it.
    R = runtime(STARPU_SCHED=sched, STARPU_NCPU=4)
    R.set_cpu_kernels(**addresses())
    rng = np.random.default_rng(n + nrhs)
    Bm = rng.uniform(-1, 1, (n, n))
    A = np.asfortranarray(Bm @ Bm.T + n * np.eye(n))
    A0 = A.copy(order="F")
    b = np.asfortranarray(rng.uniform(-1, 1, (n, nrhs)))
    X = b.copy(order="F")
    st = R.factor_solve(A.ctypes.data, n, n, nblocks, X.ctypes.data, n, nrhs)
    Lref, info = oracle.tiled_cholesky(A0, nblocks)
    Xref = oracle.tiled_potrs(np.tril(Lref), b, nblocks)
    assert st.info == info == 0
    assert np.array_equal(np.tril(A), np.tril(Lref))
    assert np.array_equal(X, Xref)
    nt = nblocks
    # factorisation + one forward and one backward task per block row (the
    # POTRF tasks form the inverses the solve uses)
    assert st.ntasks == nt + nt * (nt - 1) + nt * (nt - 1) * (nt - 2) // 6 + 2 * nt
    # a second solve with the resident factor reproduces X
    X2 = b.copy(order="F")
    st2 = R.solve(A.ctypes.data, n, n, nblocks, X2.ctypes.data, n, nrhs)
    assert np.array_equal(X2, Xref)
    assert st2.flops == 2.0 * n * n * nrhs


def test_not_spd_on_cpu_workers(runtime):
    R = runtime(STARPU_NCPU=3)
    R.set_cpu_kernels(**addresses())
    n, nb = 200, 5
    i = np.arange(n, dtype=float)
    A = np.asfortranarray(1.0 / (1.0 + i[:, None] + i[None, :]) + n * np.eye(n))
    A[123, 123] = -5.0
    W = A.copy(order="F")
    st = R.factor(W.ctypes.data, n, n, nb)
    _, info = oracle.tiled_cholesky(A, nb)
    assert st.info == info == 124


def test_history_model_calibrates_saves_and_reloads(runtime, tmp_path):
    # tests/perfmodels/feed.c semantics via the Cholesky models: run the DAG
    # with calibration on, check the saved v45 files parse back.
    R = runtime(STARPU_NCPU=4, STARPU_CALIBRATE=1)
    R.set_cpu_kernels(**addresses())
    n, nb = 512, 8
    rng = np.random.default_rng(0)
    B = rng.uniform(-1, 1, (n, n))
    for _ in range(3):
        A = np.asfortranarray(B @ B.T + n * np.eye(n))
        R.factor(A.ctypes.data, n, n, nb)
    R.set_cpu_kernels()
    R.shutdown()
    R.init(STARPU_NHIP=0, STARPU_NCPU=1, STARPU_SCHED="dmda", STARPU_PERF_MODEL_DIR=str(tmp_path / "pm"))
    d = tmp_path / "pm" / "codelets" / "45"
    files = sorted(os.listdir(d))
    assert any(f.startswith("chol_model_gemm_d.") for f in files)
    gemm = [f for f in files if f.startswith("chol_model_gemm_d.")][0]
    text = open(d / gemm).read()
    assert text.splitlines()[2] == "45"
    # 3 runs x 56 GEMMs, the first sample of the footprint dropped
    lines = [ln for ln in text.splitlines() if ln and not ln.startswith("#")]
    entry = lines[-1].split()
    assert int(entry[-1]) + int(0) <= 3 * 56 - 1
    assert float(entry[2]) == pytest.approx(2 * 64 ** 3)       # flops of a 64^3 tile GEMM


def test_reference_sampling_file_parses(runtime, tmp_path):
    # v45 file written by the reference (tools/perfmodels/sampling fixture)
    R = runtime(STARPU_NCPU=1)
    L = R.rt()

    class Model(C.Structure):
        _fields_ = [("type", C.c_int), ("cost_function", C.c_void_p), ("arch_cost_function", C.c_void_p),
                    ("worker_cost_function", C.c_void_p), ("size_base", C.c_void_p), ("footprint", C.c_void_p),
                    ("symbol", C.c_char_p), ("is_loaded", C.c_uint), ("benchmarking", C.c_uint),
                    ("is_init", C.c_uint), ("state", C.c_void_p)]

    class Dev(C.Structure):
        _fields_ = [("type", C.c_int), ("devid", C.c_int), ("ncores", C.c_int)]

    class Arch(C.Structure):
        _fields_ = [("ndevices", C.c_int), ("devices", C.POINTER(Dev))]

    m = Model(type=4, symbol=b"chol_model_gemm")   # STARPU_HISTORY_BASED
    L.starpu_perfmodel_load_file.argtypes = [C.c_char_p, C.POINTER(Model)]
    L.starpu_perfmodel_load_file.restype = C.c_int
    path = os.path.join(HERE, "golden", "chol_model_gemm.sirocco")
    assert L.starpu_perfmodel_load_file(path.encode(), C.byref(m)) == 0
    dev = Dev(1, 1, 1)   # CUDA device 1 (comb 2 of the file)
    arch = Arch(1, C.pointer(dev))
    L.starpu_perfmodel_history_count.argtypes = [C.POINTER(Model), C.POINTER(Arch)]
    L.starpu_perfmodel_history_entry.argtypes = [C.POINTER(Model), C.POINTER(Arch), C.c_int,
                                                 C.POINTER(C.c_uint32), C.POINTER(C.c_size_t),
                                                 C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                 C.POINTER(C.c_double), C.POINTER(C.c_uint)]
    assert L.starpu_perfmodel_history_count(C.byref(m), C.byref(arch)) == 4
    fp, size, flops, mean, dev_, ns = C.c_uint32(), C.c_size_t(), C.c_double(), C.c_double(), C.c_double(), C.c_uint()
    found = {}
    for i in range(4):
        L.starpu_perfmodel_history_entry(C.byref(m), C.byref(arch), i, C.byref(fp), C.byref(size), C.byref(flops),
                                         C.byref(mean), C.byref(dev_), C.byref(ns))
        found[fp.value] = (size.value, flops.value, mean.value, ns.value)
    # chol_model_gemm.sirocco:40, the 960-tile SGEMM on a K40m: 2*960^3 flop
    assert found[0x24c84a50][1] == 2 * 960 ** 3
    assert found[0x24c84a50][2] == pytest.approx(754.5142)
    assert found[0x24c84a50][3] == 15561
    # dump and re-read round trip
    L.starpu_perfmodel_dump_file.argtypes = [C.POINTER(Model), C.c_char_p]
    out = str(tmp_path / "roundtrip")
    assert L.starpu_perfmodel_dump_file(C.byref(m), out.encode()) == 0
    m2 = Model(type=4, symbol=b"chol_model_gemm")
    assert L.starpu_perfmodel_load_file(out.encode(), C.byref(m2)) == 0
    assert L.starpu_perfmodel_history_count(C.byref(m2), C.byref(arch)) == 4
    # the writer reproduces the reference writer's file (perfmodel_history.c:
    # 834-968) byte for byte outside its comment lines (the comments carry
    # the writer's own combination numbers)
    data = lambda p: [l for l in open(p, "rb").read().split(b"\n") if not l.startswith(b"#")]
    assert data(out) == data(path)
    assert len(open(out, "rb").read().split(b"\n")) == len(open(path, "rb").read().split(b"\n"))


def test_crc32c_footprint_hash_properties(runtime):
    # CRC32C (poly 0x1EDC6F41, MSB first, src/common/hash.c:22-81): appending
    # the CRC (big-endian bytes) to the message gives a zero remainder, and
    # the hash is linear over GF(2) in the message for a fixed length.
    R = runtime(STARPU_NCPU=1)
    L = R.rt()
    rng = np.random.default_rng(3)
    for _ in range(20):
        msg = rng.integers(0, 256, rng.integers(1, 40)).astype(np.uint8)
        crc = L.starpu_hash_crc32c_be_n(msg.ctypes.data, msg.size, 0)
        full = np.concatenate([msg, np.frombuffer(int(crc).to_bytes(4, "big"), dtype=np.uint8)])
        assert L.starpu_hash_crc32c_be_n(full.ctypes.data, full.size, 0) == 0
        m2 = rng.integers(0, 256, msg.size).astype(np.uint8)
        c2 = L.starpu_hash_crc32c_be_n(m2.ctypes.data, m2.size, 0)
        x = np.bitwise_xor(msg, m2)
        cx = L.starpu_hash_crc32c_be_n(x.ctypes.data, x.size, 0)
        assert cx == crc ^ c2
    # crc32c_be(x) over the 4 little-endian bytes of x
    v = 0x12345678
    b = np.frombuffer(v.to_bytes(4, "little"), dtype=np.uint8).copy()
    assert L.starpu_hash_crc32c_be(v, 7) == L.starpu_hash_crc32c_be_n(b.ctypes.data, 4, 7)


def test_blas_library_found_and_dgemm_agrees():
    from oracle.codelets import blas_addresses, blas_dgemm_nt, blas_library
    assert blas_library() is not None
    blas_addresses()
    rng = np.random.default_rng(3)
    A, B, C0 = rng.uniform(-1, 1, (37, 19)), rng.uniform(-1, 1, (23, 19)), rng.uniform(-1, 1, (37, 23))
    got = blas_dgemm_nt(-1.0, A, B, 0.5, C0)
    exp = oracle.dgemm_nt(-1.0, A, B, 0.5, C0)
    assert np.linalg.norm(got - exp) <= 1e-14 * np.linalg.norm(exp)
    # small integers: exact whatever the summation order
    Ai, Bi = rng.integers(-9, 10, (40, 30)).astype(float), rng.integers(-9, 10, (50, 30)).astype(float)
    assert np.array_equal(blas_dgemm_nt(1.0, Ai, Bi, 0.0, np.zeros((40, 50))), Ai @ Bi.T)


@pytest.mark.parametrize("sched", ["dmda", "eager"])
@pytest.mark.parametrize("n,nblocks,nrhs", [(512, 4, 16), (300, 7, 3), (1000, 3, 1)])
def test_blas_codelets_dag_matches_oracle(runtime, sched, n, nblocks, nrhs):
    # the reference's CPU path: its codelets' BLAS/LAPACK calls on the tiles
    # (the CPU baseline of the benchmark); a different summation order than
    # the oracle's loops, so the bar is the FP64 parity bar, not bit equality
    from oracle.codelets import blas_addresses
    R = runtime(STARPU_SCHED=sched, STARPU_NCPU=4)
    R.set_cpu_kernels(**blas_addresses())
    rng = np.random.default_rng(n + 7 * nrhs)
    Bm = rng.uniform(-1, 1, (n, n))
    A = np.asfortranarray(Bm @ Bm.T + n * np.eye(n))
    A0 = A.copy(order="F")
    b = np.asfortranarray(rng.uniform(-1, 1, (n, nrhs)))
    X = b.copy(order="F")
    st = R.factor_solve(A.ctypes.data, n, n, nblocks, X.ctypes.data, n, nrhs)
    Lref, info = oracle.tiled_cholesky(A0, nblocks)
    assert st.info == info == 0
    assert oracle.rel_diff_lower(np.tril(A), np.tril(Lref)) <= 1e-13
    be, be_ref = oracle.backward_error_llt(A0, np.tril(A)), oracle.backward_error_llt(A0, np.tril(Lref))
    assert be <= 10 * max(be_ref, 1e-17)
    Xref = oracle.tiled_potrs(np.tril(Lref), b, nblocks)
    assert np.linalg.norm(X - Xref) <= 1e-12 * np.linalg.norm(Xref)
    assert np.array_equal(np.triu(A, 1), np.triu(A0, 1))


def test_blas_codelets_not_spd_info(runtime):
    from oracle.codelets import blas_addresses
    R = runtime(STARPU_NCPU=2)
    R.set_cpu_kernels(**blas_addresses())
    n, nb = 256, 4
    rng = np.random.default_rng(9)
    Bm = rng.uniform(-1, 1, (n, n))
    A = np.asfortranarray(Bm @ Bm.T + n * np.eye(n))
    A[100, 100] = -5.0
    A0 = A.copy(order="F")
    st = R.factor(A.ctypes.data, n, n, nb)
    _, info = oracle.tiled_cholesky(A0, nb)
    assert st.info == info == 101


def test_memory_reservations_against_node_budget(runtime):
    # starpu_memory_allocate / _deallocate / _wait_available
    # (memory_manager.c:66-141,207-222) on main RAM with a 1 MiB budget
    import threading
    import time as _t
    R = runtime(STARPU_NCPU=1, STARPU_LIMIT_CPU_MEM=1)
    L = R.rt()
    MiB = 1 << 20
    try:
        base = L.starpu_memory_get_used(0)
        assert L.starpu_memory_get_total(0) == MiB
        assert L.starpu_memory_allocate(0, 600 * 1024, 0) == 0
        assert L.starpu_memory_get_used(0) == base + 600 * 1024
        assert L.starpu_memory_allocate(0, 600 * 1024, 0) == -12          # -ENOMEM
        assert L.starpu_memory_allocate(0, 600 * 1024, R.STARPU_MEMORY_OVERFLOW) == 0
        assert L.starpu_memory_get_available(0) < 0                        # overflowed
        assert L.starpu_memory_allocate(0, 2 * MiB, R.STARPU_MEMORY_WAIT) == -12   # never fits
        L.starpu_memory_deallocate(0, 600 * 1024)
        # a waiter blocks until another thread returns enough
        got = []

        def waiter():
            got.append(L.starpu_memory_allocate(0, 700 * 1024, R.STARPU_MEMORY_WAIT))

        th = threading.Thread(target=waiter)
        th.start()
        _t.sleep(0.2)
        assert not got                     # still waiting: 600 KiB held of 1 MiB
        L.starpu_memory_deallocate(0, 100 * 1024)   # 500 KiB held: still short
        _t.sleep(0.1)
        assert not got
        L.starpu_memory_deallocate(0, 500 * 1024)   # all returned
        th.join(5)
        assert got == [0]
        assert R.memory_node_stats(0)[1] == 1
        # wait_available reserves nothing
        L.starpu_memory_wait_available(0, 200 * 1024)
        assert L.starpu_memory_get_used(0) == base + 700 * 1024
        L.starpu_memory_deallocate(0, 700 * 1024)
        assert L.starpu_memory_get_used(0) == base
    finally:
        os.environ.pop("STARPU_LIMIT_CPU_MEM", None)
