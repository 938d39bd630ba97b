"""ctypes binding of the starpu_amd C runtime (libstarpu_amd.so) and of the
tiled Cholesky application (libstarpu_amd_chol.so).

The runtime reads its configuration from the environment at init, with the
reference's variable names (STARPU_SCHED, STARPU_NCPU, STARPU_NHIP,
STARPU_NWORKER_PER_HIP, STARPU_HIP_PIPELINE, STARPU_CALIBRATE,
STARPU_PERF_MODEL_DIR, STARPU_LIMIT_HIP_MEM, ...); ``init(**env)`` sets the
given ones before calling starpu_init.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_RT = None
_CHOL = None

STARPU_MAIN_RAM = 0
STARPU_CPU_WORKER = 0
STARPU_HIP_WORKER = 7


def _load(name: str) -> C.CDLL:
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build with `python -m starpu_amd._build`")
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


class CholStats(C.Structure):
    _fields_ = [
        ("elapsed_us", C.c_double),
        ("flops", C.c_double),
        ("info", C.c_int),
        ("ntasks", C.c_long),
        ("kernel_us", C.c_double * 7),
        ("kernel_n", C.c_long * 7),
    ]


class CholMixedStats(C.Structure):
    _fields_ = [
        ("elapsed_us", C.c_double),
        ("convert_us", C.c_double),
        ("factor_us", C.c_double),
        ("refine_us", C.c_double),
        ("flops_factor", C.c_double),
        ("info", C.c_int),
        ("iterations", C.c_int),
        ("converged", C.c_int),
        ("ntasks", C.c_long),
        ("final_berr", C.c_double),
        ("berr", C.c_double * 32),
    ]


PREC = {"fp32": 0, "bf16": 1, "bf16x3": 2}


class ServeRequest(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("nrhs", C.c_int), ("matrix", C.c_int),
        ("arrival_us", C.c_double),
        ("A", C.c_void_p), ("lda", C.c_int),
        ("B", C.c_void_p), ("ldb", C.c_int),
        ("X", C.c_void_p), ("ldx", C.c_int),
        ("start_us", C.c_double), ("finish_us", C.c_double),
        ("berr", C.c_double),
        ("iterations", C.c_int), ("status", C.c_int), ("factored", C.c_int),
        ("factor_start_us", C.c_double), ("factor_end_us", C.c_double),
    ]


class ServeConfig(C.Structure):
    _fields_ = [
        ("budget_bytes", C.c_double), ("slots", C.c_int), ("nb", C.c_int), ("prec", C.c_int),
        ("maxiter", C.c_int), ("prio", C.c_int), ("tol", C.c_double),
    ]


class ServeStats(C.Structure):
    _fields_ = [
        ("wall_us", C.c_double), ("peak_bytes", C.c_double),
        ("factorizations", C.c_int), ("cache_hits", C.c_int), ("evictions", C.c_int),
        ("memory_waits", C.c_int), ("max_inflight", C.c_int),
    ]

KINDS = ("potrf", "trsm", "syrk", "gemm", "solve_fwd", "solve_bwd", "trtri")


class CholCpuKernels(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in KINDS]


class TraceRecord(C.Structure):
    _fields_ = [
        ("name", C.c_char_p),
        ("workerid", C.c_int),
        ("submit_us", C.c_double),
        ("start_us", C.c_double),
        ("end_us", C.c_double),
        ("kernel_us", C.c_double),
        ("bytes_in", C.c_size_t),
        ("priority", C.c_int),
        ("push_us", C.c_double),
    ]


def rt() -> C.CDLL:
    global _RT
    if _RT is None:
        # the kernels library first: the Cholesky app links both
        _load("libstarpu_amd_kernels.so")
        L = _load("libstarpu_amd.so")
        vp, i, u, d = C.c_void_p, C.c_int, C.c_uint, C.c_double
        sig = {
            "starpu_init": (i, [vp]),
            "starpu_shutdown": (None, []),
            "starpu_is_initialized": (i, []),
            "starpu_worker_get_count": (u, []),
            "starpu_cpu_worker_get_count": (u, []),
            "starpu_hip_worker_get_count": (u, []),
            "starpu_worker_get_type": (i, [i]),
            "starpu_hip_memory_node": (i, [i]),
            "starpu_hip_reserved_worker": (i, [i]),
            "starpu_task_wait_for_all": (i, []),
            "starpu_timing_now": (d, []),
            "starpu_sched_get_policy_name": (C.c_char_p, []),
            "starpu_trace_count": (C.c_size_t, []),
            "starpu_trace_get": (i, [C.c_size_t, C.POINTER(TraceRecord)]),
            "starpu_trace_reset": (None, []),
            "starpu_trace_dump_chrome": (i, [C.c_char_p]),
            "starpu_profiling_status_set": (i, [i]),
            "starpu_worker_stats": (i, [u, C.POINTER(C.c_ulong), C.POINTER(d), C.POINTER(d)]),
            "starpu_memory_get_used": (C.c_longlong, [u]),
            "starpu_memory_get_available": (C.c_longlong, [u]),
            "starpu_transfer_bandwidth": (d, [u, u]),
            "starpu_transfer_latency": (d, [u, u]),
            "starpu_memory_allocate": (i, [u, C.c_size_t, i]),
            "starpu_memory_deallocate": (None, [u, C.c_size_t]),
            "starpu_memory_wait_available": (None, [u, C.c_size_t]),
            "starpu_memory_node_stats": (i, [u, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                             C.POINTER(C.c_longlong)]),
            "starpu_memory_get_total": (C.c_longlong, [u]),
            "starpu_hash_crc32c_be": (C.c_uint32, [C.c_uint32, C.c_uint32]),
            "starpu_hash_crc32c_be_n": (C.c_uint32, [vp, C.c_size_t, C.c_uint32]),
            "starpu_hash_crc32c_string": (C.c_uint32, [C.c_char_p, C.c_uint32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _RT = L
    return _RT


def chol() -> C.CDLL:
    global _CHOL
    if _CHOL is None:
        rt()
        L = _load("libstarpu_amd_chol.so")
        L.chol_factor.restype = C.c_int
        L.chol_factor.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(CholStats)]
        L.chol_solve.restype = C.c_int
        L.chol_solve.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                 C.c_int, C.c_int, C.POINTER(CholStats)]
        L.chol_factor_solve.restype = C.c_int
        L.chol_factor_solve.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                        C.c_int, C.c_int, C.POINTER(CholStats)]
        L.chol_mixed_solve.restype = C.c_int
        L.chol_mixed_solve.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                       C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                       C.POINTER(CholMixedStats)]
        L.chol_serve.restype = C.c_int
        L.chol_serve.argtypes = [C.c_int, C.POINTER(ServeConfig), C.POINTER(ServeRequest), C.c_int,
                                 C.POINTER(ServeStats)]
        L.chol_flops_potrs.restype = C.c_double
        L.chol_flops_potrs.argtypes = [C.c_double, C.c_double]
        L.chol_set_cpu_kernels.restype = None
        L.chol_set_cpu_kernels.argtypes = [C.POINTER(CholCpuKernels)]
        L.chol_flops_potrf.restype = C.c_double
        L.chol_flops_potrf.argtypes = [C.c_double]
        _CHOL = L
    return _CHOL


def init(**env) -> None:
    """Set STARPU_* variables (values converted to str) then starpu_init(NULL)."""
    for k, v in env.items():
        os.environ[k] = str(v)
    r = rt().starpu_init(None)
    if r != 0:
        raise RuntimeError(f"starpu_init failed: {r}")


def shutdown() -> None:
    rt().starpu_shutdown()


def worker_counts() -> tuple[int, int]:
    L = rt()
    return L.starpu_cpu_worker_get_count(), L.starpu_hip_worker_get_count()


STARPU_MEMORY_WAIT = 1 << 4
STARPU_MEMORY_OVERFLOW = 1 << 5


def memory_node_stats(node: int) -> tuple[int, int, int]:
    """(replicates evicted, memory-reservation waits, replicates prefetched) on a memory node."""
    ev, wt, pf = C.c_longlong(), C.c_longlong(), C.c_longlong()
    rt().starpu_memory_node_stats(node, C.byref(ev), C.byref(wt), C.byref(pf))
    return ev.value, wt.value, pf.value


def hip_node(dev: int = 0) -> int:
    return rt().starpu_hip_memory_node(dev)


def set_cpu_kernels(**kernels) -> None:
    """Install CPU codelet implementations (C function addresses, keyed by
    the names in KINDS) for the Cholesky tasks; no arguments removes them."""
    if not any(kernels.values()):
        chol().chol_set_cpu_kernels(None)
        return
    unknown = set(kernels) - set(KINDS)
    if unknown:
        raise KeyError(f"unknown codelets {sorted(unknown)}")
    k = CholCpuKernels(*(kernels.get(n) for n in KINDS))
    chol().chol_set_cpu_kernels(C.byref(k))


PRIO_SCHEMES = {"none": 0, "base": 1, "kut": 2, "timings": 3, "parsec": 4}


def _prio(prio) -> int:
    """True/False, a scheme name of PRIO_SCHEMES, or its number."""
    if isinstance(prio, str):
        return PRIO_SCHEMES[prio]
    if isinstance(prio, bool):
        return 1 if prio else 0
    return int(prio)


def factor(ptr: int, lda: int, n: int, nblocks: int, home_node: int = STARPU_MAIN_RAM,
           prio=True) -> CholStats:
    st = CholStats()
    r = chol().chol_factor(home_node, ptr, lda, n, nblocks, _prio(prio), C.byref(st))
    if r != 0:
        raise RuntimeError(f"chol_factor failed: {r}")
    return st


def solve(lptr: int, ldl: int, n: int, nblocks: int, bptr: int, ldb: int, nrhs: int,
          home_node: int = STARPU_MAIN_RAM, prio=True) -> CholStats:
    st = CholStats()
    r = chol().chol_solve(home_node, lptr, ldl, n, nblocks, bptr, ldb, nrhs, _prio(prio), C.byref(st))
    if r != 0:
        raise RuntimeError(f"chol_solve failed: {r}")
    return st


def factor_solve(ptr: int, lda: int, n: int, nblocks: int, bptr: int, ldb: int, nrhs: int,
                 home_node: int = STARPU_MAIN_RAM, prio=True) -> CholStats:
    st = CholStats()
    r = chol().chol_factor_solve(home_node, ptr, lda, n, nblocks, bptr, ldb, nrhs, _prio(prio),
                                 C.byref(st))
    if r != 0:
        raise RuntimeError(f"chol_factor_solve failed: {r}")
    return st


def mixed_solve(aptr: int, lda: int, n: int, nblocks: int, bptr: int, ldb: int, nrhs: int, xptr: int,
                ldx: int, prec: str = "bf16x3", tol: float = 1e-14, maxiter: int = 30, hip_node: int | None = None,
                prio=True) -> CholMixedStats:
    """Mixed-precision factorisation + FP64 iterative refinement, all on the
    GPU (A, B, X are device pointers; A holds both triangles)."""
    st = CholMixedStats()
    node = hip_node if hip_node is not None else hip_node_of(0)
    r = chol().chol_mixed_solve(node, aptr, lda, n, nblocks, bptr, ldb, nrhs, xptr, ldx, PREC[prec], tol,
                                maxiter, _prio(prio), C.byref(st))
    if r != 0:
        raise RuntimeError(f"chol_mixed_solve failed: {r}")
    return st


def serve(cfg: ServeConfig, reqs, hip_node: int | None = None) -> ServeStats:
    """Run chol_serve over a ctypes array of ServeRequest (results filled in)."""
    st = ServeStats()
    node = hip_node if hip_node is not None else hip_node_of(0)
    r = chol().chol_serve(node, C.byref(cfg), reqs, len(reqs), C.byref(st))
    if r != 0:
        raise RuntimeError(f"chol_serve failed: {r}")
    return st


def hip_node_of(dev: int = 0) -> int:
    return hip_node(dev)


def trace() -> list[dict]:
    L = rt()
    out = []
    rec = TraceRecord()
    for i in range(L.starpu_trace_count()):
        if L.starpu_trace_get(i, C.byref(rec)) == 0:
            out.append({"name": rec.name.decode() if rec.name else "", "worker": rec.workerid,
                        "submit": rec.submit_us, "start": rec.start_us, "end": rec.end_us,
                        "kernel_us": rec.kernel_us, "bytes_in": rec.bytes_in, "prio": rec.priority,
                        "push": rec.push_us})
    return out
