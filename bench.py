#!/usr/bin/env python3
"""Benchmark: tiled FP64 Cholesky factor-and-solve, N=32768, tile 1024
(32 x 32 tiles), 16 right-hand sides, on one MI355X through the
StarPU-style runtime.

Workload: A = B*B^T/N + I with B i.i.d. N(0,1) from a seed, and 16
right-hand sides N(0,1), generated on the GPU.  A step is one
factorisation by the POTRF/TRSM/SYRK/GEMM task DAG of
examples/cholesky/cholesky_implicit.c followed, in the same task graph,
by the forward and backward tile solves, under dmda with HIP workers.  A
and B are resident in HBM before the timed region; each step first
restores them from pristine copies (untimed), then is timed from the
first task submission to starpu_task_wait_for_all
(cholesky_implicit.c:233,325-326).
value = (FLOPS_DPOTRF(N) + 2 N^2 nrhs) / mean step time.

Prints one JSON line (see the keys below).  --gpus above 1 is not
supported: the benchmark measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Twelve HIP streams per GPU (STARPU_NWORKER_PER_HIP) need a hardware queue
# each to run concurrently, and the library's other streams take some more;
# HIP reads this once, at its initialisation.
os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(16, int(os.environ.get("GPU_MAX_HW_QUEUES") or 0))))

FP64_PEAK_TFLOPS = 78.6     # MI355X dense FP64 matrix peak (AMD public spec)
# the other MI355X dense peaks (AMD public spec) and HBM3E bandwidth
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
HBM_TBS = 8.0


def mixed_roofline(n: int, nb: int, prec: str, solves: int) -> dict:
    """Time to solution of the mixed solve if every phase ran at its own
    roof: the conversion (A's lower triangle read in FP64, written as the
    FP32 upper triangle) at HBM bandwidth; the diagonal tiles' factor and
    inverse (2 nb^3/3 flop each) at the FP64 MFMA peak; the other factor
    flops (N^3/3 minus the diagonal tiles') at the update precision's peak,
    three BF16 products each for BF16x3; per refinement step the solve's
    two passes over the FP32 upper triangle and the residual's pass over
    A's FP64 lower triangle (4 N^2 bytes each) at HBM bandwidth."""
    nt = n // nb
    f_diag = nt * 2.0 * nb ** 3 / 3.0
    f_upd = n ** 3 / 3.0 - nt * nb ** 3 / 3.0
    upd_peak = {"bf16x3": PEAK_TFLOPS["bf16"] / 3.0, "bf16": PEAK_TFLOPS["bf16"], "fp32": PEAK_TFLOPS["fp32"]}[prec]
    t_conv = (4.0 * n * n + 2.0 * n * n) / (HBM_TBS * 1e12)
    t_fact = f_diag / (PEAK_TFLOPS["fp64"] * 1e12) + f_upd / (upd_peak * 1e12)
    t_ref = solves * (4.0 * n * n + 4.0 * n * n) / (HBM_TBS * 1e12)
    return {"convert_ms": round(t_conv * 1e3, 3), "factor_ms": round(t_fact * 1e3, 3),
            "refine_ms": round(t_ref * 1e3, 3), "total_ms": round((t_conv + t_fact + t_ref) * 1e3, 3)}

# HBM bytes of one tile-GEMM launch (nb=1024, ld=32768), from the
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes over `bench.py
# --traffic-probe`, each counter scaled by what it reads for a plain copy of
# a known size (FETCH_SIZE reads 4106.5 KiB for an 8192 KiB read: x1.995;
# WRITE_SIZE reads 8192 KiB for 8192 KiB written: x1.0).
TILE_GEMM_PMC = {"nb": 1024, "ld": 32768, "fetch_kib": 28725.5, "write_kib": 8192.0,
                 "fetch_scale": 8192.0 / 4106.5, "write_scale": 1.0}


def tile_gemm_traffic(nb: int, ld: int):
    t = TILE_GEMM_PMC
    if t["fetch_kib"] is None or (nb, ld) != (t["nb"], t["ld"]):
        return None
    return round((t["fetch_kib"] * t["fetch_scale"] + t["write_kib"] * t["write_scale"]) * 1024)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_spd(K, n: int, seed: int, shift: float = 1.0):
    """A = B B^T / n + shift*I on the device (B ~ N(0,1), counter-based, seeded)."""
    L = K.lib()
    A = K.DeviceArray(n, n)
    B = K.DeviceArray(n, n)
    K.check(L.sa_fill_random(None, n, n, B.ptr, n, seed, 1), "fill")
    K.dsyrk_ln(n, n, 1.0 / n, B.ptr, n, 0.0, A.ptr, n)
    K.check(L.sa_symmetrize_lower(None, n, A.ptr, n), "symmetrize")
    K.check(L.sa_add_diagonal(None, n, A.ptr, n, shift), "diag")
    K.sync()
    B.free()
    return A


def solve_backward_error(K, n, nrhs, A0, X, B0) -> float:
    """max over right-hand sides of ||b - A x|| / (||A||_F ||x|| + ||b||),
    computed on the device (A0 holds both triangles of A)."""
    L = K.lib()
    Rm = K.DeviceArray(n, nrhs)
    K.check(L.sa_memcpy_d2d(Rm.ptr, B0.ptr, n * nrhs * 8), "copy b")
    K.dgemm_nn(n, nrhs, n, -1.0, A0.ptr, n, X.ptr, n, 1.0, Rm.ptr, n)
    a = math.sqrt(float(np.sum(K.colsumsq(n, n, A0.ptr, n))))
    r = np.sqrt(K.colsumsq(n, nrhs, Rm.ptr, n))
    x = np.sqrt(K.colsumsq(n, nrhs, X.ptr, n))
    b = np.sqrt(K.colsumsq(n, nrhs, B0.ptr, n))
    Rm.free()
    return float(np.max(r / (a * x + b)))


def cpu_info() -> dict:
    """Host CPU model and the cores this process may use."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": len(os.sched_getaffinity(0))}


def cpu_baseline(A_lead: np.ndarray, b_lead: np.ndarray, nb: int, cores: int) -> dict:
    """The same factor-and-solve DAG on `cores` CPU workers under dmda, each
    running the reference's CPU codelets on BLAS/LAPACK (DPOTRF, DTRSM,
    DSYRK, DGEMM of scipy's bundled OpenBLAS, one thread per worker), on a
    leading principal submatrix and the leading rows of the right-hand
    sides.  One calibration run of the history models first, then a timed
    run."""
    import oracle
    from oracle.codelets import blas_addresses, blas_library
    from starpu_amd import runtime as R
    n, nrhs = b_lead.shape
    nblocks = max(1, n // nb)
    R.init(STARPU_NHIP=0, STARPU_NCPU=cores, STARPU_SCHED="dmda")
    R.set_cpu_kernels(**blas_addresses())
    runs = []
    for _ in range(2):
        A = np.asfortranarray(A_lead.copy())
        X = np.asfortranarray(b_lead.copy())
        st = R.factor_solve(A.ctypes.data, n, n, nblocks, X.ctypes.data, n, nrhs)
        runs.append(st)
    R.set_cpu_kernels()
    R.shutdown()
    # sanity: the CPU run is a correct factorisation and solve
    be = oracle.backward_error_llt(A_lead[:512, :512], np.tril(A[:512, :512]))
    r = oracle.residual_sym_l(A_lead, X[:, :1], b_lead[:, :1])
    sbe = np.linalg.norm(r) / (np.linalg.norm(A_lead) * np.linalg.norm(X[:, :1]) + np.linalg.norm(b_lead[:, :1]))
    st = runs[-1]
    assert st.info == 0 and be < 1e-13 and sbe < 1e-14, (st.info, be, sbe)
    sec = st.elapsed_us / 1e6
    return {"value": round(st.flops / sec / 1e12, 6), "unit": "TFLOP/s", "cores": cores,
            "kind": "port", "cpu": cpu_info(), "blas": os.path.basename(blas_library() or ""),
            "sample": f"leading {n}x{n} principal submatrix of the same A and the leading {n} rows of "
                      f"the same {nrhs} right-hand sides, tile {nb} ({nblocks}x{nblocks} tiles), "
                      f"factor+solve DAG on {cores} CPU workers (dmda) running the reference's CPU "
                      f"codelets' BLAS/LAPACK calls (scipy's OpenBLAS, 1 thread each); second of two "
                      f"runs: {sec:.2f} s"}


def c1_section(cores: int, seed: int, reps: int = 5) -> dict:
    """The reference's CPU configuration: N=2048, tile 256 (8x8 tiles),
    A = B B^T + N I with B uniform in [-1,1], one right-hand side uniform in
    [-1,1]; the factor+solve DAG under dmda on `cores` CPU workers running
    the BLAS codelets.  Three calibration runs, then the median of `reps`;
    the factor and the solution are checked against the oracle."""
    import oracle
    from oracle.codelets import blas_addresses
    from starpu_amd import runtime as R
    n, nb, nrhs = 2048, 256, 1
    rng = np.random.default_rng(seed)
    B = rng.uniform(-1, 1, (n, n))
    A0 = np.asfortranarray(B @ B.T + n * np.eye(n))
    b0 = np.asfortranarray(rng.uniform(-1, 1, (n, nrhs)))
    R.init(STARPU_NHIP=0, STARPU_NCPU=cores, STARPU_SCHED="dmda")
    R.set_cpu_kernels(**blas_addresses())
    times, fl = [], 0.0
    for _ in range(3 + reps):
        A, X = A0.copy(order="F"), b0.copy(order="F")
        st = R.factor_solve(A.ctypes.data, n, n, n // nb, X.ctypes.data, n, nrhs)
        assert st.info == 0
        times.append(st.elapsed_us / 1e3)
        fl = st.flops
    R.set_cpu_kernels()
    R.shutdown()
    Lref, _ = oracle.tiled_cholesky(A0, n // nb)
    rel = oracle.rel_diff_lower(np.tril(A), np.tril(Lref))
    Xref = oracle.tiled_potrs(np.tril(Lref), b0, n // nb)
    xrel = float(np.linalg.norm(X - Xref) / np.linalg.norm(Xref))
    assert rel < 1e-12 and xrel < 1e-12, (rel, xrel)
    ms = float(np.median(times[3:]))
    return {"N": n, "tile": nb, "nrhs": nrhs, "sched": "dmda", "cpu_workers": cores, "cpu": cpu_info(),
            "codelets": "reference CPU codelets on BLAS/LAPACK (scipy OpenBLAS, 1 thread per worker)",
            "ms": round(ms, 3), "gflops": round(fl / ms / 1e6, 2), "runs_ms": [round(t, 2) for t in times[3:]],
            "factor_rel_diff_vs_oracle": float(f"{rel:.3e}"), "solution_rel_diff_vs_oracle": float(f"{xrel:.3e}")}


def host_to_host(K, R, A0, B0, n: int, nb: int, nrhs: int, prio: str, reps: int = 2) -> dict:
    """The C2 factor+solve with A and B in pinned main memory (the
    reference's setting, cholesky_implicit.c:423): tiles reach HBM through
    the runtime's MSI transfers as tasks need them, and the timer covers
    those uploads, from the first submission to wait_for_all
    (cholesky_implicit.c:233,325).  The gather of the factor and X back to
    main memory at unpartition is timed separately (the reference leaves it
    out, :456)."""
    import ctypes as C
    L = K.lib()
    rt = R.rt()
    rt.starpu_malloc_flags.restype = C.c_int
    rt.starpu_malloc_flags.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_int]
    rt.starpu_free_flags.restype = C.c_int
    rt.starpu_free_flags.argtypes = [C.c_void_p, C.c_size_t, C.c_int]
    pa, pb = C.c_void_p(), C.c_void_p()
    abytes, bbytes = n * n * 8, n * nrhs * 8
    if rt.starpu_malloc_flags(C.byref(pa), abytes, 2) or rt.starpu_malloc_flags(C.byref(pb), bbytes, 2):
        return {"skipped": "pinned host allocation failed"}
    out = []
    try:
        for _ in range(reps + 1):
            K.check(L.sa_memcpy_d2h(pa.value, A0.ptr, abytes), "d2h")
            K.check(L.sa_memcpy_d2h(pb.value, B0.ptr, bbytes), "d2h")
            t0 = time.perf_counter()
            st = R.factor_solve(pa.value, n, n, n // nb, pb.value, n, nrhs, home_node=0, prio=prio)
            wall = (time.perf_counter() - t0) * 1e3
            assert st.info == 0
            out.append((st.elapsed_us / 1e3, wall, st.flops))
    finally:
        rt.starpu_free_flags(pa, abytes, 2)
        rt.starpu_free_flags(pb, bbytes, 2)
    ms = float(np.median([o[0] for o in out[1:]]))
    wall = float(np.median([o[1] for o in out[1:]]))
    return {"tflops": round(out[-1][2] / ms / 1e9, 3), "ms": round(ms, 2),
            "ms_with_gather": round(wall, 2),
            "note": "A and B in pinned main memory; timed from the first submission to wait_for_all, "
                    "tile uploads included (the reference's timing); ms_with_gather adds the write-back "
                    "of the factor's tiles and X at unpartition"}


def mixed_section(K, R, n: int, nb: int, nrhs: int, seed: int, precs: list[str], reps: int,
                  mixed_nb: int = 0, mixed_prio: str = "parsec", runtime_for=None) -> dict:
    """Mixed precision + iterative refinement on the ill-conditioned matrix
    A = B B^T/N + 1e-3 I (cond ~4e3), 16 right-hand sides: time to
    solution (FP32 copy, factorisation DAG, refinement to backward error
    1e-14) per update precision, next to the FP64 factor+solve on the same
    A.  The runtime must be initialised with HIP workers; the FP64 run uses
    it as it is, and `runtime_for("mixed")` / `runtime_for("fp64")`, when
    given, switch it to the mixed solves' configuration and back."""
    L = K.lib()
    A3 = make_spd(K, n, seed, shift=1e-3)
    B0 = K.DeviceArray(n, nrhs)
    K.check(L.sa_fill_random(None, n, nrhs, B0.ptr, n, seed + 1, 1), "fill b")
    X = K.DeviceArray(n, nrhs)
    node = R.hip_node(0)
    out = {"matrix": "A = B*B^T/N + 1e-3*I (cond ~4e3), B i.i.d. N(0,1)", "nrhs": nrhs, "tol": 1e-14}
    # FP64 time to solution on the same matrix
    Aw = K.DeviceArray(n, n)
    f64 = []
    for _ in range(reps + 1):
        K.check(L.sa_memcpy_d2d(Aw.ptr, A3.ptr, n * n * 8), "copy")
        K.check(L.sa_memcpy_d2d(X.ptr, B0.ptr, n * nrhs * 8), "copy")
        K.sync()
        st = R.factor_solve(Aw.ptr, n, n, n // nb, X.ptr, n, nrhs, home_node=node)
        K.sync()
        f64.append(st.elapsed_us / 1e3)
    Aw.free()
    t64 = float(np.median(f64[1:] if len(f64) > 1 else f64))
    out["fp64_tts_ms"] = round(t64, 2)
    out["fp64_backward_error"] = float(f"{solve_backward_error(K, n, nrhs, A3, X, B0):.3e}")
    x64 = X.download().copy()          # the FP64 solution, for the forward differences below
    log(f"mixed: FP64 factor+solve on the cond~4e3 matrix {t64:.1f} ms")
    if runtime_for:
        out["runtime"] = runtime_for("mixed")
        node = R.hip_node(0)
    for prec in precs:
        runs = []
        for _ in range(reps + 1):
            st = R.mixed_solve(A3.ptr, n, n, n // (mixed_nb or nb), B0.ptr, n, nrhs, X.ptr, n, prec=prec, tol=1e-14,
                               maxiter=20 if prec == "bf16" else 40, hip_node=node, prio=mixed_prio)
            runs.append(st)
        best = runs[1:] if len(runs) > 1 else runs
        tts = float(np.median([r.elapsed_us for r in best])) / 1e3
        r = best[-1]
        berr = solve_backward_error(K, n, nrhs, A3, X, B0)
        fwd = float(np.linalg.norm(X.download() - x64) / np.linalg.norm(x64))
        roof = mixed_roofline(n, mixed_nb or nb, prec, int(r.iterations))
        out[prec] = {"tts_ms": round(tts, 2), "speedup_vs_fp64": round(t64 / tts, 3) if r.converged else None,
                     "rel_diff_vs_fp64_solution": float(f"{fwd:.3e}"),
                     "roofline": dict(roof, frac=round(roof["total_ms"] / tts, 4),
                                      model=mixed_roofline.__doc__.split("\n\n")[0].replace("\n    ", " ")),
                     "converged": bool(r.converged), "refinement_solves": int(r.iterations),
                     "final_backward_error": float(f"{r.final_berr:.3e}"),
                     "backward_error_recomputed": float(f"{berr:.3e}"),
                     "convert_ms": round(r.convert_us / 1e3, 2), "factor_ms": round(r.factor_us / 1e3, 2),
                     "refine_ms": round(r.refine_us / 1e3, 2),
                     "factor_tflops_equiv": round(r.flops_factor / r.factor_us / 1e6, 2)}
        log(f"mixed {prec}: {tts:.1f} ms ({r.iterations} solves, berr {r.final_berr:.2e}, "
            f"factor {r.factor_us / 1e3:.1f} ms, refine {r.refine_us / 1e3:.1f} ms)")
    if runtime_for:
        runtime_for("fp64")
    A3.free()
    return out


def serving_section(K, R, nreq: int, seed: int, slots: int, nb: int, prec: str, util: float,
                    tight_gb: float) -> dict:
    """The solve server on a trace of SPD solve requests: N log-uniform in
    [256, 16384] rounded to 256, nrhs uniform in 1..32, 30 % re-solves,
    A = B B^T/N + 0.01 I, FP32 factor + FP64 refinement to 1e-14, 200 GB
    HBM budget.  Four runs over the same requests:
      1. one slot, all requests queued at once: the results each request
         gets alone (the reference of the alone-versus-in-the-mix check);
      2. `slots` slots, all queued at once: the server's capacity
         (requests per second at saturation);
      3. Poisson arrivals at `util` times that capacity: latency p50/p99;
      4. the same arrivals under a `tight_gb` budget that binds (requests
         wait for memory, resident factors are evicted).
    The distinct matrices of one size share their device array (at most 64
    sizes), so a long trace's inputs fit in HBM next to the budget."""
    from starpu_amd import serving as S
    tr0 = S.make_trace(nreq, 0.0, seed)
    inp = S.Inputs(tr0, shift=0.01, seed=seed, share_by_size=True)

    def clear():
        for x in inp.X:
            K.check(K.lib().sa_dlaset(None, x.m, x.n, 0.0, 0.0, x.ptr, x.ld), "clear X")
        K.sync()

    reqs = inp.requests(tr0, alone=True)
    st1 = R.serve(S.config(budget_gb=200, slots=1, nb=nb, prec=prec), reqs)
    mean_service_ms = st1.wall_us / 1e3 / nreq
    alone = [x.download() for x in inp.X]
    log(f"serving: {nreq} requests alone in {st1.wall_us / 1e6:.1f} s")
    clear()
    reqs = inp.requests(tr0, alone=True)
    st2 = R.serve(S.config(budget_gb=200, slots=slots, nb=nb, prec=prec), reqs)
    capacity = nreq / (st2.wall_us / 1e6)
    log(f"serving: capacity {capacity:.1f} requests/s with {slots} slots")
    qps = util * capacity
    tr = S.make_trace(nreq, qps, seed)           # same requests, Poisson arrivals at qps
    assert tr.n == tr0.n and tr.matrix == tr0.matrix
    clear()
    reqs = inp.requests(tr)
    st = R.serve(S.config(budget_gb=200, slots=slots, nb=nb, prec=prec), reqs)
    out = S.summarise(reqs, st, tr)
    out["factor_roofline"] = S.factor_share(reqs, prec)
    diff = max(float(np.linalg.norm(x.download() - a) / max(np.linalg.norm(a), 1e-300))
               for x, a in zip(inp.X, alone))
    out.update({"qps": round(qps, 2), "utilization_of_capacity": util,
                "capacity_rps": round(capacity, 2), "mean_service_alone_ms": round(mean_service_ms, 3),
                "slots": slots, "tile": nb, "prec": prec, "budget_gb": 200,
                "trace": f"{nreq} requests, N log-uniform [256,16384] /256, nrhs U{{1..32}}, 30% re-solves, "
                         f"A = B*B^T/N + 0.01*I, {len(tr.matrix_n)} distinct matrices "
                         f"({len(set(tr.matrix_n))} sizes sharing device inputs)",
                "alone_vs_mix_max_rel_diff": float(f"{diff:.3e}")})
    log(f"serving: qps {qps:.1f}, p50 {out['p50_ms']} ms, p99 {out['p99_ms']} ms, "
        f"{out['solves_per_s']} solves/s, max berr {out['max_backward_error']}, alone-vs-mix {diff:.1e}, "
        f"factor share {out['factor_roofline']['achieved']} TF/s")
    if tight_gb > 0:
        clear()
        reqs = inp.requests(tr)
        stt = R.serve(S.config(budget_gb=tight_gb, slots=slots, nb=nb, prec=prec), reqs)
        tight = S.summarise(reqs, stt, tr)
        tdiff = max(float(np.linalg.norm(x.download() - a) / max(np.linalg.norm(a), 1e-300))
                    for x, a in zip(inp.X, alone))
        tight.update({"budget_gb": tight_gb, "qps": round(qps, 2), "alone_vs_mix_max_rel_diff": float(f"{tdiff:.3e}")})
        out["tight_budget"] = tight
        log(f"serving under {tight_gb} GB: p50 {tight['p50_ms']} ms, p99 {tight['p99_ms']} ms, evictions "
            f"{tight['evictions']}, memory waits {tight['memory_waits']}, peak {tight['peak_gb']} GB")
    inp.free()
    return out


def batched_section(K, seed: int, nsys: int, reps: int) -> dict:
    """Batched uneven small systems: nsys SPD systems, N uniform in
    {32, 64, ..., 1024}, A = B B^T/N + 0.1 I, one right-hand side each,
    factored and solved by the persistent kernels, in FP64 and as FP32
    factor + FP64 refinement; launch shapes from the autotuner; each launch
    captured as a hipGraph and replayed (inputs restored between replays,
    outside the timed region)."""
    import ctypes as C
    import oracle
    from starpu_amd import autotune, batched
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, 33, nsys) * 32
    b = batched.Batch(sizes, nrhs=1, shift=0.1, seed=seed, ir=True)
    L = K.lib()
    s = C.c_void_p()
    K.check(L.sa_stream_create(C.byref(s)), "stream")
    BL = batched._lib()

    def timed(mode):
        tune = autotune.tune(b, stream=s, reps=2, use_cache=False, mode=mode)
        exe = b.capture(tune["grid"], tune["order"], s, ir=(mode == "ir"), rows=tune["rows"])
        times = []
        for _ in range(reps + 1):
            b.restore()
            K.sync()
            t = C.c_void_p()
            K.check(L.sa_timer_start(s, C.byref(t)), "timer")
            K.check(BL.sa_graph_launch(exe, s), "graph launch")
            ms = C.c_float()
            K.check(L.sa_timer_stop(t, C.byref(ms)), "timer")
            times.append(ms.value)
        K.check(BL.sa_graph_destroy(exe), "graph destroy")
        return float(np.median(times[1:])), tune

    ir_ms, ir_tune = timed("ir")
    ir_infos = b.infos()
    it_all = b.iters.download()[:nsys, 0].copy()
    be_all = b.berr.download()[:nsys, 0].copy()
    t_ms, tune = timed("fp64")
    infos = b.infos()
    # parity on a sample: the largest, the smallest and random systems
    idx = sorted(set([int(np.argmax(sizes)), int(np.argmin(sizes))] + list(rng.integers(0, nsys, 6))))
    worst_rel, worst_berr = 0.0, 0.0
    ir_rel = 0.0
    for i in idx:
        A, bb, x, Lf = b.system(i)
        Lref, _ = oracle.dpotrf_l(A)
        worst_rel = max(worst_rel, oracle.rel_diff_lower(Lf, np.tril(Lref)))
        r = bb - A @ x
        worst_berr = max(worst_berr, float(np.linalg.norm(r) / (np.linalg.norm(A) * np.linalg.norm(x)
                                                              + np.linalg.norm(bb))))
        xr = oracle.dpotrs_l(np.tril(Lref), bb)
        xi = b.ir_result(i)[0]
        ir_rel = max(ir_rel, float(np.linalg.norm(xi - xr) / np.linalg.norm(xr)))
    # CPU baseline: the oracle's per-system factor + solve, the systems spread
    # over the host cores (one system per thread at a time; the oracle's
    # ctypes calls release the GIL), on a bounded sample of the same systems:
    # the matrices are copied to the host first, then each thread takes the
    # next system until about 10 s have passed
    from concurrent.futures import ThreadPoolExecutor
    cores = min(16, len(os.sched_getaffinity(0)))
    sample = [(int(i), b.system(int(i))[:2]) for i in rng.permutation(nsys)]
    t_limit = 10.0
    t0 = time.perf_counter()

    def cpu_one(item):
        i, (A, bb) = item
        if time.perf_counter() - t0 > t_limit:
            return None
        Lc, _ = oracle.dpotrf_l(A)
        oracle.dpotrs_l(np.tril(Lc), bb)
        v = float(sizes[i])
        return v ** 3 / 3 + 2 * v * v

    with ThreadPoolExecutor(max_workers=cores) as pool:
        done = [f for f in pool.map(cpu_one, sample) if f is not None]
    cpu_s = time.perf_counter() - t0
    cpu_flops, ncpu = float(sum(done)), len(done)
    del sample
    flops, nbytes = b.flops(), b.bytes()
    out = {"systems": nsys, "sizes": "N uniform in {32,64,...,1024}, A = B*B^T/N + 0.1*I, 1 RHS",
           "batch_ms": round(t_ms, 3), "systems_per_s": round(nsys / (t_ms / 1e3), 1),
           "tflops": round(flops / (t_ms / 1e3) / 1e12, 3),
           "fp64_peak_frac": round(flops / (t_ms / 1e3) / 1e12 / FP64_PEAK_TFLOPS, 4),
           "algorithmic_gbps": round(nbytes / (t_ms / 1e3) / 1e9, 1),
           "all_spd": bool((infos == 0).all()), "sample_factor_rel_diff": float(f"{worst_rel:.3e}"),
           "sample_backward_error": float(f"{worst_berr:.3e}"),
           "autotune": {"grid": tune["grid"], "order": tune["order"], "rows": tune["rows"],
                        "candidates": len(tune["table"])},
           "graph_replays": reps, "dtype": "f64",
           "fp32_ir": {"batch_ms": round(ir_ms, 3), "systems_per_s": round(nsys / (ir_ms / 1e3), 1),
                       "speedup_vs_fp64": round(t_ms / ir_ms, 3),
                       "fp64_equiv_tflops": round(flops / (ir_ms / 1e3) / 1e12, 3),
                       "all_converged": bool((ir_infos == 0).all() and (be_all <= 1e-14).all()),
                       "tol": 1e-14, "max_backward_error": float(f"{be_all.max():.3e}"),
                       "iterations_max": int(it_all.max()), "iterations_mean": round(float(it_all.mean()), 3),
                       "sample_rel_diff_vs_fp64_solve": float(f"{ir_rel:.3e}"),
                       "autotune": {"grid": ir_tune["grid"], "order": ir_tune["order"],
                                    "candidates": len(ir_tune["table"])},
                       "dtype": "f32 factor, f64 refinement"},
           "cpu_baseline": {"systems_per_s": round(ncpu / cpu_s, 2), "tflops": round(cpu_flops / cpu_s / 1e12, 6),
                            "cores": cores, "kind": "port",
                            "sample": f"{ncpu} random systems of the batch, oracle dpotrf + dpotrs, one system per "
                                      f"thread on {cores} threads, {cpu_s:.1f} s"}}
    log(f"batched: {nsys} systems in {t_ms:.2f} ms ({out['systems_per_s']} systems/s, {out['tflops']} TF/s), "
        f"tuned grid {tune['grid']} {tune['order']} rows {tune['rows']}; FP32+IR {ir_ms:.2f} ms ({t_ms / ir_ms:.2f}x), "
        f"iterations max {int(it_all.max())}, max berr {be_all.max():.1e}, grid {ir_tune['grid']}")
    K.check(L.sa_stream_destroy(s), "stream")
    b.free()
    return out


def gemm_probe(K, n: int, reps: int = 3) -> None:
    """Launches for rocprofv3 counter passes over the FP64 GEMM kernel:
    `reps` products C = A A^T of one n x n matrix."""
    L = K.lib()
    A = K.DeviceArray(n, n)
    K.check(L.sa_fill_random(None, n, n, A.ptr, n, 2, 1), "fill")
    Cm = K.DeviceArray(n, n)
    for _ in range(reps):
        K.dgemm_nt(n, n, n, 1.0, A.ptr, n, A.ptr, n, 0.0, Cm.ptr, n)
    K.sync()
    print(json.dumps({"gemm_probe": {"n": n, "launches": reps}}))
    A.free()
    Cm.free()


def batched_probe(K, seed: int, nsys: int, grid: int) -> None:
    """Launches for counter passes over the batched kernels: the batch of
    the benchmark (nsys systems, N uniform in {32, ..., 1024}), one FP64
    factor+solve launch and one FP32-factor + FP64-refinement launch, each
    after the inputs are restored."""
    from starpu_amd import batched
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, 33, nsys) * 32
    log(f"batched probe: building {nsys} systems")
    b = batched.Batch(sizes, nrhs=1, shift=0.1, seed=seed, ir=True)
    log("batched probe: launching")
    b.restore()
    K.sync()
    b.launch(grid)
    K.sync()
    b.restore()
    K.sync()
    b.launch_ir(grid)
    K.sync()
    print(json.dumps({"batched_probe": {"systems": nsys, "grid": grid, "flops": b.flops(), "bytes": b.bytes(),
                                        "launches": ["k_batched_potrs", "k_batched_posv_ir"]}}), flush=True)
    b.free()


def traffic_probe(K, n: int, nb: int, reps: int = 10) -> None:
    """Launches for HBM byte counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE):
    `reps` tile GEMMs exactly as the factorisation's GEMM task issues them
    (nb x nb tiles, K = nb, leading dimension n), then `reps` plain copies of
    one nb x nb tile (nb*nb*8 bytes read and written) that scale the counters."""
    L = K.lib()
    assert n >= 4 * nb
    big = K.DeviceArray(n, 3 * nb)            # three tile columns, ld = n
    K.check(L.sa_fill_random(None, n, 3 * nb, big.ptr, n, 7, 0), "fill")
    A, B, Cm = big.at(0, 0), big.at(nb, nb), big.at(2 * nb, 2 * nb)
    K.sync()
    for _ in range(reps):
        K.dgemm_nt(nb, nb, nb, -1.0, A, n, B, n, 1.0, Cm, n)
    K.sync()
    for _ in range(reps):
        K.check(L.sa_dlacpy(None, nb, nb, A, n, big.at(3 * nb, 0), n), "copy")
    K.sync()
    print(json.dumps({"traffic_probe": {"gemm_launches": reps, "copy_launches": reps,
                                        "tile": nb, "ld": n, "copy_bytes_each_way": nb * nb * 8}}),
          flush=True)


def resid_probe(K, n: int, nrhs: int, reps: int = 10) -> None:
    """The refinement's FP64 residual R = B - A X at n x n, nrhs right-hand
    sides, two ways, `reps` launches each, HIP-event timed: the general thin
    product over both triangles (sa_dlacpy + sa_dgemm_nn) and the product
    from the stored lower triangle (sa_dsymm_lower_resid).  For kernel
    traces and counter passes."""
    import ctypes as C
    L = K.lib()
    A = K.DeviceArray(n, n)
    X, B, R1, R2 = (K.DeviceArray(n, nrhs) for _ in range(4))
    K.check(L.sa_fill_random(None, n, n, A.ptr, n, 3, 1), "fill A")
    K.check(L.sa_symmetrize_lower(None, n, A.ptr, n), "symmetrize")
    K.check(L.sa_fill_random(None, n, nrhs, X.ptr, n, 4, 1), "fill X")
    K.check(L.sa_fill_random(None, n, nrhs, B.ptr, n, 5, 1), "fill B")
    ws = max(L.sa_dsymm_lower_resid_worksize(n, nrhs), L.sa_dgemm_thin_worksize(n, nrhs, n))
    W = K.DeviceArray(ws // 8 + 1, 1)
    K.sync()

    def general():
        K.check(L.sa_dlacpy(None, n, nrhs, B.ptr, n, R1.ptr, n), "copy")
        K.check(L.sa_dgemm_nn(None, n, nrhs, n, -1.0, A.ptr, n, X.ptr, n, 1.0, R1.ptr, n, W.ptr, W.nbytes), "gemm")

    def lower():
        K.check(L.sa_dsymm_lower_resid(None, n, nrhs, A.ptr, n, X.ptr, n, B.ptr, n, R2.ptr, n, W.ptr, W.nbytes),
                "symm")

    out = {}
    for name, fn in (("general", general), ("lower_triangle", lower)):
        ts = []
        for _ in range(reps):
            t = C.c_void_p()
            K.check(L.sa_timer_start(None, C.byref(t)), "timer")
            fn()
            ms = C.c_float()
            K.check(L.sa_timer_stop(t, C.byref(ms)), "timer")
            ts.append(ms.value)
        out[name + "_ms"] = round(float(np.median(ts[2:])), 4)
    K.sync()
    out["max_abs_diff"] = float(np.abs(R1.download() - R2.download()).max())
    print(json.dumps({"resid_probe": dict(out, n=n, nrhs=nrhs, launches_each=reps)}), flush=True)
    for d in (A, X, B, R1, R2, W):
        d.free()


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--nb", type=int, default=1024)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--workers-per-gpu", type=int, default=int(os.environ.get("STARPU_NWORKER_PER_HIP", "12")))
    ap.add_argument("--cpu-n", type=int, default=16384, help="CPU-baseline sample size (0: skip)")
    ap.add_argument("--cpu-cores", type=int, default=0)
    ap.add_argument("--c1", type=int, default=1, help="time the reference's CPU configuration (N=2048, tile 256)")
    ap.add_argument("--host-to-host", type=int, default=2,
                    help="timed runs of the factor+solve with A and B in main memory (0: skip)")
    ap.add_argument("--trace", default="", help="write a Chrome trace of the last step here")
    ap.add_argument("--prio", default="timings", choices=["none", "base", "kut", "timings", "parsec"],
                    help="task priority attribution of the factorisation (the reference's -p)")
    ap.add_argument("--mixed", default="bf16x3,fp32,bf16",
                    help="update precisions of the mixed-precision + refinement runs ('' to skip)")
    ap.add_argument("--mixed-reps", type=int, default=3)
    ap.add_argument("--mixed-workers", type=int, default=16, help="HIP streams of the mixed solves' runtime")
    ap.add_argument("--mixed-chain-stream", type=int, default=1,
                    help="keep one stream of the mixed solves for the factorisation's critical chain (0: off)")
    ap.add_argument("--mixed-nb", type=int, default=0, help="tile of the mixed-precision factor (0: --nb)")
    ap.add_argument("--mixed-prio", default="parsec", choices=["none", "base", "kut", "parsec"],
                    help="task priorities of the mixed-precision factor")
    ap.add_argument("--serve-requests", type=int, default=1000, help="solve-server trace length (0: skip)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-prec", default="fp32")
    ap.add_argument("--serve-util", type=float, default=0.7, help="arrival rate / measured capacity")
    ap.add_argument("--serve-tight-gb", type=float, default=2.5,
                    help="budget of the binding-budget run (0: skip): about twice the largest request's footprint")
    ap.add_argument("--batched-systems", type=int, default=2048, help="batched small systems (0: skip)")
    ap.add_argument("--gemm-probe", type=int, default=0,
                    help="launch three N^3 FP64 GEMMs (C = A A^T) for counter passes, then exit")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only launch the tile GEMM and a plain tile copy (for HBM byte counters)")
    ap.add_argument("--resid-probe", type=int, default=0,
                    help="only time the refinement residual at this N, both forms (kernel traces, counters)")
    ap.add_argument("--batched-probe", type=int, default=0,
                    help="only launch the batched kernels once each, with this many workgroups (counter passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the tile DAG runs on "
              f"one GPU (sharding the tile grid over several GPUs is not implemented)", file=sys.stderr)
        return 2
    if os.environ.get("WORLD_SIZE", "1") not in ("", "1"):
        print("bench.py runs as a single process on one GPU", file=sys.stderr)
        return 2

    from starpu_amd import kernels as K
    from starpu_amd import runtime as R
    if K.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    n, nb, nrhs = args.n, args.nb, args.nrhs
    if args.traffic_probe:
        traffic_probe(K, n, nb)
        return 0
    if args.resid_probe > 0:
        resid_probe(K, args.resid_probe, args.nrhs)
        return 0
    if args.gemm_probe > 0:
        gemm_probe(K, args.gemm_probe)
        return 0
    if args.batched_probe > 0:
        batched_probe(K, args.seed, args.batched_systems, args.batched_probe)
        return 0
    assert n % nb == 0
    nblocks = n // nb
    L = K.lib()
    log(f"generating A (N={n}) and {nrhs} right-hand sides on the GPU")
    t0 = time.perf_counter()
    A0 = make_spd(K, n, args.seed)
    B0 = K.DeviceArray(n, nrhs)
    K.check(L.sa_fill_random(None, n, nrhs, B0.ptr, n, args.seed + 1, 1), "fill b")
    Aw = K.DeviceArray(n, n)
    Xw = K.DeviceArray(n, nrhs)
    K.sync()
    log(f"generated in {time.perf_counter() - t0:.2f} s")

    pmdir = os.environ.get("STARPU_PERF_MODEL_DIR") or os.path.join(tempfile.gettempdir(), "starpu_amd_perfmodels")
    R.init(STARPU_NHIP=1, STARPU_NCPU=0, STARPU_SCHED="dmda", STARPU_PERF_MODEL_DIR=pmdir,
           STARPU_NWORKER_PER_HIP=args.workers_per_gpu)
    node = R.hip_node(0)
    flops = R.chol().chol_flops_potrf(float(n)) + R.chol().chol_flops_potrs(float(n), float(nrhs))

    def step():
        K.check(L.sa_memcpy_d2d(Aw.ptr, A0.ptr, n * n * 8), "restore A")
        K.check(L.sa_memcpy_d2d(Xw.ptr, B0.ptr, n * nrhs * 8), "restore B")
        K.sync()
        st = R.factor_solve(Aw.ptr, n, n, nblocks, Xw.ptr, n, nrhs, home_node=node, prio=args.prio)
        K.sync()
        if st.info != 0:
            raise RuntimeError(f"POTRF info {st.info}")
        return st

    for w in range(args.warmup):
        st = step()
        log(f"warmup {w}: {st.elapsed_us / 1e3:.1f} ms  {flops / st.elapsed_us / 1e6:.2f} TFLOP/s")
    times = []
    kinds = np.zeros(7)
    kn = np.zeros(7)
    for s in range(args.steps):
        st = step()
        times.append(st.elapsed_us / 1e6)
        kinds += np.array(st.kernel_us[:])
        kn += np.array(st.kernel_n[:])
        log(f"step {s}: {st.elapsed_us / 1e3:.1f} ms  {flops / st.elapsed_us / 1e6:.2f} TFLOP/s")
    if args.trace:
        R.rt().starpu_trace_dump_chrome(args.trace.encode())
    h2h = None
    if args.host_to_host > 0:
        h2h = host_to_host(K, R, A0, B0, n, nb, nrhs, args.prio, args.host_to_host)
        log(f"host-to-host: {h2h}")
    mixed = None
    if args.mixed:
        def runtime_for(kind):
            """The FP64 configuration (the headline's streams), or the mixed
            solves' one: more streams, with one of them kept for the
            factorisation's critical chain (POTRF, the TRSM and SYRK that
            feed the next one) so that its tasks never queue behind updates."""
            R.shutdown()
            mixed_cfg = kind == "mixed"
            w = args.mixed_workers if mixed_cfg else args.workers_per_gpu
            res = -1 if mixed_cfg and args.mixed_chain_stream else 0
            R.init(STARPU_NHIP=1, STARPU_NCPU=0, STARPU_SCHED="dmda", STARPU_PERF_MODEL_DIR=pmdir,
                   STARPU_NWORKER_PER_HIP=w, STARPU_HIP_RESERVE_CUS=res)
            return {"hip_workers": w, "chain_stream": bool(res)}
        cfg_differs = args.mixed_workers != args.workers_per_gpu or args.mixed_chain_stream
        mixed = mixed_section(K, R, n, nb, nrhs, args.seed, [p for p in args.mixed.split(",") if p],
                              args.mixed_reps, args.mixed_nb, args.mixed_prio,
                              runtime_for if cfg_differs else None)
    serving = None
    if args.serve_requests > 0:
        serving = serving_section(K, R, args.serve_requests, args.seed, args.serve_slots, 512,
                                  args.serve_prec, args.serve_util, args.serve_tight_gb)
    R.shutdown()
    batched = batched_section(K, args.seed, args.batched_systems, 5) if args.batched_systems > 0 else None

    # parity: (1) the leading block of the tiled GPU factor against the CPU
    # oracle (L11 of the factor is the Cholesky factor of A11); (2) the
    # normwise backward error of the full solve, computed on the device.
    import oracle
    m = min(2048, n)
    host = np.empty((n, m), order="F")   # first m columns
    K.check(L.sa_memcpy_d2h(host.ctypes.data, Aw.ptr, n * m * 8), "d2h")
    L11 = np.asfortranarray(host[:m, :m])
    K.check(L.sa_memcpy_d2h(host.ctypes.data, A0.ptr, n * m * 8), "d2h")
    lead = np.asfortranarray(host[:m, :m])
    ref, _ = oracle.tiled_cholesky(lead, max(1, m // nb))
    rel = oracle.rel_diff_lower(np.tril(L11), np.tril(ref))
    berr = solve_backward_error(K, n, nrhs, A0, Xw, B0)
    log(f"leading {m}x{m} factor vs oracle: rel diff {rel:.2e}; solve backward error {berr:.2e}")

    mean_t = float(np.mean(times))
    value = flops / mean_t / 1e12
    gemm_us, gemm_n = float(kinds[3]), int(kn[3])
    gemm_avg_us = gemm_us / max(1, gemm_n)
    gemm_flops = 2.0 * nb ** 3
    achieved = gemm_flops / (gemm_avg_us * 1e-6) / 1e12 if gemm_n else None
    steps_total_us = float(np.sum(times)) * 1e6
    gemm_share = float(kinds[3] / max(kinds.sum(), 1e-9))
    chip_gemm = gemm_flops * gemm_n / (steps_total_us * gemm_share) / 1e6 if gemm_n else 0.0

    cores = args.cpu_cores or min(16, len(os.sched_getaffinity(0)))
    c1 = None
    if args.c1:
        c1 = c1_section(cores, args.seed)
        log(f"C1 (N=2048, tile 256, {cores} CPU workers): {c1['ms']} ms, {c1['gflops']} GFLOP/s")
    cpu = None
    if args.cpu_n > 0:
        cn = min(args.cpu_n, n)
        host = np.empty((n, cn), order="F")
        K.check(L.sa_memcpy_d2h(host.ctypes.data, A0.ptr, n * cn * 8), "d2h")
        lead = np.asfortranarray(host[:cn, :cn])
        del host
        bh = np.empty((n, nrhs), order="F")
        K.check(L.sa_memcpy_d2h(bh.ctypes.data, B0.ptr, n * nrhs * 8), "d2h")
        log(f"CPU baseline: N={cn} on {cores} cores")
        cpu = cpu_baseline(lead, np.asfortranarray(bh[:cn]), nb, cores)
        log(f"CPU baseline: {cpu['value']:.4f} TFLOP/s")

    names = ("potrf", "trsm", "syrk", "gemm", "solve_fwd", "solve_bwd", "trtri")
    out = {
        "metric": "Cholesky TFLOP/s N=32768 FP64 vs mixed+IR at FP64 accuracy; p99 solve latency",
        "value": round(value, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mean_t * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: A = B*B^T/N + I, B i.i.d. N(0,1); right-hand sides N(0,1) "
                "(seeded, generated on the GPU)",
        "config": {"workload": "tiled FP64 Cholesky factor+solve (POTRF/TRSM/SYRK/GEMM DAG then "
                               "forward/backward tile solves, one task graph, dmda, HIP codelets), "
                               "device-resident A and B",
                   "N": n, "tile": nb, "tiles": f"{nblocks}x{nblocks}", "nrhs": nrhs,
                   "tasks": int(kn.sum()), "hip_workers": args.workers_per_gpu, "sched": "dmda", "priorities": args.prio},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3) if achieved else None,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4) if achieved else None,
                     "traffic": tile_gemm_traffic(nb, n),
                     "kernel": f"k_dgemm_nt (tile GEMM, 2*{nb}^3 flop per launch)",
                     "avg_launch_us": round(gemm_avg_us, 2),
                     # launches overlap on the runtime's streams: the average
                     # number in flight, and the chip-level GEMM rate (GEMM
                     # flops over the GEMM share of the summed kernel time)
                     "launches_in_flight": round(gemm_us / steps_total_us, 2) if gemm_n else None,
                     "chip_achieved": round(chip_gemm, 3) if gemm_n else None,
                     "chip_frac": round(chip_gemm / FP64_PEAK_TFLOPS, 4) if gemm_n else None},
        "kernel_time_share": {k: round(float(v / max(kinds.sum(), 1e-9)), 4) for k, v in zip(names, kinds)},
        "parity": {"leading_block_rel_diff_vs_oracle": float(f"{rel:.3e}"),
                   "solve_backward_error": float(f"{berr:.3e}")},
        "cpu_baseline": cpu,
        "host_to_host": h2h,
        "c1_cpu": c1,
        "mixed_precision": mixed,
        "serving": serving,
        "batched": batched,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
