"""Host side of the batched small-system path: many independent SPD systems
of uneven sizes packed in HBM, factored and solved by the persistent kernels
of sa_batched.hip, optionally replayed as a hipGraph:
  sa_dpotrs_batched    FP64 factor and solve, in place;
  sa_dposv_batched_ir  FP32 factor and solves with FP64 iterative refinement
                       (the scheme of LAPACK dsposv), A and b left intact.

Layout: the systems' matrices are packed one after the other, each
column-major with ld = n_i; likewise the right-hand sides and the block
inverses.  The queue order lists the systems largest first.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import kernels as K


def _lib():
    L = K.lib()
    if not hasattr(L, "_batched_sigs"):
        vp, i, sz = C.c_void_p, C.c_int, C.c_size_t
        L.sa_dpotrs_batched.restype = i
        L.sa_dpotrs_batched.argtypes = [vp, i, vp, vp, vp, vp, vp, vp, vp, i, vp, vp, vp, i]
        L.sa_set_batched_rows.restype = i
        L.sa_set_batched_rows.argtypes = [i]
        L.sa_batched_dinv_elems.restype = sz
        L.sa_batched_dinv_elems.argtypes = [i]
        L.sa_dposv_batched_ir.restype = i
        L.sa_dposv_batched_ir.argtypes = [vp, i, vp, vp, vp, vp, vp, vp, vp, vp, i, vp, vp, vp, vp, vp, vp, vp,
                                          i, C.c_double, vp, i]
        L.sa_graph_begin.restype = i
        L.sa_graph_begin.argtypes = [vp]
        L.sa_graph_end.restype = i
        L.sa_graph_end.argtypes = [vp, C.POINTER(vp)]
        L.sa_graph_launch.restype = i
        L.sa_graph_launch.argtypes = [vp, vp]
        L.sa_graph_destroy.restype = i
        L.sa_graph_destroy.argtypes = [vp]
        L._batched_sigs = True
    return L


def set_rows(rows: int) -> int:
    """Panel-update tile of the FP64 batched kernel: 2 or 4 16-row blocks
    per wave (0 = the default, 2).  Returns the previous setting.  A graph
    captured afterwards keeps the kernel chosen at capture time."""
    r = _lib().sa_set_batched_rows(int(rows))
    if r < 0:
        raise ValueError(f"batched panel rows must be 0, 2 or 4, not {rows}")
    return r


def _dev_from(a: np.ndarray):
    d = K.DeviceArray(max(1, a.size), 1, dtype=a.dtype)
    K.check(K.lib().sa_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes), "h2d")
    return d


class Batch:
    """`sizes[i]` x `sizes[i]` SPD systems A_i = B_i B_i^T / n_i + shift I
    (B_i i.i.d. N(0,1), seeded, generated on the GPU) with `nrhs`
    right-hand sides each (N(0,1))."""

    def __init__(self, sizes, nrhs: int = 1, shift: float = 0.1, seed: int = 0, ir: bool = False):
        L = _lib()
        self.n = np.asarray(sizes, dtype=np.int32)
        self.nsys = len(self.n)
        self.nrhs = nrhs
        n64 = self.n.astype(np.int64)
        self.offA = np.concatenate([[0], np.cumsum(n64 * n64)[:-1]]).astype(np.int64)
        self.offB = np.concatenate([[0], np.cumsum(n64 * nrhs)[:-1]]).astype(np.int64)
        dinv = np.array([L.sa_batched_dinv_elems(int(v)) for v in self.n], dtype=np.int64)
        self.offD = np.concatenate([[0], np.cumsum(dinv)[:-1]]).astype(np.int64)
        self.totA = int((n64 * n64).sum())
        self.totB = int((n64 * nrhs).sum())
        self.largest_first = np.argsort(-self.n, kind="stable").astype(np.int32)
        self.natural = np.arange(self.nsys, dtype=np.int32)
        # device arrays
        self.A0 = K.DeviceArray(self.totA, 1)
        self.A = K.DeviceArray(self.totA, 1)
        self.B0 = K.DeviceArray(max(1, self.totB), 1)
        self.B = K.DeviceArray(max(1, self.totB), 1)
        self.Dinv = K.DeviceArray(int(dinv.sum()), 1)
        self.info = K.DeviceArray(self.nsys, 1, dtype=np.int32)
        self.counter = K.DeviceArray(1, 1, dtype=np.int32)
        self.ir = ir
        if ir:
            # workspaces of the FP32 + FP64-refinement kernel: FP32 copies of
            # the matrices and block inverses, x and r in FP64, d in FP32
            self.Af = K.DeviceArray(self.totA, 1, dtype=np.float32)
            self.Dinvf = K.DeviceArray(int(dinv.sum()), 1, dtype=np.float32)
            self.X = K.DeviceArray(max(1, self.totB), 1)
            self.R = K.DeviceArray(max(1, self.totB), 1)
            self.Df = K.DeviceArray(max(1, self.totB), 1, dtype=np.float32)
            self.iters = K.DeviceArray(max(1, self.nsys * nrhs), 1, dtype=np.int32)
            self.berr = K.DeviceArray(max(1, self.nsys * nrhs), 1)
        self.d_n = _dev_from(self.n)
        self.d_offA, self.d_offB, self.d_offD = _dev_from(self.offA), _dev_from(self.offB), _dev_from(self.offD)
        self.d_order = {"largest_first": _dev_from(self.largest_first), "natural": _dev_from(self.natural)}
        nmax = int(self.n.max())
        G = K.DeviceArray(nmax, nmax)
        for i, v in enumerate(self.n):
            v = int(v)
            a = self.A0.ptr + int(self.offA[i]) * 8
            K.check(L.sa_fill_random(None, v, v, G.ptr, v, seed + 7919 * i, 1), "fill")
            K.dsyrk_ln(v, v, 1.0 / v, G.ptr, v, 0.0, a, v)
            K.check(L.sa_symmetrize_lower(None, v, a, v), "sym")
            K.check(L.sa_add_diagonal(None, v, a, v, shift), "diag")
            K.check(L.sa_fill_random(None, v, nrhs, self.B0.ptr + int(self.offB[i]) * 8, v,
                                     seed + 104729 * (i + 1), 1), "fill b")
        K.sync()
        G.free()
        self.restore()

    def flops(self) -> float:
        n = self.n.astype(np.float64)
        return float((n ** 3 / 3.0 + n ** 2 / 2.0 + n / 6.0 + 2.0 * n * n * self.nrhs).sum())

    def bytes(self) -> float:
        """Algorithmic HBM bytes of one batch: each matrix read and written
        once (lower half, 8 n^2/2 each way), right-hand sides twice."""
        n = self.n.astype(np.float64)
        return float((8.0 * n * n + 2 * 8.0 * n * self.nrhs).sum())

    def restore(self, stream=None):
        L = K.lib()
        K.check(L.sa_memcpy_d2d(self.A.ptr, self.A0.ptr, self.totA * 8), "restore A")
        if self.totB:
            K.check(L.sa_memcpy_d2d(self.B.ptr, self.B0.ptr, self.totB * 8), "restore B")

    def launch(self, grid: int, order: str = "largest_first", stream=None, rows: int | None = None):
        """FP64 factor + solve of every system.  `rows` (2 or 4) selects the
        panel-update tile, 16-row blocks per wave, for this and later
        launches; None keeps the current one."""
        if rows is not None:
            set_rows(rows)
        K.check(_lib().sa_dpotrs_batched(stream, self.nsys, self.d_order[order].ptr, self.d_n.ptr,
                                         self.d_offA.ptr, self.d_offB.ptr, self.d_offD.ptr, self.A.ptr,
                                         self.B.ptr, self.nrhs, self.Dinv.ptr, self.info.ptr, self.counter.ptr,
                                         grid), "sa_dpotrs_batched")

    def launch_ir(self, grid: int, order: str = "largest_first", stream=None, maxiter: int = 10,
                  tol: float = 1e-14):
        """FP32 factor + FP64 refinement of every system into X (A, B kept)."""
        assert self.ir, "Batch(..., ir=True) allocates the refinement workspaces"
        K.check(_lib().sa_dposv_batched_ir(stream, self.nsys, self.d_order[order].ptr, self.d_n.ptr,
                                           self.d_offA.ptr, self.d_offB.ptr, self.d_offD.ptr, self.A.ptr,
                                           self.B.ptr, self.X.ptr, self.nrhs, self.Af.ptr, self.Dinvf.ptr,
                                           self.R.ptr, self.Df.ptr, self.info.ptr, self.iters.ptr, self.berr.ptr,
                                           maxiter, tol, self.counter.ptr, grid), "sa_dposv_batched_ir")

    def solve_ir(self, grid: int, order: str = "largest_first", stream=None, maxiter: int = 10,
                 tol: float = 1e-14) -> np.ndarray:
        """launch_ir, then, as dsposv does, the FP64 factor and solve for
        every system whose FP32 factorisation broke down or whose refinement
        did not reach `tol` (factored in place in A, solution into X).
        Returns the indices of the systems that took the FP64 path."""
        self.launch_ir(grid, order, stream, maxiter, tol)
        K.sync()
        info = self.infos()
        be = self.berr.download()[: self.nsys * self.nrhs, 0].reshape(self.nsys, max(1, self.nrhs))
        bad = np.nonzero((info != 0) | ~(be <= tol).all(axis=1))[0].astype(np.int32)
        if bad.size:
            L = K.lib()
            for i in bad:
                o, v = int(self.offB[i]) * 8, int(self.n[i]) * self.nrhs * 8
                K.check(L.sa_memcpy_d2d(self.X.ptr + o, self.B.ptr + o, v), "b -> x")
            d_bad = _dev_from(bad)
            K.check(_lib().sa_dpotrs_batched(stream, int(bad.size), d_bad.ptr, self.d_n.ptr, self.d_offA.ptr,
                                             self.d_offB.ptr, self.d_offD.ptr, self.A.ptr, self.X.ptr, self.nrhs,
                                             self.Dinv.ptr, self.info.ptr, self.counter.ptr, grid),
                    "sa_dpotrs_batched")
            K.sync()
            d_bad.free()
        return bad

    def capture(self, grid: int, order: str, stream, ir: bool = False, maxiter: int = 10, tol: float = 1e-14,
                rows: int | None = None):
        """An executable graph of one batched launch on `stream`."""
        L = _lib()
        if rows is not None and not ir:
            set_rows(rows)
        K.check(L.sa_graph_begin(stream), "graph begin")
        if ir:
            self.launch_ir(grid, order, stream, maxiter, tol)
        else:
            self.launch(grid, order, stream)
        exe = C.c_void_p()
        K.check(L.sa_graph_end(stream, C.byref(exe)), "graph end")
        return exe

    def system(self, i: int):
        """(A_i, b_i, x_i, L_i) on the host: the pristine matrix, its
        right-hand sides, the solution and the factor (lower)."""
        v, r = int(self.n[i]), self.nrhs
        A = np.empty((v, v), order="F")
        Lf = np.empty((v, v), order="F")
        K.check(K.lib().sa_memcpy_d2h(A.ctypes.data, self.A0.ptr + int(self.offA[i]) * 8, v * v * 8), "d2h")
        K.check(K.lib().sa_memcpy_d2h(Lf.ctypes.data, self.A.ptr + int(self.offA[i]) * 8, v * v * 8), "d2h")
        b = np.empty((v, r), order="F")
        x = np.empty((v, r), order="F")
        K.check(K.lib().sa_memcpy_d2h(b.ctypes.data, self.B0.ptr + int(self.offB[i]) * 8, v * r * 8), "d2h")
        K.check(K.lib().sa_memcpy_d2h(x.ctypes.data, self.B.ptr + int(self.offB[i]) * 8, v * r * 8), "d2h")
        return A, b, x, np.tril(Lf)

    def infos(self) -> np.ndarray:
        return self.info.download()[:, 0].copy()

    def ir_result(self, i: int):
        """(x_i, iterations, backward errors) of the refinement kernel."""
        v, r = int(self.n[i]), self.nrhs
        x = np.empty((v, r), order="F")
        K.check(K.lib().sa_memcpy_d2h(x.ctypes.data, self.X.ptr + int(self.offB[i]) * 8, v * r * 8), "d2h")
        it = self.iters.download()[i * r:(i + 1) * r, 0].copy()
        be = self.berr.download()[i * r:(i + 1) * r, 0].copy()
        return x, it, be

    def free(self):
        extra = (self.Af, self.Dinvf, self.X, self.R, self.Df, self.iters, self.berr) if self.ir else ()
        for a in (self.A0, self.A, self.B0, self.B, self.Dinv, self.info, self.counter, self.d_n, self.d_offA,
                  self.d_offB, self.d_offD, *self.d_order.values(), *extra):
            a.free()
