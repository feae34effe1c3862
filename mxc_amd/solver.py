"""Python mirror of the host API (H2MatrixSolver, solver.hpp:11-47), driving
the C++ host layer (libmxc_h2) and through it the HIP device layer.

Vectors are DeviceArray objects holding tree-ordered FP64 data.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .device import Context, DeviceArray, kernel_desc

GEOMETRY = {"cube": 0, "sphere": 1, "mesh": 2}


def generate_points(geometry: str, n: int, seed: int) -> np.ndarray:
    out = np.zeros(3 * n)
    N.check_h2(N.h2().mxc_generate_points(GEOMETRY[geometry], n, seed, out.ctypes.data))
    return out


def uniform_vector(n: int, lo: float, hi: float, seed: int) -> np.ndarray:
    x = np.zeros(n)
    N.check_h2(N.h2().mxc_uniform_vector(x.ctypes.data, n, lo, hi, seed))
    return x


def tree_levels(n: int, leaf: int) -> int:
    return int(N.h2().mxc_tree_levels(n, leaf))


def build_tree(points: np.ndarray, levels: int):
    """Tree-ordered copy of the points and the cell arrays (child, body, R, C)."""
    b = np.ascontiguousarray(points, dtype=np.float64).copy()
    nc = (2 << levels) - 1
    child = np.zeros(2 * nc, np.int64)
    body = np.zeros(2 * nc, np.int64)
    R = np.zeros(3 * nc)
    Cc = np.zeros(3 * nc)
    N.check_h2(N.h2().mxc_build_tree(b.ctypes.data, len(b) // 3, levels, child.ctypes.data, body.ctypes.data,
                                     R.ctypes.data, Cc.ctypes.data))
    return b, child, body, R, Cc


def csr(far: bool, theta: float, levels: int, child, body, R, Cc):
    nc = (2 << levels) - 1
    rp = np.zeros(nc + 1, np.int64)
    L = N.h2()
    nnz = L.mxc_csr(int(far), theta, levels, child.ctypes.data, body.ctypes.data, R.ctypes.data, Cc.ctypes.data,
                    rp.ctypes.data, None, 0)
    if nnz < 0:
        raise N.NativeError(L.mxc_last_error().decode())
    col = np.zeros(max(nnz, 1), np.int64)
    L.mxc_csr(int(far), theta, levels, child.ctypes.data, body.ctypes.data, R.ctypes.data, Cc.ctypes.data,
              rp.ctypes.data, col.ctypes.data, nnz)
    return rp, col[:nnz]


def far_sample(levels: int, body, bfar, near, far, with_near: bool, cell: int, sampling: int) -> np.ndarray:
    """Sampled far-body list of one cell; bfar/near/far are (rowptr, col) CSR pairs."""
    L = N.h2()
    args = (levels, body.ctypes.data, bfar[0].ctypes.data, bfar[1].ctypes.data, near[0].ctypes.data,
            near[1].ctypes.data, far[0].ctypes.data, far[1].ctypes.data, int(with_near), cell, sampling)
    n = L.mxc_far_sample(*args, None, 0)
    out = np.zeros(max(n, 1), np.int64)
    L.mxc_far_sample(*args, out.ctypes.data, n)
    return out[:n]


class Problem:
    """Points sorted into tree order and resident on one device."""

    def __init__(self, ctx: Context, points: np.ndarray, leaf: int):
        self.ctx = ctx
        pts = np.ascontiguousarray(points, dtype=np.float64)
        self.n = len(pts) // 3
        h = C.c_void_p()
        N.check_h2(N.h2().mxc_problem_create(ctx.handle, pts.ctypes.data, self.n, leaf, C.byref(h)))
        self.handle = h.value
        self.levels = int(N.h2().mxc_problem_levels(self.handle))
        self.bodies = np.zeros(3 * self.n)
        N.check_h2(N.h2().mxc_problem_bodies(self.handle, self.bodies.ctypes.data))
        ctx.adopt(self, Context.PROBLEMS)

    def dense_matvec(self, kernel: str, diag: float, alpha: float, x: DeviceArray, y: DeviceArray) -> None:
        """y = K x with the matrix-free FP64 dense product; x and y may hold
        several columns (n x nrhs)."""
        if len(x.shape) > 1:
            assert x.shape == y.shape and x.shape[0] == self.n
            bodies = N.h2().mxc_problem_device_bodies(self.handle)
            N.check_gpu(N.gpu().mxg_dense_matvec_batch(self.ctx.handle, kernel_desc(kernel, diag, alpha), bodies, self.n,
                                                       x.ptr, self.n, y.ptr, self.n, x.shape[1]))
            return
        N.check_h2(N.h2().mxc_dense_matvec(self.handle, kernel_desc(kernel, diag, alpha), x.ptr, y.ptr))

    def close(self):
        if self.handle and self.ctx.handle:
            N.h2().mxc_problem_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class H2:
    """An H^2 matrix on the device (H2MatrixSolver)."""

    def __init__(self, problem: Problem, kernel="laplace", diag=1e3, alpha=1.0, theta=2.0, tol=0.0, rank=0,
                 leveled_rank=0, orth=False, hss_far=False, sampling=0, factorable=False, householder_rows=False,
                 explicit_far=False):
        self.problem = problem
        self.kernel = (kernel, diag, alpha)
        o = N.MxcH2Opts(N.KERNELS[kernel], diag, alpha, theta, tol, rank, leveled_rank, int(orth), int(hss_far), sampling,
                        int(factorable), int(householder_rows), int(explicit_far))
        h = C.c_void_p()
        N.check_h2(N.h2().mxc_solver_create(problem.handle, C.byref(o), C.byref(h)))
        self.handle = h.value
        self.levels = problem.levels
        problem.ctx.adopt(self, Context.SOLVERS)

    @classmethod
    def operator(cls, problem, eps, sampling, kernel="laplace", diag=1e3, alpha=1.0, theta=2.0, householder_rows=False,
                 explicit_far=False):
        """The accurate operator of lorasp: interpolative bases at tolerance eps (solver.cpp:28-36, epi < 1)."""
        return cls(problem, kernel, diag, alpha, theta, tol=eps, sampling=sampling, householder_rows=householder_rows,
                   explicit_far=explicit_far)

    @classmethod
    def preconditioner(cls, problem, rank, sampling, kernel="laplace", diag=1e3, alpha=1.0, theta=2.0, eps=0.0,
                       leveled_rank=0, householder_rows=False):
        """The ULV preconditioner of lorasp: orthogonal bases of fixed rank (epi == 0), or truncated at eps
        with the rank as a cap, basis far lists from the sibling lists (solver.cpp:28-36)."""
        return cls(problem, kernel, diag, alpha, theta, tol=eps, rank=rank, leveled_rank=leveled_rank, orth=True,
                   hss_far=True, sampling=sampling, factorable=True, householder_rows=householder_rows)

    def dims(self, level: int):
        nb = 1 << level
        d = np.zeros(nb, np.int64)
        r = np.zeros(nb, np.int64)
        N.check_h2(N.h2().mxc_solver_dims(self.handle, level, d.ctypes.data, r.ctypes.data))
        return d, r

    def counts(self, level: int):
        c = np.zeros(2, np.int64)
        N.check_h2(N.h2().mxc_solver_counts(self.handle, level, c.ctypes.data))
        return int(c[0]), int(c[1])

    def download(self, level: int, what: int) -> np.ndarray:
        """Level data on the host: what 0 = near blocks, 1 = bases Q, 2 = R, 3 = couplings, 4 = skeleton indices."""
        dev = N.h2().mxc_solver_device(self.handle)
        n = N.gpu().mxg_h2_level_size(dev, level, what)
        out = np.zeros(max(n, 0))
        N.check_gpu(N.gpu().mxg_h2_download(dev, level, what, out.ctypes.data, n))
        return out

    def matvec(self, x: DeviceArray, y: DeviceArray, precision: int = N.MXG_FP64) -> None:
        N.check_h2(N.h2().mxc_solver_matvec(self.handle, x.ptr, y.ptr, precision))

    def autotune(self, precision: int = N.MXG_FP64, nrhs: int = 8, reps: int = 3):
        """Times the matvec's candidate launch shapes for nrhs right-hand sides
        and keeps the fastest (mxg_h2_autotune); returns (ms per product,
        tuning dict)."""
        ms = C.c_double()
        dev = N.h2().mxc_solver_device(self.handle)
        N.check_gpu(N.gpu().mxg_h2_autotune(dev, precision, nrhs, reps, C.byref(ms)))
        return ms.value, self.tuning(precision)

    def tuning(self, precision: int = N.MXG_FP64) -> dict:
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        dev = N.h2().mxc_solver_device(self.handle)
        N.check_gpu(N.gpu().mxg_h2_get_tuning(dev, precision, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return {"near_symmetric": a.value, "far_symmetric": b.value, "split_rows": c.value, "matrix_cores": d.value}

    def set_tuning(self, precision: int, near_symmetric: int, far_symmetric: int, split_rows: int,
                   matrix_cores: int = 0) -> None:
        dev = N.h2().mxc_solver_device(self.handle)
        N.check_gpu(N.gpu().mxg_h2_set_tuning(dev, precision, near_symmetric, far_symmetric, split_rows, matrix_cores))

    def factorize(self, precision: int = N.MXG_FP64) -> np.ndarray:
        info = np.zeros(self.levels + 1, np.int32)
        N.check_h2(N.h2().mxc_solver_factorize(self.handle, precision, info.ctypes.data))
        return info

    @staticmethod
    def factorize_batch(solvers, precision: int = N.MXG_FP32):
        """ULV factorizations of several independent H^2 matrices on one
        context, every phase of a level sweep one launch shared by all of them;
        returns the singular-block counts per level of each."""
        hs = (C.c_void_p * len(solvers))(*[s.handle for s in solvers])
        info = np.zeros(sum(s.levels + 1 for s in solvers), np.int32)
        N.check_h2(N.h2().mxc_solver_factorize_batch(hs, len(solvers), precision, info.ctypes.data))
        out, o = [], 0
        for s in solvers:
            out.append(info[o:o + s.levels + 1])
            o += s.levels + 1
        return out

    def factor_seconds(self) -> float:
        return float(N.gpu().mxg_ulv_factor_seconds(N.h2().mxc_solver_ulv(self.handle)))

    def precondition(self, x: DeviceArray) -> None:
        N.check_h2(N.h2().mxc_solver_precondition(self.handle, x.ptr))

    def gmres(self, M: "H2", b: DeviceArray, x: DeviceArray, tol: float, restart: int = 10, max_cycles: int = 50,
              mode: int = N.GMRES_REFERENCE, inner_precision: int = N.MXG_FP64, inner: "H2" = None):
        """GMRES on this operator with left preconditioner M. `inner` (optional)
        is the operator applied in the inner iterations; every cycle's residual
        uses this one."""
        inner_dev = N.h2().mxc_solver_device(inner.handle) if inner is not None else None
        o = N.MxgGmresOpts(tol, restart, max_cycles, mode, inner_precision, inner_dev)
        resid = np.zeros(max_cycles + 1)
        it = C.c_longlong()
        inner = C.c_longlong()
        N.check_h2(N.h2().mxc_solver_gmres(self.handle, M.handle, C.byref(o), b.ptr, x.ptr, resid.ctypes.data,
                                           C.byref(it), C.byref(inner)))
        return resid[: it.value + 1], it.value, inner.value

    def gmres_batch(self, M: "H2", B: DeviceArray, X: DeviceArray, tol: float, restart: int = 10, max_cycles: int = 50,
                    mode: int = N.GMRES_REFERENCE, inner_precision: int = N.MXG_FP64, inner: "H2" = None,
                    graph: bool = False):
        """GMRES for the columns of B (n x nrhs) in lockstep, solutions into X
        (initial guesses on entry). Returns the residual histories (nrhs x
        (max_cycles + 1)), cycles and inner iterations per right-hand side.
        graph: replay the inner steps' launches as one HIP graph (CGS2 with
        FP32 inner iterations)."""
        n = self.problem.n
        nrhs = B.shape[1] if len(B.shape) > 1 else 1
        assert X.shape == B.shape and B.shape[0] == n
        inner_dev = N.h2().mxc_solver_device(inner.handle) if inner is not None else None
        o = N.MxgGmresOpts(tol, restart, max_cycles, mode, inner_precision, inner_dev, int(graph))
        resid = np.zeros((nrhs, max_cycles + 1))
        it = np.zeros(nrhs, np.int64)
        inn = np.zeros(nrhs, np.int64)
        N.check_h2(N.h2().mxc_solver_gmres_batch(self.handle, M.handle, C.byref(o), nrhs, B.ptr, n, X.ptr, n,
                                                 resid.ctypes.data, it.ctypes.data, inn.ctypes.data))
        return [resid[q, : it[q] + 1] for q in range(nrhs)], it, inn

    def close(self):
        if self.handle and self.problem.ctx.handle:
            N.h2().mxc_solver_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostLevels:
    """An H^2 matrix handed over level by level from host data, as the
    reference's H2MatrixSolver does when it factorizes on the device
    (factorizeDeviceM, solver.cpp:102-119: createMatrixDesc +
    copyDataInMatrixDesc, device_precondition.cu:39-154,198-202).

    levels[l] (l = 0 .. L) is a dict with the level-local structure "dims",
    "ranks", "near" = (rowptr, col), "far" = (rowptr, col), and the data the
    caller has: "A" (near blocks; at an upper level the couplings of the far
    child pairs), "Q" (full M x M bases), "R", "C" (couplings), each in the
    layout of mxg_h2_download. Vectors are tree-ordered over problem.bodies."""

    def __init__(self, problem: Problem, levels, kernel="laplace", diag=1e3, alpha=1.0, factorable=True):
        self.problem = problem
        self.levels = len(levels) - 1
        assert self.levels == problem.levels
        ctx = problem.ctx
        h = C.c_void_p()
        N.check_gpu(N.gpu().mxg_h2_create(ctx.handle, kernel_desc(kernel, diag, alpha), problem.bodies.ctypes.data,
                                          problem.n, self.levels, N.MXG_H2_FACTORABLE if factorable else 0, C.byref(h)))
        self.handle = h.value
        self.ulv = None
        ctx.adopt(self, Context.SOLVERS)
        i64 = lambda a: np.ascontiguousarray(a, dtype=np.int64)  # noqa: E731
        for l in range(self.levels, -1, -1):
            lv = levels[l]
            dims, ranks = i64(lv["dims"]), i64(lv["ranks"])
            (nrp, ncol), (frp, fcol) = (tuple(map(i64, lv["near"])), tuple(map(i64, lv["far"])))
            N.check_gpu(N.gpu().mxg_h2_define_level(self.handle, l, len(dims), dims.ctypes.data, ranks.ctypes.data,
                                                    nrp.ctypes.data, ncol.ctypes.data, frp.ctypes.data, fcol.ctypes.data))
            for key, what in (("A", N.H2_NEAR), ("Q", N.H2_BASIS), ("R", N.H2_R), ("C", N.H2_COUPLING)):
                if key in lv:
                    self.upload(l, what, lv[key])

    def upload(self, level: int, what: int, data) -> None:
        a = np.ascontiguousarray(data, dtype=np.float64)
        N.check_gpu(N.gpu().mxg_h2_upload(self.handle, level, what, a.ctypes.data, a.size))

    def download(self, level: int, what: int) -> np.ndarray:
        n = N.gpu().mxg_h2_level_size(self.handle, level, what)
        out = np.zeros(max(n, 0))
        N.check_gpu(N.gpu().mxg_h2_download(self.handle, level, what, out.ctypes.data, n))
        return out

    def matvec(self, x: DeviceArray, y: DeviceArray, precision: int = N.MXG_FP64) -> None:
        nrhs = x.shape[1] if len(x.shape) > 1 else 1
        N.check_gpu(N.gpu().mxg_h2_matvec(self.handle, x.ptr, y.ptr, nrhs, self.problem.n, precision))

    def factorize(self, precision: int = N.MXG_FP64) -> np.ndarray:
        if self.ulv:
            N.gpu().mxg_ulv_destroy(self.ulv)
            self.ulv = None
        f = C.c_void_p()
        N.check_gpu(N.gpu().mxg_ulv_create(self.handle, precision, C.byref(f)))
        self.ulv = f.value
        info = np.zeros(self.levels + 1, np.int32)
        N.check_gpu(N.gpu().mxg_ulv_info(self.ulv, info.ctypes.data))
        return info

    def precondition(self, x: DeviceArray) -> None:
        nrhs = x.shape[1] if len(x.shape) > 1 else 1
        N.check_gpu(N.gpu().mxg_ulv_solve(self.ulv, x.ptr, nrhs, self.problem.n))

    def close(self):
        if self.handle and self.problem.ctx.handle:
            if self.ulv:
                N.gpu().mxg_ulv_destroy(self.ulv)
            N.gpu().mxg_h2_destroy(self.handle)
        self.handle = None
        self.ulv = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
