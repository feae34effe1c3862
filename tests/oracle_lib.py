"""ctypes loader of the CPU oracle (oracle/_build/liboracle.so).

The oracle is the checker: tests compare the product's results against it.
Nothing in mxc_amd/ imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

c_ll = C.c_longlong
c_d = C.c_double
c_int = C.c_int
P = C.c_void_p

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, stdout=subprocess.DEVNULL)
        L = C.CDLL(LIB)
        sig = {
            "orc_uniform_cube_rnd": (None, [P, c_ll, c_d, c_ll, C.c_uint]),
            "orc_sphere_rnd": (None, [P, c_ll, C.c_uint]),
            "orc_mesh_sphere": (None, [P, c_ll, c_d]),
            "orc_uniform_vec": (None, [P, c_ll, c_d, c_d, C.c_ulonglong]),
            "orc_levels": (c_ll, [c_ll, c_ll]),
            "orc_build_tree": (None, [P, c_ll, c_ll, P, P, P, P]),
            "orc_csr": (c_ll, [c_int, c_d, c_ll, P, P, P, P, P, c_ll]),
            "orc_far_sample": (c_ll, [c_ll, P, P, P, P, P, P, P, c_int, c_ll, c_ll, P, c_ll]),
            "orc_kernel_eval": (c_d, [c_int, c_d, c_d, c_d]),
            "orc_set_sketch": (None, [c_int]),
            "orc_rademacher": (c_d, [C.c_ulonglong, C.c_ulonglong]),
            "orc_gen_matrix": (None, [c_int, c_d, c_d, c_ll, c_ll, P, P, P]),
            "orc_dense_matvec": (None, [c_int, c_d, c_d, P, c_ll, P, P, c_int]),
            "orc_qr_r": (None, [c_ll, c_ll, P, c_ll]),
            "orc_cpqr": (c_ll, [c_ll, c_ll, P, c_ll, P, c_d, P, P]),
            "orc_lu_solve": (c_d, [c_ll, c_ll, P, P]),
            "orc_problem_create": (P, [P, c_ll, c_ll]),
            "orc_problem_destroy": (None, [P]),
            "orc_problem_levels": (c_ll, [P]),
            "orc_problem_bodies": (None, [P, P]),
            "orc_h2_create": (P, [P, c_int, c_d, c_d, c_d, c_d, c_ll, c_ll, c_int, c_int, c_ll, c_int]),
            "orc_h2_destroy": (None, [P]),
            "orc_h2_set_threads": (None, [P, c_int]),
            "orc_h2_nboxes": (c_ll, [P, c_ll]),
            "orc_h2_dims": (None, [P, c_ll, P, P]),
            "orc_h2_near_count": (c_ll, [P, c_ll]),
            "orc_h2_download": (c_ll, [P, c_ll, c_int, P]),
            "orc_h2_level_list": (c_ll, [P, c_ll, c_int, P, P]),
            "orc_h2_far_count": (c_ll, [P, c_ll]),
            "orc_h2_matvec": (None, [P, P]),
            "orc_h2_factorize": (c_int, [P]),
            "orc_h2_far_level": (None, [P, c_ll, P, P, P]),
            "orc_h2_precondition": (None, [P, P]),
            "orc_h2_gmres": (c_ll, [P, P, c_d, P, P, c_ll, c_ll, P]),
            "orc_rel_err": (c_d, [c_ll, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


KID = {"laplace": 0, "yukawa": 1, "gaussian": 2}


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def cube_points(n: int, seed: int) -> np.ndarray:
    p = np.zeros(3 * n)
    lib().orc_uniform_cube_rnd(ptr(p), n, 1.0, 3, seed)
    return p


def sphere_points(n: int, seed: int) -> np.ndarray:
    p = np.zeros(3 * n)
    lib().orc_sphere_rnd(ptr(p), n, seed)
    return p


def uniform_vec(n: int, lo: float, hi: float, seed: int) -> np.ndarray:
    x = np.zeros(n)
    lib().orc_uniform_vec(ptr(x), n, lo, hi, seed)
    return x


def build_tree(points: np.ndarray, levels: int):
    b = points.copy()
    nc = (2 << levels) - 1
    child = np.zeros(2 * nc, np.int64)
    body = np.zeros(2 * nc, np.int64)
    R = np.zeros(3 * nc)
    Cc = np.zeros(3 * nc)
    lib().orc_build_tree(ptr(b), len(points) // 3, levels, ptr(child), ptr(body), ptr(R), ptr(Cc))
    return b, child, body, R, Cc


def csr(far: bool, theta: float, levels: int, child, R, Cc):
    nc = (2 << levels) - 1
    rp = np.zeros(nc + 1, np.int64)
    nnz = lib().orc_csr(int(far), theta, levels, ptr(child), ptr(R), ptr(Cc), ptr(rp), None, 0)
    col = np.zeros(max(nnz, 1), np.int64)
    lib().orc_csr(int(far), theta, levels, ptr(child), ptr(R), ptr(Cc), ptr(rp), ptr(col), nnz)
    return rp, col[:nnz]


def dense_matvec(kernel: str, diag: float, alpha: float, bodies: np.ndarray, x: np.ndarray, nthreads: int = 8):
    y = np.zeros_like(x)
    lib().orc_dense_matvec(KID[kernel], diag, alpha, ptr(bodies), len(x), ptr(x), ptr(y), nthreads)
    return y


class Problem:
    """Oracle problem: tree-ordered bodies and the two H^2 matrices of lorasp."""

    def __init__(self, points: np.ndarray, leaf: int):
        self.n = len(points) // 3
        self.h = lib().orc_problem_create(ptr(points), self.n, leaf)
        self.levels = lib().orc_problem_levels(self.h)
        self.bodies = np.zeros(3 * self.n)
        lib().orc_problem_bodies(self.h, ptr(self.bodies))

    def h2(self, kernel="laplace", diag=1e3, alpha=1.0, theta=2.0, tol=0.0, rank=0, leveled_rank=0, orth=False,
           hss_far=False, sampling=0, nthreads=8):
        return H2(self, lib().orc_h2_create(self.h, KID[kernel], diag, alpha, theta, tol, rank, leveled_rank, int(orth),
                                            int(hss_far), sampling, nthreads))

    def __del__(self):
        try:
            lib().orc_problem_destroy(self.h)
        except Exception:
            pass


class H2:
    def __init__(self, prob: Problem, h):
        self.prob = prob
        self.h = h

    def dims(self, level: int):
        nb = lib().orc_h2_nboxes(self.h, level)
        d = np.zeros(nb, np.int64)
        r = np.zeros(nb, np.int64)
        lib().orc_h2_dims(self.h, level, ptr(d), ptr(r))
        return d, r

    def download(self, level: int, what: int) -> np.ndarray:
        out = np.zeros(lib().orc_h2_download(self.h, level, what, None))
        lib().orc_h2_download(self.h, level, what, ptr(out))
        return out

    def level_list(self, level: int, far: bool):
        """Level-local near (far=False) or far (far=True) CSR list: (rowptr, col)."""
        nb = lib().orc_h2_nboxes(self.h, level)
        nnz = lib().orc_h2_level_list(self.h, level, int(far), None, None)
        rp = np.zeros(nb + 1, np.int64)
        col = np.zeros(max(nnz, 1), np.int64)
        lib().orc_h2_level_list(self.h, level, int(far), ptr(rp), ptr(col))
        return rp, col[:nnz]

    def matvec(self, x: np.ndarray) -> np.ndarray:
        y = np.ascontiguousarray(x, dtype=np.float64).copy()
        lib().orc_h2_matvec(self.h, ptr(y))
        return y

    def far_level(self, level: int, x: np.ndarray):
        """(approximate, exact) far field of the couplings of one level."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        ya, ye = np.zeros_like(x), np.zeros_like(x)
        lib().orc_h2_far_level(self.h, level, ptr(x), ptr(ya), ptr(ye))
        return ya, ye

    def set_threads(self, n: int) -> None:
        lib().orc_h2_set_threads(self.h, n)

    def factorize(self) -> int:
        return lib().orc_h2_factorize(self.h)

    def precondition(self, x: np.ndarray) -> np.ndarray:
        y = np.ascontiguousarray(x, dtype=np.float64).copy()
        lib().orc_h2_precondition(self.h, ptr(y))
        return y

    def gmres(self, M: "H2", b: np.ndarray, tol: float, inner: int = 10, outer: int = 50):
        x = np.zeros_like(b)
        res = np.zeros(outer + 1)
        it = lib().orc_h2_gmres(self.h, M.h, tol, ptr(x), ptr(b), inner, outer, ptr(res))
        return x, res[: it + 1]

    def __del__(self):
        try:
            lib().orc_h2_destroy(self.h)
        except Exception:
            pass


def rel_err(x: np.ndarray, ref: np.ndarray) -> float:
    return float(np.sqrt(np.sum((x - ref) ** 2) / np.sum(ref**2)))
