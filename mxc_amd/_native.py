"""ctypes bindings of the in-tree native libraries.

``libmxc_gpu.so`` is the HIP device layer (include/mxc_gpu.h) and
``libmxc_h2.so`` the C++ host layer (include/mxc_h2.h). Both are built into
``mxc_amd/lib`` by ``mxc_amd.build.build()``; loading fails loudly when they
are missing, there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

c_ll = C.c_longlong
c_int = C.c_int
c_d = C.c_double
P = C.c_void_p

MXG_ERR_ARG = 1  # invalid argument or size beyond a documented limit
MXG_FP64 = 0
MXG_FP32 = 1
MXG_BF16 = 2  # ULV only: FP32 factorization, bfloat16-stored factors for the substitution
MXG_BF16_SCHUR = 3  # ULV only: as MXG_BF16, with bfloat16 inputs to the Schur-complement products
KERNELS = {"laplace": 0, "yukawa": 1, "gaussian": 2}
GMRES_REFERENCE = 0
GMRES_CGS2 = 1
MXG_H2_FACTORABLE = 1
# level data kinds of mxg_h2_download / mxg_h2_upload
H2_NEAR, H2_BASIS, H2_R, H2_COUPLING, H2_SKELETON = 0, 1, 2, 3, 4
# kernel timing classes of mxg_profile_read used from Python
MXG_PROF_SKETCH = 9


class MxgKernel(C.Structure):
    _fields_ = [("id", c_int), ("diag", c_d), ("alpha", c_d)]


class MxgGmresOpts(C.Structure):
    _fields_ = [("tol", c_d), ("restart", c_int), ("max_cycles", c_int), ("mode", c_int), ("inner_precision", c_int),
                ("inner_operator", C.c_void_p), ("graph", c_int)]


class MxcH2Opts(C.Structure):
    _fields_ = [
        ("kernel", c_int),
        ("diag", c_d),
        ("alpha", c_d),
        ("theta", c_d),
        ("tol", c_d),
        ("rank", c_ll),
        ("leveled_rank", c_ll),
        ("orth", c_int),
        ("hss_far", c_int),
        ("sampling", c_ll),
        ("factorable", c_int),
        ("householder_rows", c_int),
        ("explicit_far", c_int),
    ]


class NativeError(RuntimeError):
    pass


def _load(name: str) -> C.CDLL:
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise NativeError(f"{path} is missing: build it with mxc_amd.build.build() (no fallback exists)")
    return C.CDLL(path)


_gpu = None
_h2 = None


def gpu() -> C.CDLL:
    global _gpu
    if _gpu is None:
        L = _load("libmxc_gpu.so")
        L.mxg_last_error.restype = C.c_char_p
        L.mxg_stream.restype = P
        L.mxg_stream.argtypes = [P]
        sigs = {
            "mxg_init": [c_int, C.POINTER(P)],
            "mxg_finalize": [P],
            "mxg_device_count": [C.POINTER(c_int)],
            "mxg_synchronize": [P],
            "mxg_mem_info": [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
            "mxg_malloc": [P, C.c_size_t, C.POINTER(P)],
            "mxg_free": [P, P],
            "mxg_memcpy_h2d": [P, P, P, C.c_size_t],
            "mxg_memcpy_d2h": [P, P, P, C.c_size_t],
            "mxg_memcpy_d2d": [P, P, P, C.c_size_t],
            "mxg_memset": [P, P, c_int, C.c_size_t],
            "mxg_gemm": [P, c_int, c_int, c_int, c_ll, c_ll, c_ll, c_d, P, c_ll, P, c_ll, c_d, P, c_ll],
            "mxg_gen_block": [P, MxgKernel, c_ll, c_ll, P, P, P, c_ll],
            "mxg_dense_matvec": [P, MxgKernel, P, c_ll, P, P],
            "mxg_dense_matvec_batch": [P, MxgKernel, P, c_ll, P, c_ll, P, c_ll, c_ll],
            "mxg_dense_matvec_batch_noalloc": [P, MxgKernel, P, c_ll, P, c_ll, P, c_ll, c_ll],
            "mxg_block_gemv": [P, c_int, c_ll, c_ll, P, c_ll, P, c_ll, P, c_ll, c_ll, c_int],
            "mxg_h2_create": [P, MxgKernel, P, c_ll, c_ll, c_int, C.POINTER(P)],
            "mxg_h2_destroy": [P],
            "mxg_h2_define_level": [P, c_ll, c_ll, P, P, P, P, P, P],
            "mxg_h2_construct_level": [P, c_ll, c_ll, P, P, P, P, P, P, P, P, c_d, c_ll, c_int, P],
            "mxg_h2_construct_level_runs": [P, c_ll, c_ll, P, P, P, P, P, P, P, P, c_d, c_ll, c_int, P],
            "mxg_h2_upload": [P, c_ll, c_int, P, c_ll],
            "mxg_ulv_create": [P, c_int, C.POINTER(P)],
            "mxg_ulv_destroy": [P],
            "mxg_h2_matvec": [P, P, P, c_ll, c_ll, c_int],
            "mxg_h2_autotune": [P, c_int, c_ll, c_int, C.POINTER(c_d)],
            "mxg_h2_get_tuning": [P, c_int, C.POINTER(c_int), C.POINTER(c_int), C.POINTER(c_int), C.POINTER(c_int)],
            "mxg_h2_set_tuning": [P, c_int, c_int, c_int, c_int, c_int],
            "mxg_h2_download": [P, c_ll, c_int, P, c_ll],
            "mxg_ulv_solve": [P, P, c_ll, c_ll],
            "mxg_ulv_info": [P, P],
            "mxg_gmres": [P, P, C.POINTER(MxgGmresOpts), P, P, P, C.POINTER(c_ll), C.POINTER(c_ll)],
            "mxg_profile_enable": [P, c_int],
            "mxg_profile_reset": [P],
            "mxg_profile_read": [P, c_int, C.POINTER(c_ll), C.POINTER(c_d)],
            "mxg_profile_work": [P, c_int, C.POINTER(c_d)],
        }
        for name, args in sigs.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = c_int
        L.mxg_h2_level_size.argtypes = [P, c_ll, c_int]
        L.mxg_h2_level_size.restype = c_ll
        if hasattr(L, "mxg_ulv_factor_seconds"):
            L.mxg_ulv_factor_seconds.argtypes = [P]
            L.mxg_ulv_factor_seconds.restype = c_d
        _gpu = L
    return _gpu


def h2() -> C.CDLL:
    global _h2
    if _h2 is None:
        gpu()
        L = _load("libmxc_h2.so")
        L.mxc_last_error.restype = C.c_char_p
        sigs = {
            "mxc_generate_points": [c_int, c_ll, C.c_uint, P],
            "mxc_uniform_vector": [P, c_ll, c_d, c_d, C.c_ulonglong],
            "mxc_build_tree": [P, c_ll, c_ll, P, P, P, P],
            "mxc_problem_create": [P, P, c_ll, c_ll, C.POINTER(P)],
            "mxc_problem_destroy": [P],
            "mxc_problem_bodies": [P, P],
            "mxc_solver_create": [P, C.POINTER(MxcH2Opts), C.POINTER(P)],
            "mxc_solver_destroy": [P],
            "mxc_solver_dims": [P, c_ll, P, P],
            "mxc_solver_counts": [P, c_ll, P],
            "mxc_solver_matvec": [P, P, P, c_int],
            "mxc_solver_factorize": [P, c_int, P],
            "mxc_solver_factorize_batch": [P, c_ll, c_int, P],
            "mxc_solver_precondition": [P, P],
            "mxc_solver_gmres": [P, P, C.POINTER(MxgGmresOpts), P, P, P, C.POINTER(c_ll), C.POINTER(c_ll)],
            "mxc_solver_gmres_batch": [P, P, C.POINTER(MxgGmresOpts), c_ll, P, c_ll, P, c_ll, P, P, P],
            "mxc_dense_matvec": [P, MxgKernel, P, P],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = c_int
        L.mxc_tree_levels.argtypes = [c_ll, c_ll]
        L.mxc_tree_levels.restype = c_ll
        L.mxc_problem_levels.argtypes = [P]
        L.mxc_problem_levels.restype = c_ll
        L.mxc_problem_device_bodies.argtypes = [P]
        L.mxc_problem_device_bodies.restype = P
        L.mxc_problem_ctx.argtypes = [P]
        L.mxc_problem_ctx.restype = P
        L.mxc_solver_device.argtypes = [P]
        L.mxc_solver_device.restype = P
        L.mxc_solver_ulv.argtypes = [P]
        L.mxc_solver_ulv.restype = P
        L.mxc_csr.argtypes = [c_int, c_d, c_ll, P, P, P, P, P, P, c_ll]
        L.mxc_csr.restype = c_ll
        L.mxc_far_sample.argtypes = [c_ll, P, P, P, P, P, P, P, c_int, c_ll, c_ll, P, c_ll]
        L.mxc_far_sample.restype = c_ll
        _h2 = L
    return _h2


def check_gpu(rc: int) -> None:
    if rc != 0:
        raise NativeError(f"mxg error {rc}: {gpu().mxg_last_error().decode()}")


def check_h2(rc: int) -> None:
    if rc != 0:
        raise NativeError(f"mxc error {rc}: {h2().mxc_last_error().decode()}")
