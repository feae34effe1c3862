"""Writes dense_kats.npz: known answers from LAPACK (through scipy/numpy) and
from closed forms, used to pin the oracle's dense building blocks.

  qr_*     : R of a Householder QR (numpy.linalg.qr, LAPACK dgeqrf)
  cpqr_*   : R and the column order of a column-pivoted QR
             (scipy.linalg.qr(pivoting=True), LAPACK dgeqp3)
  lu_*     : X = A^{-1} B and log|det A| (numpy.linalg.solve / slogdet, dgesv)
  kern_*   : y = K x for Laplace, Yukawa and Gaussian kernels from the closed
             forms 1/r, exp(-a r)/r, exp(-a r^2) with the diagonal values of
             the reference's kernel.hpp:30-60 (1/s, 1/s, 1.0001)

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.linalg as sl

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dense_kats.npz")


def cpqr_cases(rng):
    cases = {}
    cases["gauss_12x9"] = rng.standard_normal((12, 9))
    cases["rank6_30x20"] = rng.standard_normal((30, 6)) @ rng.standard_normal((6, 20))
    U, _ = np.linalg.qr(rng.standard_normal((24, 16)))
    V, _ = np.linalg.qr(rng.standard_normal((16, 16)))
    cases["graded_24x16"] = (U * 10.0 ** -np.linspace(0, 14, 16)) @ V.T
    cases["wide_8x20"] = rng.standard_normal((8, 20))
    # far-field block of the Laplace kernel: 40 targets around (3,0,0), 16 sources in the unit cube
    src = rng.random((16, 3))
    tgt = rng.random((40, 3)) + np.array([3.0, 0.0, 0.0])
    cases["laplace_far_40x16"] = 1.0 / np.linalg.norm(tgt[:, None, :] - src[None, :, :], axis=-1)
    return cases


def kernel_matrix(kind, pts, diag, alpha):
    d = np.linalg.norm(pts[:, None, :] - pts[None, :, :], axis=-1)
    off = ~np.eye(len(pts), dtype=bool)
    K = np.empty_like(d)
    if kind == "laplace":
        K[off] = 1.0 / d[off]
        K[~off] = diag
    elif kind == "yukawa":
        K[off] = np.exp(-alpha * d[off]) / d[off]
        K[~off] = diag
    else:
        K[off] = np.exp(-alpha * d[off] ** 2)
        K[~off] = 1.0001
    return K


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    for name, A in cpqr_cases(rng).items():
        out[f"cpqr_{name}_A"] = A
        R, piv = sl.qr(A, mode="r", pivoting=True)
        out[f"cpqr_{name}_R"] = R
        out[f"cpqr_{name}_perm"] = piv.astype(np.int64)
        out[f"qr_{name}_R"] = np.linalg.qr(A, mode="r")
    for name, (n, nrhs) in {"lu_10": (10, 3), "lu_33": (33, 2)}.items():
        A = rng.standard_normal((n, n))
        if name == "lu_33":
            A = A + np.diag(rng.standard_normal(n) * 1e-3)  # forces row exchanges
        B = rng.standard_normal((n, nrhs))
        out[f"{name}_A"] = A
        out[f"{name}_B"] = B
        out[f"{name}_X"] = np.linalg.solve(A, B)
        out[f"{name}_logdet"] = np.array(np.linalg.slogdet(A)[1])
    pts = rng.random((60, 3))
    x = rng.random(60)
    out["kern_points"] = pts
    out["kern_x"] = x
    for kind, diag, alpha in (("laplace", 1e3, 0.0), ("yukawa", 1e3, 1.0), ("gaussian", 0.0, 2.0)):
        out[f"kern_{kind}_y"] = kernel_matrix(kind, pts, diag, alpha) @ x
    np.savez(OUT, **out)
    print("wrote", OUT, len(out), "arrays")


if __name__ == "__main__":
    main()
