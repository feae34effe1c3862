"""The symmetric block products (symv.hip) against the general ones.

An operator built on the device from a kernel with interpolative bases reads
each stored pair of transposed leaf near blocks (A_ji = A_ij^T) and of
couplings (C_ji = C_ij^T, generated on the skeleton points) once. The same
level data handed back through the upload path (mxg_h2_define_level /
mxg_h2_upload, which makes no symmetry assumption) runs the general block
products, which read every block. Both must give the same matvec to rounding, in FP64 and FP32,
on leaves of 256 points (16-byte panel loads), ragged leaves of 93/94 points
(element loads) and leaves of 312/313 points (two row chunks per block); and
every column of a batch must equal its single-column product bit for bit."""
import numpy as np
import pytest

import oracle_lib as O
from mxc_amd import _native as N
from mxc_amd.solver import H2, HostLevels, Problem, build_tree, csr, generate_points

pytestmark = pytest.mark.gpu

CASES = [("cube", 4096, 256, "laplace", 0.0), ("sphere", 3000, 128, "yukawa", 1.0), ("cube", 20000, 256, "laplace", 0.0)]


def level_list(rp, col, l):
    c0, nb = (1 << l) - 1, 1 << l
    return rp[c0:c0 + nb + 1] - rp[c0], col[rp[c0]:rp[c0 + nb]] - c0


def downloaded_levels(A, prob, points, theta):
    _, child, body, R, Cc = build_tree(points, prob.levels)
    near = csr(False, theta, prob.levels, child, body, R, Cc)
    far = csr(True, theta, prob.levels, child, body, R, Cc)
    out = []
    for l in range(prob.levels + 1):
        d, r = A.dims(l)
        out.append(dict(dims=d, ranks=r, near=level_list(*near, l), far=level_list(*far, l), A=A.download(l, 0),
                        Q=A.download(l, 1), R=A.download(l, 2), C=A.download(l, 3)))
    return out


@pytest.mark.timeout(300)
@pytest.mark.parametrize("geometry,n,leaf,kernel,alpha", CASES, ids=["cube4096-leaf256", "sphere3000-leaf128",
                                                                        "cube20000-leaf312"])
def test_symmetric_near_products_equal_general_ones(ctx, geometry, n, leaf, kernel, alpha):
    theta = 2.0
    pts = generate_points(geometry, n, 7)
    prob = Problem(ctx, pts, leaf)
    A = H2.operator(prob, 1e-8, 2000, kernel, 1e3, alpha, theta)
    G = HostLevels(prob, downloaded_levels(A, prob, pts, theta), kernel, 1e3, alpha, factorable=False)
    X = np.asfortranarray(np.random.default_rng(3).uniform(-1, 1, size=(n, 11)))
    dX = ctx.array(X)
    for prec, tol in ((N.MXG_FP64, 1e-14), (N.MXG_FP32, 2e-6)):
        dS, dG = ctx.empty((n, 11)), ctx.empty((n, 11))
        N.check_gpu(N.gpu().mxg_h2_matvec(N.h2().mxc_solver_device(A.handle), dX.ptr, dS.ptr, 11, n, prec))
        G.matvec(dX, dG, prec)
        S, Gm = dS.download(), dG.download()
        err = O.rel_err(S.ravel(), Gm.ravel())
        print(f"{geometry} N={n} leaf {leaf} precision {prec}: symmetric vs general products {err:.2e}")
        assert err < tol
        # each column of the batch of 11 (launches of 8, 2 and 1) equals its single product
        for q in (0, 7, 8, 10):
            d1, y1 = ctx.array(X[:, q].copy()), ctx.empty(n)
            N.check_gpu(N.gpu().mxg_h2_matvec(N.h2().mxc_solver_device(A.handle), d1.ptr, y1.ptr, 1, n, prec))
            assert np.array_equal(y1.download(), S[:, q]), q
    for o in (G, A, prob):
        o.close()
