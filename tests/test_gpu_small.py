"""Problems of one or two boxes (N = 200 with leaf 256: a tree of one box;
N = 600: two leaves of 300 that are near each other). Nothing is compressed,
so the operator is the dense kernel matrix and the ULV factorization is an
exact solve. Checks against closed forms computed here with numpy: the
matvec equals the dense product, the substitution equals
numpy.linalg.solve, and GMRES stops after one inner step at the dense
solution."""
import numpy as np
import pytest

import oracle_lib as O
from mxc_amd import _native as N
from mxc_amd.solver import H2, Problem

pytestmark = pytest.mark.gpu


def dense_laplace(p, diag):
    x = p.reshape(-1, 3)
    d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    with np.errstate(divide="ignore"):
        K = 1.0 / d
    np.fill_diagonal(K, diag)
    return K


@pytest.mark.timeout(120)
@pytest.mark.parametrize("n", [200, 600])
def test_one_or_two_boxes_are_dense(ctx, n):
    prob = Problem(ctx, O.cube_points(n, 7), 256)
    assert prob.levels == (0 if n < 512 else 1)
    K = dense_laplace(prob.bodies, 1e3)
    A = H2.operator(prob, 1e-10, 2000)
    # a rank cap above the box sizes: the preconditioner's bases (built from
    # the sibling's bodies) keep every column, so its ULV is the dense solve
    M = H2.preconditioner(prob, 512, 2000)
    for l in range(prob.levels + 1):
        _, r = A.dims(l)
        assert not r.any(), "no box has a far field, so no box has a skeleton"
    x = O.uniform_vec(n, -1, 1, 3)
    dx, dy = ctx.array(x), ctx.empty(n)
    A.matvec(dx, dy)
    assert np.linalg.norm(dy.download() - K @ x) <= 1e-14 * np.linalg.norm(K @ x)
    b = K @ x
    xs = np.linalg.solve(K, b)
    for prec, tol in ((N.MXG_FP64, 1e-12), (N.MXG_FP32, 1e-5)):
        M.factorize(prec)
        db = ctx.array(b)
        M.precondition(db)
        assert np.linalg.norm(db.download() - xs) <= tol * np.linalg.norm(xs), prec
    M.factorize(N.MXG_FP64)
    db, dsol = ctx.array(b), ctx.zeros(n)
    res, it, inner = A.gmres(M, db, dsol, 1e-10, restart=10, max_cycles=5, mode=N.GMRES_CGS2)
    assert res[-1] < 1e-10 and it == 1 and inner <= 2
    assert np.linalg.norm(dsol.download() - xs) <= 1e-12 * np.linalg.norm(xs)
    for a in (dx, dy, db, dsol):
        a.free()
