"""The device factorization of level data handed over from the host
(mxg_h2_define_level + mxg_h2_upload, the counterpart of the reference's
copyDataInMatrixDesc, device_precondition.cu:198-202, as used by
H2MatrixSolver::factorizeDeviceM, solver.cpp:102-119).

The oracle's own preconditioner (bases, near blocks, couplings) is uploaded,
so both sides factorize identical data: the GPU substitution must equal the
oracle's to rounding, and on two near leaves (L = 1) it must equal the closed
form U A'^-1 U^T that pins the single Schur complement (h2matrix.cpp:346-425)."""
import numpy as np
import pytest

import oracle_lib as O
from mxc_amd import _native as N
from mxc_amd.solver import HostLevels, Problem

pytestmark = pytest.mark.gpu


def oracle_levels(oM):
    """Level structure and data of an oracle H^2 matrix (before factorization)."""
    out = []
    for l in range(oM.prob.levels + 1):
        d, r = oM.dims(l)
        out.append(dict(dims=d, ranks=r, near=oM.level_list(l, False), far=oM.level_list(l, True),
                        A=oM.download(l, 0), Q=oM.download(l, 1), R=oM.download(l, 2), C=oM.download(l, 3)))
    return out


def kernel_dense(kind, bodies, diag=1e3, alpha=1.0):
    p = bodies.reshape(-1, 3)
    d = np.linalg.norm(p[:, None] - p[None, :], axis=-1)
    off = ~np.eye(len(p), dtype=bool)
    K = np.full_like(d, diag)
    K[off] = (np.exp(-alpha * d[off]) if kind == "yukawa" else 1.0) / d[off]
    return K


@pytest.fixture(scope="module")
def uploaded(ctx):
    n, leaf, rank, samp = 4096, 256, 100, 2000
    pts = O.cube_points(n, 0)
    prob = Problem(ctx, pts, leaf)
    oprob = O.Problem(pts, leaf)
    assert np.array_equal(prob.bodies, oprob.bodies)
    oM = oprob.h2(rank=rank, orth=True, hss_far=True, sampling=samp)
    levels = oracle_levels(oM)
    G = HostLevels(prob, levels)
    return dict(ctx=ctx, prob=prob, oprob=oprob, oM=oM, levels=levels, G=G, n=n)


def test_upload_round_trips(uploaded):
    G, levels = uploaded["G"], uploaded["levels"]
    for l, lv in enumerate(levels):
        for key, what in (("A", N.H2_NEAR), ("Q", N.H2_BASIS), ("R", N.H2_R), ("C", N.H2_COUPLING)):
            assert np.array_equal(G.download(l, what), lv[key]), (l, key)


def test_uploaded_matvec_matches_oracle(uploaded):
    ctx, G, oM, n = uploaded["ctx"], uploaded["G"], uploaded["oM"], uploaded["n"]
    x = O.uniform_vec(n, -1, 1, 3)
    dx, dy = ctx.array(x), ctx.empty(n)
    G.matvec(dx, dy)
    assert O.rel_err(dy.download(), oM.matvec(x)) < 1e-14


def test_uploaded_factorization_matches_oracle(uploaded):
    """Identical level data on both sides: the substitutions agree to rounding
    (FP64; the only differences are summation orders and the explicit rr
    inverse of the device against the oracle's LU solves)."""
    ctx, G, oM, n = uploaded["ctx"], uploaded["G"], uploaded["oM"], uploaded["n"]
    assert not G.factorize(N.MXG_FP64).any()
    assert oM.factorize() == 0
    for seed in (1, 2):
        b = O.uniform_vec(n, -1, 1, seed)
        d = ctx.array(b)
        G.precondition(d)
        err = O.rel_err(d.download(), oM.precondition(b))
        print(f"uploaded FP64 factorization vs oracle, seed {seed}: {err:.3e}")
        assert err < 1e-12


@pytest.mark.parametrize("n,leaf,kernel,tol,rank", [(256, 128, "laplace", 0.0, 24), (257, 128, "yukawa", 1e-4, 60)])
def test_two_near_boxes_closed_form(ctx, n, leaf, kernel, tol, rank):
    pts = O.cube_points(n, 11)
    prob = Problem(ctx, pts, leaf)
    oprob = O.Problem(pts, leaf)
    assert prob.levels == 1 and np.array_equal(prob.bodies, oprob.bodies)
    oM = oprob.h2(kernel=kernel, alpha=1.0, tol=tol, rank=rank, orth=True, hss_far=True, sampling=0)
    assert O.lib().orc_h2_near_count(oM.h, 1) == 4
    levels = oracle_levels(oM)
    d, s = levels[1]["dims"], levels[1]["ranks"]
    assert np.all(s > 0) and np.all(s < d)
    U = np.zeros((n, n))
    o = oq = 0
    q = levels[1]["Q"]
    for m in map(int, d):
        U[o:o + m, o:o + m] = q[oq:oq + m * m].reshape(m, m, order="F")
        o += m
        oq += m * m
    Ah = U.T @ kernel_dense(kernel, prob.bodies) @ U
    r0, r1 = np.arange(s[0], d[0]), np.arange(d[0] + s[1], n)
    Ah[np.ix_(r0, r1)] = 0.
    Ah[np.ix_(r1, r0)] = 0.
    b = O.uniform_vec(n, -1, 1, 5)
    expect = U @ np.linalg.solve(Ah, U.T @ b)
    G = HostLevels(prob, levels, kernel=kernel)
    for prec, tol_ in ((N.MXG_FP64, 1e-12), (N.MXG_FP32, 1e-5)):
        assert not G.factorize(prec).any()
        dv = ctx.array(b)
        G.precondition(dv)
        err = O.rel_err(dv.download(), expect)
        print(f"L=1 {kernel} n={n} precision {prec}: {err:.3e}")
        assert err < tol_
    G.close()
    prob.close()
