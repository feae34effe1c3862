"""Variable-size batching: 64 independent problems (N log-uniform in
[4096, 65536], theta in {1, 2, 3}, leaf 128 or 256, rank in {32, 64, 128},
operator eps 1e-8, Laplace with K(0) = 1e3, points uniform in [0,1]^3) whose
preconditioners are factorized in FP32 in shared launches
(mxg_ulv_create_batch), then solved with FP64 GMRES.

Every problem's batched factorization equals its factorization alone bit for
bit; every solve meets its tolerance on the dense FP64 check; the smallest
problems' substitution errors are within 10x of the oracle's (FP64) on the
same points."""
import numpy as np
import pytest

import oracle_lib as O
from mxc_amd import _native as N
from mxc_amd.batch import Batch, log_uniform_specs

pytestmark = pytest.mark.gpu

COUNT, TOL = 64, 1e-8


@pytest.fixture(scope="module")
def batch(ctx):
    specs = log_uniform_specs(COUNT, seed=4)
    assert min(s.n for s in specs) >= 4096 and max(s.n for s in specs) <= 65536
    b = Batch(ctx, specs)
    infos = b.factorize(N.MXG_FP32)
    assert all(not i.any() for i in infos)
    print(f"{COUNT} problems, N {min(s.n for s in specs)}..{max(s.n for s in specs)}, total "
          f"{sum(s.n for s in specs)} points; batched FP32 factorization {b.factor_seconds() * 1e3:.1f} ms")
    yield b
    b.close()


def apply_precond(ctx, M, b):
    d = ctx.array(b)
    M.precondition(d)
    return d.download()


@pytest.mark.timeout(900)
def test_batched_factorization_equals_single(batch):
    ctx = batch.ctx
    order = sorted(range(COUNT), key=lambda k: batch.members[k].spec.n)
    picks = order[:3] + order[-3:] + order[COUNT // 2 - 1:COUNT // 2 + 1]
    for k in picks:
        m = batch.members[k]
        b = O.uniform_vec(m.spec.n, -1, 1, 17 + k)
        zb = apply_precond(ctx, m.M, b)
        assert not m.M.factorize(N.MXG_FP32).any()
        zs = apply_precond(ctx, m.M, b)
        assert np.array_equal(zb, zs), (k, m.spec)
    # restore the batch's factorizations for the tests that follow
    batch.factorize(N.MXG_FP32)


@pytest.mark.timeout(900)
def test_every_problem_solves_to_tolerance(batch):
    ctx = batch.ctx
    worst, worst_u = 0., 0.
    for k, m in enumerate(batch.members):
        n = m.spec.n
        b = O.uniform_vec(n, -1, 1, 100 + k)
        db, dx = ctx.array(b), ctx.zeros(n)
        res, cycles, inner = m.A.gmres(m.M, db, dx, TOL, restart=30, max_cycles=20, mode=N.GMRES_CGS2)
        assert res[-1] < TOL, (k, m.spec, res)
        dax = ctx.empty(n)
        m.problem.dense_matvec("laplace", 1e3, 0.0, dx, dax)
        ax = dax.download()
        rr = O.rel_err(ax, b)
        # the dense residual is the H^2 residual plus the operator's
        # construction error on the solution
        dh = ctx.empty(n)
        m.A.matvec(dx, dh)
        cerr = O.rel_err(dh.download(), ax)
        assert rr <= 1.01 * (res[-1] + cerr) + 1e-14, (k, m.spec, rr, cerr)
        # construction error on x ~ U(0,1), the reference's own check
        # (main.cpp:104-117); the oracle gives 2.5e-9..1.5e-8 on these specs
        xu = O.uniform_vec(n, 0, 1, 5489)
        dxu, dbu = ctx.array(xu), ctx.empty(n)
        m.problem.dense_matvec("laplace", 1e3, 0.0, dxu, dbu)
        m.A.matvec(dxu, dh)
        cu = O.rel_err(dh.download(), dbu.download())
        assert cu < 1e-7, (k, m.spec, cu)
        worst, worst_u = max(worst, rr), max(worst_u, cu)
        for a in (db, dx, dax, dh, dxu, dbu):
            a.free()
    print(f"over {COUNT} problems: worst dense residual of the solution {worst:.3e}, worst construction error "
          f"{worst_u:.3e}")


@pytest.mark.timeout(900)
def test_smallest_problems_match_oracle(batch):
    ctx = batch.ctx
    order = sorted(range(COUNT), key=lambda k: batch.members[k].spec.n)
    for k in order[:3]:
        m = batch.members[k]
        s = m.spec
        pts = O.cube_points(s.n, s.seed)
        assert np.array_equal(m.problem.bodies, O.Problem(pts, s.leaf).bodies)
        oprob = O.Problem(pts, s.leaf)
        oM = oprob.h2(rank=s.rank, orth=True, hss_far=True, sampling=s.sampling, theta=s.theta, nthreads=16)
        assert oM.factorize() == 0
        x = O.uniform_vec(s.n, 0, 1, 5489)
        b = O.dense_matvec("laplace", 1e3, 0.0, oprob.bodies, x, 16)
        serr = O.rel_err(apply_precond(ctx, m.M, b), x)
        oserr = O.rel_err(oM.precondition(b), x)
        oA = oprob.h2(tol=s.eps, sampling=s.sampling, theta=s.theta, nthreads=16)
        ocerr = O.rel_err(oA.matvec(x), b)
        dxd, dh = ctx.array(x), ctx.empty(s.n)
        m.A.matvec(dxd, dh)
        cerr = O.rel_err(dh.download(), b)
        print(f"problem {k} {s}: substitution error gpu (FP32, batched) {serr:.4e} oracle (FP64) {oserr:.4e}; "
              f"construction error gpu {cerr:.3e} oracle {ocerr:.3e}")
        assert serr < 10 * oserr
        assert cerr < 10 * ocerr
