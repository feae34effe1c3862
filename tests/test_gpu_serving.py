"""Serving: resident factorized problems on one GPU, HBM-budget admission,
continuous batching of arriving right-hand sides, work of different problems
overlapping on the device (one context and stream per problem).

Every served solution meets 1e-10 on the H^2 residual and on the dense FP64
check; each request's solution in the mix (its columns sharing a lockstep
GMRES call with other requests', while other problems run beside it) equals
bit for bit the solution of the same request solved alone; a budget smaller
than the working set evicts idle problems and rebuilds them on demand."""
import numpy as np
import pytest

import oracle_lib as O
from mxc_amd import _native as N
from mxc_amd.serving import ProblemSpec, Server, latency_stats, poisson_load

pytestmark = pytest.mark.gpu

SIZES = [16384, 32768, 65536, 16384, 32768, 65536]


def specs():
    return {pid: ProblemSpec(n=n, seed=300 + pid) for pid, n in enumerate(SIZES)}


def dense_resid(srv, pid, B, X):
    res = srv.resident[pid]
    ctx = res.ctx
    dX, dY = ctx.array(X), ctx.empty(X.shape)
    s = res.spec
    res.problem.dense_matvec(s.kernel, s.diag, s.alpha, dX, dY)
    Y = dY.download()
    dX.free()
    dY.free()
    return max(float(np.linalg.norm(B[:, q] - Y[:, q]) / np.linalg.norm(B[:, q])) for q in range(B.shape[1]))


@pytest.mark.timeout(900)
def test_mix_equals_alone_and_meets_tolerance(ctx):
    free, total = ctx.mem_info()
    srv = Server(hbm_budget=int(0.5 * free), slots=3, max_rhs=32)
    try:
        sp = specs()
        for pid, s in sp.items():
            srv.register(pid, s)
        reqs = poisson_load(srv, list(sp), [s.n for s in sp.values()], rate=40.0, count=48, seed=9,
                            max_rhs_per_request=12)
        st = latency_stats(reqs, srv.calls)
        print(f"mix: {st}")
        print(f"server: {srv.state()}", flush=True)
        assert srv.evictions == 0, srv.state()
        assert any(c[2] > 1 for c in srv.calls)  # calls carried several requests
        for r in reqs:
            assert max(r.resid) < srv.h2_tol, (r.pid, r.resid)
        # the same requests, one at a time
        for i, r in enumerate(reqs[:16]):
            alone = srv.submit(r.pid, r.B).wait(120)
            d = float(np.max(np.abs(alone - r.X)))
            assert np.array_equal(alone, r.X), (i, r.pid, r.nrhs, d)
        print("alone == mix for 16 requests", flush=True)
        for pid in sp:
            mine = [r for r in reqs if r.pid == pid][:2]
            for r in mine:
                rr = dense_resid(srv, pid, r.B, r.X)
                assert rr < 1e-10, (pid, rr)
    finally:
        srv.close()


@pytest.mark.timeout(900)
def test_budget_evicts_and_readmits(ctx):
    sp = specs()
    srv = Server(hbm_budget=1 << 62, slots=1, max_rhs=8)
    try:
        srv.register(0, sp[2])  # the largest problem sets the scale of the budget
        srv.submit(0, np.ones((sp[2].n, 1))).wait(120)
        big = srv.resident[0].footprint
        print(f"footprint of N={sp[2].n}: {big / 2**30:.2f} GiB", flush=True)
    finally:
        srv.close()
    srv = Server(hbm_budget=int(2.2 * big), slots=2, max_rhs=8)
    try:
        for pid, s in sp.items():
            srv.register(pid, s)
        rng = np.random.default_rng(3)
        reqs = []
        for rnd in range(2):
            for pid, s in sp.items():
                reqs.append(srv.submit(pid, rng.uniform(-1, 1, size=(s.n, 2))))
            for r in reqs:
                try:
                    r.wait(120)
                except TimeoutError as e:
                    raise TimeoutError(f"{e}: {srv.state()}") from None
            print(f"round {rnd}: {srv.state()}", flush=True)
            assert srv.resident_bytes() <= srv.budget
        print(f"budget {srv.budget / 2**30:.2f} GiB, footprint of N=65536 {big / 2**30:.2f} GiB, "
              f"evictions {srv.evictions}")
        assert srv.evictions > 0
        for r in reqs:
            assert max(r.resid) < 1e-10
    finally:
        srv.close()
