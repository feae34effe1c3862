"""HIP-graph replay of the GMRES inner steps (mxg_gmres_opts_t.graph): a call
captures its first inner step's operator and preconditioner launches (the
matvec's two streams included) into one graph and replays it for every later
step. Solutions, residual histories and iteration counts must equal those of
plain launches bit for bit, for one and for eleven right-hand sides (matvec
and substitution launches of 8, 2 and 1 columns), and over several calls."""
import numpy as np
import pytest

from mxc_amd import _native as N
from mxc_amd.solver import H2, Problem, generate_points, uniform_vector

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sphere(ctx):
    n = 6000
    prob = Problem(ctx, generate_points("sphere", n, 3), 128)
    kern = ("yukawa", 1e3, 1.0)
    A = H2.operator(prob, 1e-10, 2000, *kern)
    A_in = H2.operator(prob, 1e-7, 2000, *kern)
    M = H2.preconditioner(prob, 48, 2000, *kern, eps=1e-6)
    assert not M.factorize(N.MXG_BF16).any()
    yield dict(ctx=ctx, n=n, A=A, A_in=A_in, M=M)
    for o in (M, A_in, A, prob):
        o.close()


@pytest.mark.timeout(300)
@pytest.mark.parametrize("nrhs", [1, 11])
def test_graph_replay_equals_plain_launches(sphere, nrhs):
    ctx, n, A, M = sphere["ctx"], sphere["n"], sphere["A"], sphere["M"]
    B = np.stack([uniform_vector(n, -1.0, 1.0, 40 + q) for q in range(nrhs)], axis=1)
    dB = ctx.array(B)
    out = {}
    for graph in (False, True, True):
        dX = ctx.zeros((n, nrhs))
        res, cyc, inn = A.gmres_batch(M, dB, dX, 1e-10, 30, 10, N.GMRES_CGS2, N.MXG_FP32, sphere["A_in"], graph=graph)
        X = dX.download()
        assert max(r[-1] for r in res) < 1e-10
        if graph in out:
            assert np.array_equal(X, out[graph][0])  # a second graph call repeats the first
        out[graph] = (X, [r.tolist() for r in res], list(cyc), list(inn))
        print(f"nrhs {nrhs} graph {graph}: cycles {list(cyc)} inner {list(inn)}")
    assert max(out[False][3]) > 1  # the graph was replayed
    assert np.array_equal(out[True][0], out[False][0])
    assert out[True][1:] == out[False][1:]
