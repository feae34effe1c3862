"""Kernel timing classes (mxg_profile_*): the launches each class counts for one
construction, one matvec and one substitution, the sketch class's flop count,
and that reset clears and a disabled context records nothing."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from mxc_amd import _native as N
from mxc_amd.solver import H2, Problem

pytestmark = pytest.mark.gpu

MATVEC_LEAF_FP64, MATVEC_LEAF_FP32, MATVEC_FP64, MATVEC_FP32, ULV_SOLVE = 0, 1, 6, 7, 8


def read(ctx, cls):
    cnt, ms, work = C.c_longlong(), C.c_double(), C.c_double()
    N.check_gpu(N.gpu().mxg_profile_read(ctx.handle, cls, C.byref(cnt), C.byref(ms)))
    N.check_gpu(N.gpu().mxg_profile_work(ctx.handle, cls, C.byref(work)))
    return cnt.value, ms.value, work.value


@pytest.mark.timeout(120)
def test_profile_classes_count_their_launches(ctx):
    n, leaf = 4096, 256
    prob = Problem(ctx, O.cube_points(n, 0), leaf)
    G = N.gpu()
    N.check_gpu(G.mxg_profile_reset(ctx.handle))
    N.check_gpu(G.mxg_profile_enable(ctx.handle, 1))
    A = H2.operator(prob, 1e-10, 2000)
    sk_launches, sk_ms, sk_work = read(ctx, N.MXG_PROF_SKETCH)
    # boxes with more far rows than columns are sketched, in at most one
    # launch per level; each counts 2 M^2 nf flops with M < nf < n
    assert 1 <= sk_launches <= prob.levels + 1
    assert sk_ms > 0
    m2 = [A.dims(l)[0].astype(np.float64) ** 2 for l in range(prob.levels + 1)]
    assert 2.0 * min(float(np.min(m * np.sqrt(m))) for m in m2 if m.size) <= sk_work
    assert sk_work <= 2.0 * n * sum(float(np.sum(m)) for m in m2)
    # the Householder row reduction takes no sketch
    H2.operator(prob, 1e-10, 2000, householder_rows=True)
    assert read(ctx, N.MXG_PROF_SKETCH)[0] == sk_launches

    x = ctx.array(np.random.default_rng(0).standard_normal(n))
    y = ctx.empty(n)
    A.matvec(x, y, N.MXG_FP64)
    assert read(ctx, MATVEC_FP64)[0] == 1 and read(ctx, MATVEC_LEAF_FP64)[0] == 1
    assert read(ctx, MATVEC_FP32)[0] == 0 and read(ctx, MATVEC_LEAF_FP32)[0] == 0
    A.matvec(x, y, N.MXG_FP32)
    assert read(ctx, MATVEC_FP32)[0] == 1 and read(ctx, MATVEC_LEAF_FP32)[0] == 1
    M = H2.preconditioner(prob, 64, 2000)
    M.factorize(N.MXG_FP64)
    M.precondition(y)
    assert read(ctx, ULV_SOLVE)[0] == 1

    N.check_gpu(G.mxg_profile_reset(ctx.handle))
    for cls in range(N.MXG_PROF_SKETCH + 1):
        assert read(ctx, cls) == (0, 0.0, 0.0)
    N.check_gpu(G.mxg_profile_enable(ctx.handle, 0))
    A.matvec(x, y, N.MXG_FP64)
    assert read(ctx, MATVEC_FP64)[0] == 0
    for a in (x, y):
        a.free()
