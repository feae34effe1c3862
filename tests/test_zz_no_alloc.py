"""The allocation-free dense check (mxg_dense_matvec_batch_noalloc) against
the allocating one (mxg_dense_matvec_batch) through the C ABI: equal bit for
bit on two calls in a row, the second smaller than the first. The first call
needs several row slabs of the 64 MiB scratch (N = 300000 and 8 columns give
37 column chunks, so 28160 rows per slab: 11 slabs), the second one slab.
The allocation-free call must also leave free device memory unchanged."""
import ctypes as C

import numpy as np
import pytest

from mxc_amd import _native as N
from mxc_amd.device import kernel_desc

pytestmark = pytest.mark.gpu


def points(n, seed):
    return np.random.default_rng(seed).uniform(0.0, 1.0, size=(n, 3))


def both(ctx, kern, pts, X):
    n, k = X.shape
    G = N.gpu()
    db = ctx.array(np.ascontiguousarray(pts).reshape(-1))
    dX = ctx.array(X)
    dY0, dY1 = ctx.zeros((n, k)), ctx.zeros((n, k))
    N.check_gpu(G.mxg_dense_matvec_batch(ctx.handle, kern, db.ptr, n, dX.ptr, n, dY0.ptr, n, k))
    free0, _ = ctx.mem_info()
    N.check_gpu(G.mxg_dense_matvec_batch_noalloc(ctx.handle, kern, db.ptr, n, dX.ptr, n, dY1.ptr, n, k))
    free1, _ = ctx.mem_info()
    out = dY0.download(), dY1.download(), free0 - free1
    for a in (db, dX, dY0, dY1):
        a.free()
    return out


@pytest.mark.timeout(300)
def test_noalloc_dense_check_equals_allocating_one(ctx):
    kern = kernel_desc("yukawa", 1e3, 1.0)
    # first call: N = 300000, 8 columns (11 row slabs)
    n1 = 300000
    p1 = points(n1, 11)
    X1 = np.asfortranarray(np.random.default_rng(12).uniform(-1, 1, size=(n1, 8)))
    print(f"noalloc dense check: N={n1}, 8 columns", flush=True)
    Y0, Y1, used = both(ctx, kern, p1, X1)
    assert np.array_equal(Y0, Y1)
    assert used == 0, used
    # spot check against a host evaluation of a few rows
    d = np.linalg.norm(p1[:5, None, :] - p1[None, :, :], axis=2)
    with np.errstate(divide="ignore"):
        K = np.where(d > 0, np.exp(-d) / d, 1e3)
    assert np.allclose(K @ X1, Y0[:5], rtol=1e-11, atol=0)
    # second call: smaller, 3 columns (one slab, groups of 2 and 1)
    n2 = 5000
    p2 = points(n2, 13)
    X2 = np.asfortranarray(np.random.default_rng(14).uniform(-1, 1, size=(n2, 3)))
    print(f"noalloc dense check: N={n2}, 3 columns", flush=True)
    Z0, Z1, used = both(ctx, kern, p2, X2)
    assert np.array_equal(Z0, Z1)
    assert used == 0, used


def test_noalloc_dense_check_rejects_sizes_beyond_its_scratch(ctx):
    kern = kernel_desc("laplace", 1e3, 0.0)
    G = N.gpu()
    one = ctx.zeros(8)
    n = (1 << 28) + 1
    rc = G.mxg_dense_matvec_batch_noalloc(ctx.handle, kern, one.ptr, n, one.ptr, n, one.ptr, n, 1)
    assert rc == N.MXG_ERR_ARG
    assert b"exceeds" in C.string_at(G.mxg_last_error())
    one.free()
