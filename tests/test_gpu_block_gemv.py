"""The solve phase's block product by itself (mxg_block_gemv), against a host
FP64 product of the stored block: bfloat16-stored blocks (the substitution's
factors) of 256 x 256 (the one-wave path, two staged column chunks at 8
right-hand sides), 93 x 94 and 94 x 93 (ragged leaves) and 300 x 512 (the
workgroup path), for 1, 2, 3, 4 and 8 right-hand sides; FP32 and FP64 blocks
of the same shapes, on the vector units and on the matrix cores (the aligned
FP32 and FP64 blocks; bfloat16 blocks stay on the vector units). Each column
of a batch also equals its single-column product bit for bit."""
import numpy as np
import pytest

from mxc_amd import _native as N

pytestmark = pytest.mark.gpu

SHAPES = [(256, 256, 256), (93, 94, 93), (94, 93, 97), (300, 512, 304)]  # n, m, ldb


def to_bf16(a):
    """float32 -> bfloat16 bit patterns, round to nearest even."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16(u):
    return (u.astype(np.uint32) << 16).view(np.float32)


def run(ctx, storage, Bst, X, n, m, ldb, nrhs, units):
    G = N.gpu()
    dB = ctx.array(Bst)
    dX = ctx.array(X)
    dY = ctx.zeros((n, nrhs), dtype=X.dtype)
    N.check_gpu(G.mxg_block_gemv(ctx.handle, storage, n, m, dB.ptr, ldb, dX.ptr, m, dY.ptr, n, nrhs, units))
    Y = dY.download()
    for a in (dB, dX, dY):
        a.free()
    return Y


@pytest.mark.timeout(120)
@pytest.mark.parametrize("n,m,ldb", SHAPES, ids=[f"{n}x{m}" for n, m, _ in SHAPES])
@pytest.mark.parametrize("storage", ["bf16", "fp32", "fp64"])
@pytest.mark.parametrize("units", [0, 1], ids=["vector", "matrix"])
def test_block_product_matches_host(ctx, n, m, ldb, storage, units):
    rng = np.random.default_rng(n * 1000 + m)
    full = rng.standard_normal((ldb, m))
    for nrhs in (1, 2, 3, 4, 8):
        X64 = rng.standard_normal((m, nrhs))
        if storage == "bf16":
            Bst = np.asfortranarray(to_bf16(full.astype(np.float32)))
            Bval = from_bf16(np.asarray(Bst)).astype(np.float64)[:n]
            X = np.asfortranarray(X64.astype(np.float32))
            code, tol = N.MXG_BF16, 5e-6
        elif storage == "fp32":
            Bst = np.asfortranarray(full.astype(np.float32))
            Bval = Bst.astype(np.float64)[:n]
            X = np.asfortranarray(X64.astype(np.float32))
            code, tol = N.MXG_FP32, 5e-6
        else:
            Bst = np.asfortranarray(full)
            Bval = Bst[:n]
            X = np.asfortranarray(X64)
            code, tol = N.MXG_FP64, 1e-14
        Y = run(ctx, code, Bst, X, n, m, ldb, nrhs, units)
        ref = Bval @ X.astype(np.float64)
        err = np.linalg.norm(Y.astype(np.float64) - ref) / np.linalg.norm(ref)
        assert err < tol, (storage, n, m, nrhs, err)
        if nrhs > 1:
            for q in (0, nrhs - 1):
                y1 = run(ctx, code, Bst, np.asfortranarray(X[:, q:q + 1]), n, m, ldb, 1, units)
                assert np.array_equal(y1[:, 0], Y[:, q]), (storage, n, m, nrhs, q)


def test_block_product_rejects_bad_arguments(ctx):
    G = N.gpu()
    one = ctx.zeros(16)
    assert G.mxg_block_gemv(ctx.handle, N.MXG_FP64, 4, 4, one.ptr, 3, one.ptr, 4, one.ptr, 4, 1, 0) == N.MXG_ERR_ARG
    assert G.mxg_block_gemv(ctx.handle, N.MXG_FP64, 4, 4, one.ptr, 4, one.ptr, 4, one.ptr, 4, 9, 0) == N.MXG_ERR_ARG
    assert G.mxg_block_gemv(ctx.handle, 7, 4, 4, one.ptr, 4, one.ptr, 4, one.ptr, 4, 1, 0) == N.MXG_ERR_ARG
    one.free()
