"""GPU parity of the building blocks: kernel-block generation, the dense
check and the batched MFMA GEMM, through the C ABI of libmxc_gpu."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from mxc_amd import _native as N
from mxc_amd.device import kernel_desc

pytestmark = pytest.mark.gpu


def _gen_device(ctx, kernel, diag, alpha, bi, bj):
    m, n = len(bi) // 3, len(bj) // 3
    dbi, dbj = ctx.array(bi), ctx.array(bj)
    dA = ctx.empty((m, n))
    N.check_gpu(N.gpu().mxg_gen_block(ctx.handle, kernel_desc(kernel, diag, alpha), m, n, dbi.ptr, dbj.ptr, dA.ptr, m))
    return dA.download()


def _truth(kernel, diag, alpha, bi, bj):
    """K(|bi - bj|) in x87 extended precision (64-bit mantissa)."""
    p = bi.reshape(-1, 3).astype(np.longdouble)
    q = bj.reshape(-1, 3).astype(np.longdouble)
    d2 = ((p[:, None, :] - q[None, :, :]) ** 2).sum(-1)
    d = np.sqrt(d2)
    with np.errstate(divide="ignore", invalid="ignore"):
        if kernel == "laplace":
            k = np.where(d2 == 0, diag, 1 / d)
        elif kernel == "yukawa":
            k = np.where(d2 == 0, diag, np.exp(-alpha * d) / d)
        else:
            k = np.where(d2 == 0, 1.0001, np.exp(-alpha * d2))
    return k


# Entries differ from the reference's gen_matrix only in rounding: the
# reference forms the distance with the scaled 3-argument hypot, the device
# with sqrt(dx^2 + dy^2 + dz^2) (correctly rounded sqrt and divide) and the
# device libm exp. Both must sit within a few ulp of the extended-precision
# value, and within a few ulp of each other.
@pytest.mark.parametrize("kernel,alpha,ulps", [("laplace", 0.0, 3), ("yukawa", 1.0, 6), ("gaussian", 0.5, 6)])
def test_gen_block_matches_oracle(ctx, kernel, alpha, ulps):
    rng = np.random.default_rng(3)
    bi = rng.uniform(0, 1, 3 * 97)
    bj = np.concatenate([bi[: 3 * 20], rng.uniform(0, 1, 3 * 53)])  # shared points give K(0) = diag
    A = _gen_device(ctx, kernel, 1e3, alpha, bi, bj)
    ref = np.zeros((97, 73), order="F")
    O.lib().orc_gen_matrix(O.KID[kernel], 1e3, alpha, 97, 73, O.ptr(bi), O.ptr(bj), ref.ctypes.data)
    truth = _truth(kernel, 1e3, alpha, bi, bj)
    ulp = np.spacing(np.abs(ref))
    dev_err = float(np.max(np.abs(A - truth).astype(np.float64) / ulp))
    orc_err = float(np.max(np.abs(ref - truth).astype(np.float64) / ulp))
    pair = float(np.max(np.abs(A - ref) / ulp))
    assert dev_err <= ulps, f"device {dev_err:.2f} ulp from extended precision"
    assert orc_err <= ulps, f"oracle {orc_err:.2f} ulp from extended precision"
    assert pair <= 2 * ulps, f"device vs oracle {pair:.2f} ulp"
    if kernel != "gaussian":
        assert all(A[t, t] == 1e3 for t in range(20))


def test_gen_block_empty_and_bad_args(ctx):
    g = N.gpu()
    assert g.mxg_gen_block(ctx.handle, kernel_desc("laplace"), 0, 5, None, None, None, 1) == 0
    assert g.mxg_gen_block(ctx.handle, kernel_desc("laplace"), 4, 4, None, None, None, 2) != 0  # lda < m


@pytest.mark.parametrize("kernel", ["laplace", "yukawa"])
def test_dense_matvec_matches_oracle(ctx, kernel):
    n = 3000
    pts = O.cube_points(n, 5)
    x = O.uniform_vec(n, -1, 1, 9)
    y = np.zeros(n)
    dp, dx, dy = ctx.array(pts), ctx.array(x), ctx.empty(n)
    N.check_gpu(N.gpu().mxg_dense_matvec(ctx.handle, kernel_desc(kernel, 1e3, 1.0), dp.ptr, n, dx.ptr, dy.ptr))
    y = dy.download()
    ref = O.dense_matvec(kernel, 1e3, 1.0, pts, x)
    assert O.rel_err(y, ref) < 1e-14


@pytest.mark.parametrize("prec,dtype,tol", [(0, np.float64, 1e-13), (1, np.float32, 2e-5)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (70, 45, 33), (130, 129, 65), (64, 64, 16), (5, 200, 300)])
def test_gemm_matches_numpy(ctx, prec, dtype, tol, ta, tb, m, n, k):
    rng = np.random.default_rng(m * 1000 + n * 10 + k + 7 * ta + 3 * tb)
    A = rng.standard_normal((k, m) if ta else (m, k)).astype(dtype)
    B = rng.standard_normal((n, k) if tb else (k, n)).astype(dtype)
    Cm = rng.standard_normal((m, n)).astype(dtype)
    alpha, beta = 0.75, -1.25
    dA, dB, dC = ctx.array(np.asfortranarray(A)), ctx.array(np.asfortranarray(B)), ctx.array(np.asfortranarray(Cm))
    N.check_gpu(N.gpu().mxg_gemm(ctx.handle, prec, ta, tb, m, n, k, alpha, dA.ptr, A.shape[0], dB.ptr, B.shape[0], beta,
                                 dC.ptr, m))
    got = dC.download().astype(np.float64)
    opA = A.T if ta else A
    opB = B.T if tb else B
    ref = alpha * (opA.astype(np.float64) @ opB.astype(np.float64)) + beta * Cm.astype(np.float64)
    scale = np.abs(alpha) * (np.abs(opA) @ np.abs(opB)) + abs(beta) * np.abs(Cm)
    assert np.max(np.abs(got - ref) / scale) < tol


def test_gemm_beta_zero_ignores_nan(ctx):
    m = n = k = 40
    A = np.eye(m)
    Cm = np.full((m, n), np.nan)
    dA, dC = ctx.array(A), ctx.array(Cm)
    N.check_gpu(N.gpu().mxg_gemm(ctx.handle, 0, 0, 0, m, n, k, 1.0, dA.ptr, m, dA.ptr, m, 0.0, dC.ptr, m))
    assert np.array_equal(dC.download(), np.eye(m))


def bf16_round(a):
    """float32 values rounded to bfloat16 (nearest even), as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (70, 45, 33), (130, 129, 65), (64, 64, 16), (5, 200, 300)])
def test_gemm_bf16_inputs_match_numpy(ctx, ta, tb, m, n, k):
    """FP32 matrices multiplied on the bfloat16 matrix cores: equal to the
    FP64 product of the bfloat16-rounded inputs up to FP32 accumulation."""
    rng = np.random.default_rng(m * 1000 + n * 10 + k + 7 * ta + 3 * tb + 1)
    A = rng.standard_normal((k, m) if ta else (m, k)).astype(np.float32)
    B = rng.standard_normal((n, k) if tb else (k, n)).astype(np.float32)
    Cm = rng.standard_normal((m, n)).astype(np.float32)
    alpha, beta = 0.75, -1.25
    dA, dB, dC = ctx.array(np.asfortranarray(A)), ctx.array(np.asfortranarray(B)), ctx.array(np.asfortranarray(Cm))
    N.check_gpu(N.gpu().mxg_gemm(ctx.handle, N.MXG_BF16, ta, tb, m, n, k, alpha, dA.ptr, A.shape[0], dB.ptr,
                                 B.shape[0], beta, dC.ptr, m))
    got = dC.download().astype(np.float64)
    opA = bf16_round(A.T if ta else A)
    opB = bf16_round(B.T if tb else B)
    ref = alpha * (opA @ opB) + beta * Cm.astype(np.float64)
    scale = np.abs(alpha) * (np.abs(opA) @ np.abs(opB)) + abs(beta) * np.abs(Cm)
    assert np.max(np.abs(got - ref) / scale) < 2e-5
    # and it is not the FP32 product: the inputs were rounded
    if k >= 16:
        full = alpha * (np.asarray(A.T if ta else A, np.float64) @ np.asarray(B.T if tb else B, np.float64)) \
            + beta * Cm.astype(np.float64)
        assert np.max(np.abs(got - full) / scale) > 1e-4
