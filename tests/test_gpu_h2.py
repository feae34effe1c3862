"""GPU parity of the H^2 solve path against the CPU oracle on the same seeded
inputs: operator construction and matvec, ULV factorization and
substitution (FP64 and FP32), HSS exactness, and GMRES.

Tolerances: the GPU and the oracle run the same algorithm in different
summation orders, so column-pivot ties can resolve differently; results are
compared by the reference's own error measures (construct error against the
dense matvec, substitution error, dense residual of the solve), each within
10x of the oracle's, and directly where the algorithm is deterministic."""
import numpy as np
import pytest

import oracle_lib as O
from mxc_amd import _native as N
from mxc_amd.solver import H2, Problem

pytestmark = pytest.mark.gpu

NPTS, LEAF, EPS, RANK, SAMP = 4096, 256, 1e-10, 100, 2000


@pytest.fixture(scope="module")
def setup(ctx):
    pts = O.cube_points(NPTS, 0)
    prob = Problem(ctx, pts, LEAF)
    oprob = O.Problem(pts, LEAF)
    assert np.array_equal(prob.bodies, oprob.bodies)
    x = O.uniform_vec(NPTS, 0, 1, 0)
    bref = O.dense_matvec("laplace", 1e3, 0.0, prob.bodies, x)
    return dict(ctx=ctx, prob=prob, oprob=oprob, x=x, bref=bref)


@pytest.fixture(scope="module")
def operators(setup):
    A = H2.operator(setup["prob"], EPS, SAMP)
    oA = setup["oprob"].h2(tol=EPS, sampling=SAMP)
    return A, oA


@pytest.fixture(scope="module")
def preconditioners(setup):
    M = H2.preconditioner(setup["prob"], RANK, SAMP)
    oM = setup["oprob"].h2(rank=RANK, orth=True, hss_far=True, sampling=SAMP)
    return M, oM


def gpu_matvec(ctx, A, x, precision=N.MXG_FP64):
    dx, dy = ctx.array(x), ctx.empty(len(x))
    A.matvec(dx, dy, precision)
    return dy.download()


def test_structure_matches_oracle(setup, operators, preconditioners):
    for A, oA in (operators, preconditioners):
        for l in range(A.levels + 1):
            d, r = A.dims(l)
            od, orr = oA.dims(l)
            assert A.counts(l) == (O.lib().orc_h2_near_count(oA.h, l), O.lib().orc_h2_far_count(oA.h, l))
            if l == A.levels:
                assert np.array_equal(d, od)
            # tolerance ranks may differ by pivot ties; fixed ranks may not
            assert np.max(np.abs(r - orr)) <= (2 if A is operators[0] else 0)


def test_operator_construct_error(setup, operators):
    A, oA = operators
    y = gpu_matvec(setup["ctx"], A, setup["x"])
    oy = oA.matvec(setup["x"])
    err, oerr = O.rel_err(y, setup["bref"]), O.rel_err(oy, setup["bref"])
    assert oerr < 1e-10
    assert err < 10 * oerr + 1e-14, (err, oerr)
    assert O.rel_err(y, oy) < 10 * (err + oerr)


def test_operator_householder_rows_matches_oracle(setup):
    """The reference's row reduction (R factor of the far block) on both sides."""
    A = H2.operator(setup["prob"], EPS, SAMP, householder_rows=True)
    O.lib().orc_set_sketch(0)
    try:
        oA = setup["oprob"].h2(tol=EPS, sampling=SAMP)
    finally:
        O.lib().orc_set_sketch(1)
    for l in range(A.levels + 1):
        assert np.max(np.abs(A.dims(l)[1] - oA.dims(l)[1])) <= 2
    y = gpu_matvec(setup["ctx"], A, setup["x"])
    oy = oA.matvec(setup["x"])
    err, oerr = O.rel_err(y, setup["bref"]), O.rel_err(oy, setup["bref"])
    assert oerr < 1e-10
    assert err < 10 * oerr + 1e-14, (err, oerr)


def test_preconditioner_bases_match_oracle(setup, preconditioners):
    """Fixed-rank orthogonal bases are fixed by the sketch's pivots: equal sign
    patterns on both sides give the same basis columns up to sign."""
    M, oM = preconditioners
    L = M.levels
    ql = M.download(L, 1)
    d, r = M.dims(L)
    oq = oM.download(L, 1)
    off = 0
    for i in range(len(d)):
        m, k = int(d[i]), int(r[i])
        Q = ql[off:off + m * m].reshape(m, m, order="F")[:, :k]
        oQ = oq[off:off + m * m].reshape(m, m, order="F")[:, :k]
        s = np.sign(np.sum(Q * oQ, axis=0))
        assert np.max(np.abs(Q * s - oQ)) < 1e-8, i
        off += m * m


def test_operator_fp32_matvec(setup, operators):
    A, _ = operators
    y64 = gpu_matvec(setup["ctx"], A, setup["x"])
    y32 = gpu_matvec(setup["ctx"], A, setup["x"], N.MXG_FP32)
    assert 1e-9 < O.rel_err(y32, y64) < 1e-5


def test_preconditioner_matvec_matches_oracle(setup, preconditioners):
    M, oM = preconditioners
    y = gpu_matvec(setup["ctx"], M, setup["x"])
    oy = oM.matvec(setup["x"])
    assert O.rel_err(y, oy) < 1e-9
    assert O.rel_err(y, setup["bref"]) < 1e-2


def test_factorization_substitution_matches_oracle(setup, preconditioners):
    M, oM = preconditioners
    info = M.factorize(N.MXG_FP64)
    assert not info.any()
    assert oM.factorize() == 0
    ctx = setup["ctx"]
    d = ctx.array(setup["bref"])
    M.precondition(d)
    z = d.download()
    oz = oM.precondition(setup["bref"])
    serr, oserr = O.rel_err(z, setup["x"]), O.rel_err(oz, setup["x"])
    assert serr < 10 * oserr, (serr, oserr)
    assert O.rel_err(z, oz) < 1e-8
    # FP32 factorization of the same matrix: its rounding (unit roundoff 6e-8
    # times the conditioning of the rotated blocks) stays within the
    # preconditioner's own approximation error
    M.factorize(N.MXG_FP32)
    d = ctx.array(setup["bref"])
    M.precondition(d)
    z32 = d.download()
    assert O.rel_err(z32, z) < 2 * serr, (O.rel_err(z32, z), serr)
    assert O.rel_err(z32, setup["x"]) < 2 * serr
    # FP32 factorization with bfloat16-stored factors: the substitution
    # differs from the FP32 one by bfloat16 rounding (2^-9 relative per
    # entry) amplified by the conditioning of the rotated blocks
    M.factorize(N.MXG_BF16)
    d = ctx.array(setup["bref"])
    M.precondition(d)
    z16 = d.download()
    assert 1e-4 < O.rel_err(z16, z32) < 0.1, O.rel_err(z16, z32)
    assert O.rel_err(z16, setup["x"]) < 0.1
    M.factorize(N.MXG_FP64)


def test_refactorization_is_reproducible(setup, preconditioners):
    """Factorizations of one matrix in a row, in alternating precisions, reuse
    the context's staged task lists; each gives the same substitution bit for
    bit as the first one of its precision."""
    M, _ = preconditioners
    ctx = setup["ctx"]

    def substitute():
        d = ctx.array(setup["bref"])
        M.precondition(d)
        z = d.download()
        d.free()
        return z

    first = {}
    for prec in (N.MXG_FP64, N.MXG_BF16, N.MXG_FP64, N.MXG_FP32, N.MXG_BF16, N.MXG_FP32):
        assert not M.factorize(prec).any()
        z = substitute()
        if prec in first:
            assert np.array_equal(z, first[prec]), prec
        else:
            first[prec] = z
    M.factorize(N.MXG_FP64)


@pytest.mark.parametrize("precision,tol", [(N.MXG_FP64, 1e-11), (N.MXG_FP32, 1e-4)])
def test_hss_ulv_is_exact(setup, precision, tol):
    """theta = 0 makes the preconditioner an HSS matrix, for which the ULV
    factorization is exact: M^{-1} (M x) = x up to rounding."""
    ctx = setup["ctx"]
    M = H2.preconditioner(setup["prob"], 60, SAMP, theta=0.0)
    oM = setup["oprob"].h2(rank=60, orth=True, hss_far=True, sampling=SAMP, theta=0.0)
    x = setup["x"]
    mx = gpu_matvec(ctx, M, x)
    M.factorize(precision)
    d = ctx.array(mx)
    M.precondition(d)
    assert O.rel_err(d.download(), x) < tol
    omx = oM.matvec(x)
    oM.factorize()
    assert O.rel_err(oM.precondition(omx), x) < 1e-11


def test_gmres_reference_mode_matches_oracle(setup, operators, preconditioners):
    A, oA = operators
    M, oM = preconditioners
    ctx = setup["ctx"]
    b = setup["bref"]
    db, dx = ctx.array(b), ctx.zeros(NPTS)
    resid, iters, inner = A.gmres(M, db, dx, EPS, restart=10, max_cycles=50, mode=N.GMRES_REFERENCE)
    x = dx.download()
    ox, ores = oA.gmres(oM, b, EPS, 10, 50)
    assert resid[-1] < EPS and ores[-1] < EPS
    assert inner == 10 * iters
    ax = O.dense_matvec("laplace", 1e3, 0.0, setup["prob"].bodies, x)
    oax = O.dense_matvec("laplace", 1e3, 0.0, setup["prob"].bodies, ox)
    rr, orr = O.rel_err(ax, b), O.rel_err(oax, b)
    assert rr < 10 * orr + 1e-14, (rr, orr)
    assert O.rel_err(x, ox) < 10 * (O.rel_err(ox, setup["x"]) + 1e-14)


@pytest.mark.parametrize("inner_precision", [N.MXG_FP64, N.MXG_FP32])
def test_gmres_cgs2_reaches_tolerance(setup, operators, preconditioners, inner_precision):
    A, _ = operators
    M, _ = preconditioners
    ctx = setup["ctx"]
    b = O.uniform_vec(NPTS, -1, 1, 11)
    db, dx = ctx.array(b), ctx.zeros(NPTS)
    resid, iters, inner = A.gmres(M, db, dx, 1e-11, restart=30, max_cycles=10, mode=N.GMRES_CGS2,
                                  inner_precision=inner_precision)
    assert resid[-1] < 1e-11, resid
    assert iters <= 6
    # same solution as the FP64 reference-mode solve of the same system
    dx2 = ctx.zeros(NPTS)
    A.gmres(M, db, dx2, 1e-11, restart=10, max_cycles=50, mode=N.GMRES_REFERENCE)
    assert O.rel_err(dx.download(), dx2.download()) < 1e-9


def test_gmres_bf16_preconditioner(setup, operators, preconditioners):
    """GMRES-IR with the bfloat16-stored ULV reaches the tolerance and the
    FP64 solution."""
    A, _ = operators
    M, _ = preconditioners
    ctx = setup["ctx"]
    b = O.uniform_vec(NPTS, -1, 1, 17)
    db = ctx.array(b)
    dx64 = ctx.zeros(NPTS)
    A.gmres(M, db, dx64, 1e-11, restart=30, max_cycles=10, mode=N.GMRES_CGS2)
    M.factorize(N.MXG_BF16)
    try:
        dx = ctx.zeros(NPTS)
        resid, iters, inner = A.gmres(M, db, dx, 1e-11, restart=30, max_cycles=10, mode=N.GMRES_CGS2,
                                      inner_precision=N.MXG_FP32)
        assert resid[-1] < 1e-11, resid
        assert O.rel_err(dx.download(), dx64.download()) < 1e-9
    finally:
        M.factorize(N.MXG_FP64)


def test_bf16_schur_products(setup, operators, preconditioners):
    """MXG_BF16_SCHUR: the Schur-complement products of the factorization on
    bfloat16 inputs. Its substitution differs from the FP32 one with
    bfloat16-stored factors (so the products did change) but stays within a
    few percent of the FP64 substitution, and GMRES-IR with it reaches the
    tolerance and the FP64 solution."""
    A, _ = operators
    M, _ = preconditioners
    ctx = setup["ctx"]
    b = O.uniform_vec(NPTS, -1, 1, 19)
    zs = {}
    try:
        for prec in (N.MXG_FP64, N.MXG_BF16, N.MXG_BF16_SCHUR):
            M.factorize(prec)
            d = ctx.array(b)
            M.precondition(d)
            zs[prec] = d.download()
        e16 = O.rel_err(zs[N.MXG_BF16], zs[N.MXG_FP64])
        e_schur = O.rel_err(zs[N.MXG_BF16_SCHUR], zs[N.MXG_FP64])
        print(f"substitution vs FP64: bf16-stored {e16:.3e}, bf16 Schur products {e_schur:.3e}")
        assert not np.array_equal(zs[N.MXG_BF16_SCHUR], zs[N.MXG_BF16])
        assert e_schur < 5e-2
        db = ctx.array(b)
        dx64 = ctx.zeros(NPTS)
        M.factorize(N.MXG_FP64)
        A.gmres(M, db, dx64, 1e-11, restart=30, max_cycles=10, mode=N.GMRES_CGS2)
        M.factorize(N.MXG_BF16_SCHUR)
        dx = ctx.zeros(NPTS)
        resid, iters, inner = A.gmres(M, db, dx, 1e-11, restart=30, max_cycles=10, mode=N.GMRES_CGS2,
                                      inner_precision=N.MXG_FP32)
        assert resid[-1] < 1e-11, resid
        assert O.rel_err(dx.download(), dx64.download()) < 1e-9
    finally:
        M.factorize(N.MXG_FP64)


def test_gmres_with_looser_inner_operator(setup, operators, preconditioners):
    """Iterative refinement: FP32 inner iterations on an operator built at
    1e-6, every cycle's residual on the accurate operator, reaches the same
    solution as the plain FP64 solve."""
    A, _ = operators
    M, _ = preconditioners
    ctx = setup["ctx"]
    A_in = H2.operator(setup["prob"], 1e-6, SAMP)
    b = O.uniform_vec(NPTS, -1, 1, 13)
    db, dx = ctx.array(b), ctx.zeros(NPTS)
    resid, iters, inner = A.gmres(M, db, dx, 1e-11, restart=30, max_cycles=10, mode=N.GMRES_CGS2,
                                  inner_precision=N.MXG_FP32, inner=A_in)
    assert resid[-1] < 1e-11, resid
    assert 1 <= iters <= 8
    dx2 = ctx.zeros(NPTS)
    r2, _, _ = A.gmres(M, db, dx2, 1e-11, restart=30, max_cycles=10, mode=N.GMRES_CGS2)
    assert O.rel_err(dx.download(), dx2.download()) < 1e-9
    # a zero initial guess takes r = b; a nonzero one the operator residual
    dx3 = ctx.array(dx.download())
    r3, it3, _ = A.gmres(M, db, dx3, 1e-11, restart=30, max_cycles=10, mode=N.GMRES_CGS2)
    assert it3 == 0 and r3[0] < 1e-11


def test_batched_right_hand_sides_match_single(setup, operators, preconditioners):
    """Matvec, substitution and GMRES over a batch of 5 right-hand sides (one
    launch of 4 and one of 1) give every column bit for bit what a solve of
    that column alone gives."""
    A, _ = operators
    M, _ = preconditioners
    ctx = setup["ctx"]
    nq = 5
    X = np.stack([O.uniform_vec(NPTS, -1, 1, 20 + q) for q in range(nq)], axis=1)
    devA = N.h2().mxc_solver_device(A.handle)
    for prec in (N.MXG_FP64, N.MXG_FP32):
        dX, dY = ctx.array(X), ctx.empty((NPTS, nq))
        N.check_gpu(N.gpu().mxg_h2_matvec(devA, dX.ptr, dY.ptr, nq, NPTS, prec))
        Y = dY.download()
        for q in range(nq):
            assert np.array_equal(Y[:, q], gpu_matvec(ctx, A, X[:, q], prec)), (prec, q)
    for fprec in (N.MXG_FP64, N.MXG_BF16):
        M.factorize(fprec)
        ulv = N.h2().mxc_solver_ulv(M.handle)
        dZ = ctx.array(X)
        N.check_gpu(N.gpu().mxg_ulv_solve(ulv, dZ.ptr, nq, NPTS))
        Z = dZ.download()
        for q in range(nq):
            d = ctx.array(X[:, q])
            M.precondition(d)
            assert np.array_equal(Z[:, q], d.download()), (fprec, q)
    M.factorize(N.MXG_FP64)
    dB, dXb = ctx.array(X), ctx.zeros((NPTS, nq))
    res, its, inn = A.gmres_batch(M, dB, dXb, 1e-11, restart=30, max_cycles=10, mode=N.GMRES_CGS2,
                                  inner_precision=N.MXG_FP32)
    Xb = dXb.download()
    for q in range(nq):
        db, dx = ctx.array(X[:, q]), ctx.zeros(NPTS)
        r1, i1, n1 = A.gmres(M, db, dx, 1e-11, restart=30, max_cycles=10, mode=N.GMRES_CGS2, inner_precision=N.MXG_FP32)
        assert res[q][-1] < 1e-11 and (its[q], inn[q]) == (i1, n1)
        assert np.array_equal(Xb[:, q], dx.download()), q


def test_zero_right_hand_side_in_a_batch(setup, operators, preconditioners):
    """A zero column among nonzero ones: its solution stays zero and takes no
    cycle (the reference's loop also ends at once there: |r|/|b| is NaN, which
    no tolerance test passes), and the other columns equal their single
    solves bit for bit."""
    A, _ = operators
    M, _ = preconditioners
    ctx = setup["ctx"]
    M.factorize(N.MXG_FP64)
    X = np.stack([O.uniform_vec(NPTS, -1, 1, 40), np.zeros(NPTS), O.uniform_vec(NPTS, -1, 1, 41)], axis=1)
    dB, dX = ctx.array(X), ctx.zeros((NPTS, 3))
    res, its, inn = A.gmres_batch(M, dB, dX, 1e-11, restart=30, max_cycles=10, mode=N.GMRES_CGS2)
    Xb = dX.download()
    assert its[1] == 0 and inn[1] == 0 and res[1][-1] == 0.0 and not Xb[:, 1].any()
    for q in (0, 2):
        db, dx = ctx.array(X[:, q]), ctx.zeros(NPTS)
        r1, i1, n1 = A.gmres(M, db, dx, 1e-11, restart=30, max_cycles=10, mode=N.GMRES_CGS2)
        assert res[q][-1] < 1e-11 and (its[q], inn[q]) == (i1, n1)
        assert np.array_equal(Xb[:, q], dx.download()), q
