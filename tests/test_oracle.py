"""The oracle pinned against known answers and mathematics (CPU only).

Dense building blocks against LAPACK known answers committed in
golden/dense_kats.npz (Householder QR, column-pivoted QR with the threshold
rank rule, LU solve) and against the kernels' closed forms. The H^2 path
against properties that fix its result independently of the implementation:
a matvec without compression equals the dense product; the tolerance
construction error scales with eps; ULV on an HSS matrix (theta = 0) or
without compression is an exact solve; GMRES reaches the dense solution."""
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dense_kats.npz"))
CPQR_CASES = sorted({k[len("cpqr_"):-len("_A")] for k in GOLD.files if k.startswith("cpqr_") and k.endswith("_A")})


def fortran(a):
    return np.asfortranarray(np.array(a, dtype=np.float64))


@pytest.mark.parametrize("case", CPQR_CASES)
def test_householder_qr_matches_lapack(case):
    A = fortran(GOLD[f"cpqr_{case}_A"])
    m, n = A.shape
    O.lib().orc_qr_r(m, n, O.ptr(A), m)
    R = A[: min(m, n)]
    Rl = GOLD[f"qr_{case}_R"]
    # rows of R are fixed up to sign
    scale = np.abs(Rl).max()
    assert np.max(np.abs(np.abs(R) - np.abs(Rl))) < 1e-12 * scale


@pytest.mark.parametrize("case", CPQR_CASES)
def test_cpqr_matches_lapack_and_rank_rule(case):
    A0 = GOLD[f"cpqr_{case}_A"]
    Rl, pl = GOLD[f"cpqr_{case}_R"], GOLD[f"cpqr_{case}_perm"]
    m, n = A0.shape
    dl = np.abs(np.diag(Rl))
    # pivots are determined while the remaining column norms are well above rounding
    k = int(np.sum(dl > 1e-10 * dl[0]))
    for tol in (0.0, 1e-4, 1e-8, 1e-12):
        A = fortran(A0)
        perm = np.zeros(n, np.int64)
        mp, nz = O.C.c_double(), O.C.c_longlong()
        rank = O.lib().orc_cpqr(m, n, O.ptr(A), m, O.ptr(perm), tol, O.C.byref(mp), O.C.byref(nz))
        assert np.array_equal(perm[:k], pl[:k])
        d = np.abs(np.diag(A[: min(m, n)]))
        assert np.max(np.abs(d[:k] - dl[:k])) < 1e-12 * dl[0]
        assert abs(mp.value - dl[0]) < 1e-12 * dl[0]
        if tol == 0.0:
            assert rank == min(m, n)
        else:
            # Eigen's rule: the number of |R_ii| above tol * max pivot
            thr = tol * dl[0]
            if np.min(np.abs(dl - thr)) > 0.01 * thr:
                assert rank == int(np.sum(dl > thr)), (rank, dl, thr)


@pytest.mark.parametrize("case", ["lu_10", "lu_33"])
def test_lu_solve_matches_lapack(case):
    A = fortran(GOLD[f"{case}_A"])
    B = fortran(GOLD[f"{case}_B"])
    n, nrhs = B.shape
    logdet = O.lib().orc_lu_solve(n, nrhs, O.ptr(A), O.ptr(B))
    X = GOLD[f"{case}_X"]
    assert np.max(np.abs(B - X)) < 1e-11 * np.abs(X).max()
    assert abs(logdet - float(GOLD[f"{case}_logdet"])) < 1e-12 * max(1.0, abs(float(GOLD[f"{case}_logdet"])))


@pytest.mark.parametrize("kind,diag,alpha", [("laplace", 1e3, 0.0), ("yukawa", 1e3, 1.0), ("gaussian", 0.0, 2.0)])
def test_kernels_match_closed_forms(kind, diag, alpha):
    pts = GOLD["kern_points"]
    x = GOLD["kern_x"]
    y = O.dense_matvec(kind, diag, alpha, np.ascontiguousarray(pts.reshape(-1)), x, nthreads=2)
    yl = GOLD[f"kern_{kind}_y"]
    assert np.max(np.abs(y - yl)) < 1e-13 * np.abs(yl).max()
    # single entries, including the diagonal value at d = 0
    kid = O.KID[kind]
    for d in (0.0, 1e-3, 0.5, 2.0):
        v = O.lib().orc_kernel_eval(kid, diag, alpha, d)
        if kind == "gaussian":
            ref = 1.0001 if d == 0 else np.exp(-alpha * d * d)
        elif d == 0:
            ref = diag
        else:
            ref = np.exp(-alpha * d) / d
        assert v == pytest.approx(ref, rel=1e-15)
    # an off-diagonal block through the block generator
    bi, bj = np.ascontiguousarray(pts[:7].reshape(-1)), np.ascontiguousarray(pts[20:31].reshape(-1))
    Ab = np.zeros(7 * 11)
    O.lib().orc_gen_matrix(kid, diag, alpha, 7, 11, O.ptr(bi), O.ptr(bj), O.ptr(Ab))
    r = np.linalg.norm(pts[:7, None] - pts[None, 20:31], axis=-1)
    ref = np.exp(-alpha * r * r) if kind == "gaussian" else np.exp(-alpha * r) / r
    assert np.max(np.abs(Ab.reshape(11, 7).T - ref)) < 1e-15 * np.abs(ref).max() * 4


def test_rademacher_signs_are_balanced():
    t = np.arange(4096)
    s = np.array([[O.lib().orc_rademacher(int(a), int(i)) for i in range(0, 256, 3)] for a in t[:512]])
    assert set(np.unique(s)) == {-1.0, 1.0}
    assert abs(s.mean()) < 0.02
    # different sketch rows are uncorrelated
    c = (s.T @ s) / s.shape[0]
    assert np.max(np.abs(c - np.eye(c.shape[0]))) < 0.25


# ---------------- H^2 path ----------------

def kernel_dense(kind, bodies, diag=1e3, alpha=1.0):
    p = bodies.reshape(-1, 3)
    d = np.linalg.norm(p[:, None] - p[None, :], axis=-1)
    off = ~np.eye(len(p), dtype=bool)
    K = np.full_like(d, diag)
    K[off] = (np.exp(-alpha * d[off]) if kind == "yukawa" else 1.0) / d[off]
    return K


@pytest.fixture(scope="module")
def small():
    pts = O.cube_points(2048, 3)
    prob = O.Problem(pts, 128)
    x = O.uniform_vec(2048, 0, 1, 7)
    K = kernel_dense("laplace", prob.bodies)
    return prob, x, K


def test_matvec_without_compression_is_dense(small):
    prob, x, K = small
    A = prob.h2(rank=1 << 20, orth=True, sampling=0)
    assert O.rel_err(A.matvec(x), K @ x) < 1e-13


@pytest.mark.parametrize("sketch", [1, 0])
def test_construct_error_follows_tolerance(small, sketch):
    prob, x, K = small
    b = K @ x
    O.lib().orc_set_sketch(sketch)
    try:
        errs = [O.rel_err(prob.h2(tol=eps, sampling=0).matvec(x), b) for eps in (1e-4, 1e-7, 1e-10, 1e-13)]
    finally:
        O.lib().orc_set_sketch(1)
    for eps, e in zip((1e-4, 1e-7, 1e-10, 1e-13), errs):
        assert e < 10 * eps + 1e-15, (eps, e)
    assert errs[0] > errs[1] > errs[2]


def test_sampling_keeps_operator_accurate(small):
    prob, x, K = small
    e = O.rel_err(prob.h2(tol=1e-10, sampling=600).matvec(x), K @ x)
    assert e < 1e-9


def test_ulv_is_exact_for_hss_matrix(small):
    prob, x, _ = small
    M = prob.h2(rank=40, orth=True, hss_far=True, sampling=0, theta=0.0)
    y = M.matvec(x)
    assert M.factorize() == 0
    assert O.rel_err(M.precondition(y), x) < 1e-12


def test_ulv_without_compression_is_dense_solve():
    pts = O.cube_points(512, 5)
    prob = O.Problem(pts, 64)
    K = kernel_dense("laplace", prob.bodies)
    b = O.uniform_vec(512, -1, 1, 9)
    M = prob.h2(rank=1 << 20, orth=True, hss_far=True, sampling=0)
    assert O.rel_err(M.matvec(b), K @ b) < 1e-13
    assert M.factorize() == 0
    assert O.rel_err(M.precondition(b), np.linalg.solve(K, b)) < 1e-12


def test_gmres_reaches_dense_solution(small):
    prob, _, K = small
    b = O.uniform_vec(2048, -1, 1, 4)
    A = prob.h2(tol=1e-12, sampling=0)
    M = prob.h2(rank=16, orth=True, hss_far=True, sampling=0)
    assert M.factorize() == 0
    x, res = A.gmres(M, b, 1e-12, 10, 20)
    assert res[-1] < 1e-12
    assert np.all(np.diff(res[1:]) <= 0)
    xs = np.linalg.solve(K, b)
    assert O.rel_err(x, xs) < 1e-10


@pytest.mark.parametrize("n,leaf,kernel,tol,rank", [
    (256, 128, "laplace", 0.0, 24),
    (257, 128, "yukawa", 1e-4, 60),
    (200, 100, "laplace", 1e-6, 30),
])
def test_single_schur_ulv_of_two_near_boxes_is_exact(n, leaf, kernel, tol, rank):
    """Pins the off-diagonal Schur updates of the factorization
    (h2matrix.cpp:405-424): with two leaves that are near each other (L = 1),
    every term of the single Schur complement onto the skeleton runs
    (A_ii.sr A_ij.rs, A_ij.sr B_j, A_ij.sr rr_j^-1 A_ij.sr^T) and the root is a
    full LU. The preconditioner is then exactly U A'^-1 U^T, where A' = U^T A U
    (U = blockdiag of the boxes' full orthogonal bases, skeleton columns first)
    with the r x r coupling between the two boxes' redundant parts removed:
    the one interaction a single Schur complement leaves out."""
    prob = O.Problem(O.cube_points(n, 11), leaf)
    assert prob.levels == 1
    M = prob.h2(kernel=kernel, alpha=1.0, tol=tol, rank=rank, orth=True, hss_far=True, sampling=0)
    assert O.lib().orc_h2_near_count(M.h, 1) == 4  # the two leaves are near each other
    d, s = M.dims(1)
    assert np.all(s > 0) and np.all(s < d)
    q = M.download(1, 1)
    U = np.zeros((n, n))
    o = oq = 0
    for m in map(int, d):
        U[o:o + m, o:o + m] = q[oq:oq + m * m].reshape(m, m, order="F")
        o += m
        oq += m * m
    assert np.abs(U.T @ U - np.eye(n)).max() < 1e-12
    K = kernel_dense(kernel, prob.bodies)
    Ah = U.T @ K @ U
    r0 = np.arange(s[0], d[0])
    r1 = np.arange(d[0] + s[1], n)
    Ah[np.ix_(r0, r1)] = 0.
    Ah[np.ix_(r1, r0)] = 0.
    b = O.uniform_vec(n, -1, 1, 5)
    expect = U @ np.linalg.solve(Ah, U.T @ b)
    assert M.factorize() == 0
    assert O.rel_err(M.precondition(b), expect) < 1e-12
    # the dropped coupling matters: the exact inverse of A differs
    assert O.rel_err(np.linalg.solve(K, b), expect) > 1e-9
