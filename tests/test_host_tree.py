"""Host layer (tree, block lists, far-body sampling, generators) against the
oracle, bit for bit, and the oracle against independent checks: a from-spec
MT19937, tree invariants, and brute-force block lists and far-body lists."""
import numpy as np
import pytest

import oracle_lib as O
from mxc_amd import solver as S


# ---------------- random generators ----------------

class MT19937:
    """Mersenne Twister MT19937 from Matsumoto & Nishimura (1998), the engine
    std::mt19937 is defined as."""

    def __init__(self, seed: int):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.idx = 624

    def __call__(self) -> int:
        if self.idx >= 624:
            for i in range(624):
                y = (self.mt[i] & 0x80000000) | (self.mt[(i + 1) % 624] & 0x7FFFFFFF)
                self.mt[i] = self.mt[(i + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            self.idx = 0
        y = self.mt[self.idx]
        self.idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y


def canonical(gen) -> float:
    """generate_canonical<double, 53> over a 32-bit engine: two draws,
    (g1 + g2 * 2^32) / 2^64, clamped below 1 (libstdc++)."""
    s = gen() + gen() * 4294967296.0
    r = s / 18446744073709551616.0
    return r if r < 1.0 else np.nextafter(1.0, 0.0)


def test_mt19937_known_answer():
    # [rand.predef]: the 10000th output of a default-constructed mt19937 is 4123659995.
    g = MT19937(5489)
    for _ in range(9999):
        g()
    assert g() == 4123659995


@pytest.mark.parametrize("seed", [0, 1, 2, 999])
def test_cube_points_from_spec(seed):
    n = 500
    g = MT19937(seed)
    ref = np.array([canonical(g) for _ in range(3 * n)])
    assert np.array_equal(O.cube_points(n, seed), ref)
    assert np.array_equal(S.generate_points("cube", n, seed), ref)


def test_sphere_points_match_oracle_and_lie_on_sphere():
    n = 4000
    p = S.generate_points("sphere", n, 2)
    assert np.array_equal(p, O.sphere_points(n, 2))
    P = p.reshape(-1, 3)
    assert np.max(np.abs(np.linalg.norm(P, axis=1) - 1)) < 4e-16 * 8
    # a rigid rotation of the golden-angle points: pairwise distances preserved
    m = np.zeros(3 * n)
    O.lib().orc_mesh_sphere(O.ptr(m), n, 1.0)
    Mp = m.reshape(-1, 3)
    idx = np.random.default_rng(0).choice(n, 200, replace=False)
    dp = np.linalg.norm(P[idx, None] - P[None, idx], axis=-1)
    dm = np.linalg.norm(Mp[idx, None] - Mp[None, idx], axis=-1)
    assert np.max(np.abs(dp - dm)) < 1e-14
    # the rotation (fit by least squares) is proper orthogonal and depends on the seed
    Rm, *_ = np.linalg.lstsq(Mp, P, rcond=None)
    assert np.allclose(Rm.T @ Rm, np.eye(3), atol=1e-12) and abs(np.linalg.det(Rm) - 1) < 1e-12
    assert not np.allclose(S.generate_points("sphere", n, 3), p)
    # uniform on the sphere: each coordinate has mean 0 and second moment 1/3
    assert np.all(np.abs(P.mean(0)) < 0.01) and np.all(np.abs((P**2).mean(0) - 1 / 3) < 0.01)
    # quasi-uniform: no pair closer than half the mean spacing
    assert np.min(dm + 9 * np.eye(200)) > 0.5 * np.sqrt(4 * np.pi / n)


def test_mesh_sphere_and_rhs_match_oracle():
    n = 1000
    a = S.generate_points("mesh", n, 0)
    b = np.zeros(3 * n)
    O.lib().orc_mesh_sphere(O.ptr(b), n, np.sqrt(n / (4 * np.pi)))
    assert np.array_equal(a, b)
    assert np.array_equal(S.uniform_vector(777, -1, 1, 5), O.uniform_vec(777, -1, 1, 5))


# ---------------- tree ----------------

CASES = [("cube", 2048, 128, 0), ("cube", 20000, 256, 0), ("sphere", 4096, 256, 2), ("cube", 1000, 1000, 3),
         ("cube", 300, 64, 4)]


@pytest.mark.parametrize("geom,n,leaf,seed", CASES)
def test_tree_matches_oracle(geom, n, leaf, seed):
    pts = S.generate_points(geom, n, seed)
    levels = S.tree_levels(n, leaf)
    assert levels == O.lib().orc_levels(n, leaf)
    mine = S.build_tree(pts, levels)
    ref = O.build_tree(pts, levels)
    for a, b in zip(mine, ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("geom,n,leaf,seed", CASES)
def test_tree_invariants(geom, n, leaf, seed):
    pts = S.generate_points(geom, n, seed)
    levels = S.tree_levels(n, leaf)
    bodies, child, body, R, Cc = O.build_tree(pts, levels)
    b3 = bodies.reshape(-1, 3)
    # a permutation of the input
    assert np.array_equal(np.sort(b3, axis=0), np.sort(pts.reshape(-1, 3), axis=0))
    nc = (2 << levels) - 1
    for c in range(nc):
        lo, hi = body[2 * c], body[2 * c + 1]
        blk = b3[lo:hi]
        mn, mx = blk.min(0), blk.max(0)
        assert np.array_equal(Cc[3 * c:3 * c + 3], (mn + mx) * 0.5)
        assert np.array_equal(R[3 * c:3 * c + 3], (mx - mn) * 0.5 + 1e-8)
        if c < (1 << levels) - 1:
            c0, c1 = child[2 * c], child[2 * c + 1]
            assert (c0, c1) == (2 * c + 1, 2 * c + 3)
            mid = lo + (hi - lo) // 2
            assert (body[2 * c0], body[2 * c0 + 1], body[2 * c0 + 2], body[2 * c0 + 3]) == (lo, mid, mid, hi)
            axis = int(np.argmax(R[3 * c:3 * c + 3]))
            assert b3[lo:mid, axis].max() <= b3[mid:hi, axis].min()
        else:
            assert child[2 * c] == -1


# ---------------- block lists ----------------

def brute_lists(levels, R, Cc, theta):
    """Near/far pairs by definition: a same-level pair is visited when every
    ancestor pair is inadmissible; visited inadmissible pairs are near,
    visited admissible pairs are far."""
    nc = (2 << levels) - 1
    Rr = R.reshape(-1, 3)
    Cm = Cc.reshape(-1, 3)

    def adm(i, j):
        return np.sum((Cm[i] - Cm[j]) ** 2) > theta * (np.sum(Rr[i] ** 2) + np.sum(Rr[j] ** 2))

    near, far = set(), set()
    for l in range(levels + 1):
        first = (1 << l) - 1
        for i in range(first, 2 * first + 1):
            for j in range(first, 2 * first + 1):
                a, b, ok = i, j, True
                while a > 0:
                    a, b = (a - 1) // 2, (b - 1) // 2
                    if adm(a, b):
                        ok = False
                        break
                if ok:
                    (far if adm(i, j) else near).add((i, j))
    def to_csr(s):
        rp = np.zeros(nc + 1, np.int64)
        col = []
        for i in range(nc):
            row = sorted(j for (a, j) in s if a == i)
            col += row
            rp[i + 1] = rp[i] + len(row)
        return rp, np.array(col, np.int64)
    return to_csr(near), to_csr(far)


@pytest.mark.parametrize("theta", [0.0, 1.0, 2.0, 3.0])
@pytest.mark.parametrize("geom,n,leaf,seed", [("cube", 2048, 64, 0), ("sphere", 4096, 128, 2)])
def test_csr_matches_oracle_and_brute_force(geom, n, leaf, seed, theta):
    pts = S.generate_points(geom, n, seed)
    levels = S.tree_levels(n, leaf)
    bodies, child, body, R, Cc = O.build_tree(pts, levels)
    (bn_rp, bn_col), (bf_rp, bf_col) = brute_lists(levels, R, Cc, theta)
    for far, (brp, bcol) in ((False, (bn_rp, bn_col)), (True, (bf_rp, bf_col))):
        orp, ocol = O.csr(far, theta, levels, child, R, Cc)
        mrp, mcol = S.csr(far, theta, levels, child, body, R, Cc)
        assert np.array_equal(orp, brp) and np.array_equal(ocol, bcol)
        assert np.array_equal(mrp, orp) and np.array_equal(mcol, ocol)


# ---------------- far-body lists and sampling ----------------

def full_far_list(levels, body, frp, fcol, cell):
    """The reference's far list: the parent's list followed by the bodies of
    the cell's own far cells (h-matrix.cpp:121-127,160-165)."""
    if cell == 0:
        return []
    out = full_far_list(levels, body, frp, fcol, (cell - 1) // 2)
    for k in range(frp[cell], frp[cell + 1]):
        f = fcol[k]
        out += list(range(body[2 * f], body[2 * f + 1]))
    return out


def cell_bodies(body, cells):
    out = set()
    for f in cells:
        out |= set(range(body[2 * f], body[2 * f + 1]))
    return out


@pytest.mark.parametrize("hss", [False, True])
@pytest.mark.parametrize("sampling", [0, 100, 2000, 5000])
@pytest.mark.parametrize("n,leaf", [(8192, 128), (32768, 32)])
def test_far_sample_matches_oracle(hss, sampling, n, leaf):
    theta = 2.0
    pts = S.generate_points("cube", n, 1)
    levels = S.tree_levels(n, leaf)
    bodies, child, body, R, Cc = O.build_tree(pts, levels)
    near = O.csr(False, theta, levels, child, R, Cc)
    far = O.csr(True, theta, levels, child, R, Cc)
    bfar = O.csr(True, 0.0, levels, child, R, Cc) if hss else far
    nc = (2 << levels) - 1
    rng = np.random.default_rng(0)
    for cell in list(rng.choice(nc, 24, replace=False)) + [0, 1, 2, nc - 1]:
        cell = int(cell)
        mine = S.far_sample(levels, body, bfar, near, far, hss, cell, sampling)
        out = np.zeros(n, np.int64)
        k = O.lib().orc_far_sample(levels, O.ptr(body), O.ptr(bfar[0]), O.ptr(bfar[1]), O.ptr(near[0]), O.ptr(near[1]),
                                   O.ptr(far[0]), O.ptr(far[1]), int(hss), cell, sampling, O.ptr(out), n)
        assert np.array_equal(mine, out[:k])
        full = full_far_list(levels, body, bfar[0], bfar[1], cell)
        if hss and cell > 0:
            lo, hi = body[2 * cell], body[2 * cell + 1]
            assert sorted(full) == [b for b in range(n) if not lo <= b < hi]  # every body outside the cell
        if sampling <= 0 or len(full) <= sampling:
            assert list(mine) == full
            continue
        assert set(mine.tolist()) <= set(full)
        assert len(set(mine.tolist())) == len(mine)  # distinct
        # distance shells partition the reference list
        shells = []
        if hss:
            shells.append(cell_bodies(body, [j for j in near[1][near[0][cell]:near[0][cell + 1]] if j != cell]))
        a = cell
        while a > 0:
            shells.append(cell_bodies(body, far[1][far[0][a]:far[0][a + 1]]))
            a = (a - 1) // 2
        assert sum(len(s) for s in shells) == len(full) and set().union(*shells) == set(full)
        got = set(mine.tolist())
        if hss:
            # a budget of exactly `sampling`; the closest shell keeps at least
            # min(its size, half the budget)
            assert len(mine) == sampling
            assert len(shells[0] & got) >= min(len(shells[0]), (sampling + 1) // 2)
        else:
            # the max(3, ceil(levels / 2)) closest shells whole, every farther
            # shell min(size, sampling)
            res = max(3, (levels + 1) // 2)
            for k, sk in enumerate(shells):
                want = len(sk) if k < res else min(len(sk), sampling)
                assert len(sk & got) == want
            assert len(mine) == sum(len(sk) if k < res else min(len(sk), sampling) for k, sk in enumerate(shells))
