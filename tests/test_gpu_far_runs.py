"""Far-body lists handed to the device as runs of consecutive bodies
(mxg_h2_construct_level_runs, expanded by k_expand_runs) build the same H^2
matrices, bit for bit, as the same lists given as one index per body
(mxg_h2_construct_level): every level's dims, ranks, near blocks, bases, R
factors, couplings and skeletons. Sampling caps of 300 and 2000 make the
lists mix whole cells (long runs) with sampled bodies (runs of one); the
preconditioner's lists take the shell-0 quotas."""
import numpy as np
import pytest

import oracle_lib as O
from mxc_amd.solver import H2, Problem

pytestmark = pytest.mark.gpu


def same_matrix(a, b, levels):
    for l in range(levels + 1):
        da, ra = a.dims(l)
        db, rb = b.dims(l)
        assert np.array_equal(da, db) and np.array_equal(ra, rb), l
        for what in range(5):
            xa, xb = a.download(l, what), b.download(l, what)
            assert xa.shape == xb.shape and np.array_equal(xa, xb), (l, what)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("sampling", [300, 2000])
def test_runs_build_the_same_matrices(ctx, sampling):
    prob = Problem(ctx, O.cube_points(8192, 3), 128)
    A_runs = H2.operator(prob, 1e-10, sampling, "yukawa")
    A_idx = H2.operator(prob, 1e-10, sampling, "yukawa", explicit_far=True)
    same_matrix(A_runs, A_idx, prob.levels)
    M_runs = H2(prob, "yukawa", rank=40, orth=True, hss_far=True, sampling=sampling, factorable=True)
    M_idx = H2(prob, "yukawa", rank=40, orth=True, hss_far=True, sampling=sampling, factorable=True, explicit_far=True)
    same_matrix(M_runs, M_idx, prob.levels)
