"""Variable-size batches of independent H^2 problems on one device.

Each problem has its own points, tree depth, leaf size, admissibility and
rank; their preconditioners are factorized together, every phase of a level
sweep one launch shared by all of them (mxg_ulv_create_batch), and each is
then solved with its own operator. A problem's factorization is the same as
when it is factorized alone.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import _native as N
from .solver import H2, Problem, generate_points


@dataclasses.dataclass
class Spec:
    n: int
    theta: float
    leaf: int
    rank: int
    seed: int
    eps: float = 1e-8
    kernel: str = "laplace"
    diag: float = 1e3
    alpha: float = 1.0
    geometry: str = "cube"
    sampling: int = 0  # every far body, as the reference (problems of this size afford it)


def log_uniform_specs(count: int, seed: int, n_lo: int = 4096, n_hi: int = 65536, thetas=(1.0, 2.0, 3.0),
                      leaves=(128, 256), ranks=(32, 64, 128), eps: float = 1e-8):
    """count problems with N log-uniform in [n_lo, n_hi] and theta, leaf and
    rank drawn uniformly from the given sets (numpy's PCG64 from `seed`)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))))
        out.append(Spec(n=n, theta=float(rng.choice(thetas)), leaf=int(rng.choice(leaves)), rank=int(rng.choice(ranks)),
                        seed=1000 + k, eps=eps))
    return out


class Member:
    """One problem of a batch: its points, operator and preconditioner."""

    def __init__(self, ctx, spec: Spec):
        self.spec = spec
        s = spec
        self.problem = Problem(ctx, generate_points(s.geometry, s.n, s.seed), s.leaf)
        self.A = H2.operator(self.problem, s.eps, s.sampling, s.kernel, s.diag, s.alpha, s.theta)
        self.M = H2.preconditioner(self.problem, s.rank, s.sampling, s.kernel, s.diag, s.alpha, s.theta)

    def close(self):
        for o in (self.M, self.A, self.problem):
            o.close()


class Batch:
    """Independent problems resident on one context."""

    def __init__(self, ctx, specs):
        self.ctx = ctx
        self.members = [Member(ctx, s) for s in specs]

    def factorize(self, precision: int = N.MXG_FP32):
        """Every preconditioner factorized in shared launches; returns the
        singular-block counts per level of each problem."""
        return H2.factorize_batch([m.M for m in self.members], precision)

    def factor_seconds(self) -> float:
        return self.members[0].M.factor_seconds()

    def close(self):
        for m in self.members:
            m.close()
