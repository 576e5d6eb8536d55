"""Shared builders for the tests: small problems and perturbed states."""

from __future__ import annotations

import numpy as np

from scratch_wadge_amd import operators
from scratch_wadge_amd import problem as P


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def has_rule(dim, N):
    try:
        operators.build_ref_ops(dim, N)
        return True
    except KeyError:
        return False


def small_problem(dim, N, curved=True, tau=1.0):
    if dim == 2:
        return P.vortex2d_problem(n=4, N=N, perturb=0.05, seed=3, tau=tau)
    return P.tgv3d_problem(n=3, N=N, curved=curved, tau=tau)


def rough_state(prob, seed=0, amp=0.3):
    """A smooth but strongly varying state (density and pressure vary by ~amp),
    projected like the reference's initial condition, so that both branches of
    the log mean are exercised.  Returns (Q, rhsQ, Qf) in kernel layout."""
    rng = np.random.default_rng(seed)
    xq = prob.xq
    dim = prob.dim
    ph = rng.uniform(0, 2 * np.pi, size=(4, dim))
    kv = rng.integers(1, 3 if dim == 2 else 2, size=(4, dim))

    def wave(m):
        return np.sin(sum(kv[m, d] * xq[d] * (2 * np.pi / 10.0 if dim == 2 else 1.0) + ph[m, d]
                          for d in range(dim)))

    rho = 1.0 + amp * wave(0)
    p = 1.0 + amp * wave(1)
    vel = [0.5 * wave(2 + (d % 2)) * (1 if d == 0 else -1) for d in range(dim)]
    E = p / (prob.gamma - 1.0) + 0.5 * rho * sum(v * v for v in vel)
    U = np.stack([rho] + [rho * v for v in vel] + [E], 0)
    Q, Uq, Uf = P.project_initial_state(prob.ops, U, prob.gamma)
    assert np.isfinite(Uq).all() and np.isfinite(Uf).all(), "rough state left the admissible set"
    lay = P._to_layout
    return lay(Q), lay(Uq), lay(Uf)
