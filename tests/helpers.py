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
    return P.varying_state(prob, seed=seed, amp=amp)
