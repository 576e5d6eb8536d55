"""The workloads at their full sizes (3D Taylor-Green on 32^3 x 6 = 196,608
curved tets at N = 3 and N = 4, and on 16^3 x 6 = 24,576 affine tets): the GPU
RHS against the oracle's over every element, and the size-independent
properties of the scheme — conservation of every field, and entropy
conservation of the dissipation-free flux (tau = 0) at round-off, entropy
dissipation with the penalty.

At these resolutions the Taylor-Green RHS is a small residual of O(100)
pressure fluxes that cancel (Mach 0.1, p ~ 71): its relative round-off floor is
well above 1e-16.  The tests measure that floor with the oracle itself (its RHS
at inputs perturbed by about one ulp, and its own entropy production) and hold
the GPU path to 1e-12 or to ten times that floor, whichever is larger."""

import numpy as np
import pytest
import torch

from helpers import relerr
from scratch_wadge_amd import problem as P

pytestmark = pytest.mark.gpu

CASES = [(32, 3, True), (32, 4, True), (16, 3, False)]


def _rhs_gpu(prob):
    from scratch_wadge_amd.device import DeviceSolver
    dev = DeviceSolver(prob)
    r, _ = dev.rhs(prob.rhsQ, prob.Qf)
    del dev
    torch.cuda.empty_cache()
    return r


def _entropy_terms(prob, r):
    nf = prob.nfields
    U = prob.rhsQ.reshape(prob.K, nf, prob.Nq).transpose(1, 2, 0)
    V = P.entropy_vars(U)
    R = r.reshape(prob.K, nf, prob.Nq).transpose(1, 2, 0)
    return prob.wq[:, None] * np.einsum("fik,fik->ik", V, R)


@pytest.mark.parametrize("n,N,curved", CASES)
def test_full_size_rhs_parity_and_conservation(oracle, n, N, curved):
    prob = P.tgv3d_problem(n=n, N=N, curved=curved)
    print(f"case full size n={n} N={N} curved={curved} K={prob.K}", flush=True)
    r_g = _rhs_gpu(prob)
    r_o, _ = oracle.rhs(prob, prob.rhsQ.copy(), prob.Qf.copy())
    rng = np.random.default_rng(0)
    ulp = lambda a: a * (1.0 + 2.2e-16 * rng.standard_normal(a.shape))
    r_p, _ = oracle.rhs(prob, ulp(prob.rhsQ), ulp(prob.Qf))
    floor = relerr(r_p, r_o)
    e = relerr(r_g, r_o)
    R = r_g.reshape(prob.K, prob.nfields, prob.Nq)
    tot = np.einsum("i,kfi->f", prob.wq, R)
    scale = np.einsum("i,kfi->f", prob.wq, np.abs(R)).max()
    print(f"rel L2 vs oracle {e:.2e}; oracle round-off floor {floor:.2e}; "
          f"conservation {np.abs(tot).max() / scale:.2e}", flush=True)
    assert e < max(1e-12, 10 * floor)
    assert np.abs(tot).max() <= 1e-11 * scale


def _rough(prob, amp=1e-3, seed=0):
    """The projected state of the initial condition with per-node noise of
    relative size amp: traces jump across every face, so the penalty acts."""
    nf, K, Nq = prob.nfields, prob.K, prob.Nq
    rng = np.random.default_rng(seed)
    U = prob.Q.reshape(K, nf, Nq).transpose(1, 2, 0)
    U = U * (1.0 + amp * rng.standard_normal(U.shape))
    _, uq, uf = P.project_initial_state(prob.ops, U, prob.gamma)
    prob.rhsQ, prob.Qf = P._to_layout(uq), P._to_layout(uf)
    return prob


@pytest.mark.parametrize("n,N,curved", CASES[:2])
def test_full_size_entropy_production(oracle, n, N, curved):
    prob0 = _rough(P.tgv3d_problem(n=n, N=N, curved=curved, tau=0.0))
    print(f"case entropy n={n} N={N} K={prob0.K}", flush=True)
    t0 = _entropy_terms(prob0, _rhs_gpu(prob0))
    rel0 = abs(t0.sum()) / np.abs(t0).sum()
    r_o, _ = oracle.rhs(prob0, prob0.rhsQ.copy(), prob0.Qf.copy())
    to = _entropy_terms(prob0, r_o)
    rel_o = abs(to.sum()) / np.abs(to).sum()
    print(f"oracle's own entropy production (relative) {rel_o:.2e}", flush=True)
    prob1 = _rough(P.tgv3d_problem(n=n, N=N, curved=curved, tau=1.0))
    t1 = _entropy_terms(prob1, _rhs_gpu(prob1))
    print(f"entropy production tau=0: {t0.sum():.3e} (relative {rel0:.2e}); tau=1: {t1.sum():.3e}",
          flush=True)
    assert rel0 < max(1e-12, 10 * rel_o)
    # du/dt = -2 M^-1 rhs, so a dissipative penalty makes sum wq v.rhs positive
    assert t1.sum() > 0 and abs(t1.sum()) > 1e3 * abs(t0.sum())
