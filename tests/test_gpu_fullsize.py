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


@pytest.mark.parametrize("n,N,curved", CASES)
def test_full_size_rough_state_flat_bar(oracle, n, N, curved):
    """A well-conditioned state at full size: density and pressure varying by
    30 % and velocities of 0.5 (both log-mean branches taken), where the RHS
    is O(1) rather than a cancellation residual.  Held to a flat 1e-12
    relative L2 against the oracle, the RHS and the face terms."""
    from helpers import rough_state
    prob = P.tgv3d_problem(n=n, N=N, curved=curved)
    _, rq, qf = rough_state(prob, seed=7, amp=0.3)
    print(f"case rough full size n={n} N={N} curved={curved} K={prob.K}", flush=True)
    from scratch_wadge_amd.device import DeviceSolver
    dev = DeviceSolver(prob)
    r_g, rf_g = dev.vol_only(rq, qf)
    r_gs, _ = dev.rhs(rq, qf)
    del dev
    torch.cuda.empty_cache()
    r_o = rq.copy()
    rf_o = np.zeros_like(qf)
    oracle.vol(prob, r_o, qf, rf_o)
    e_vol, e_f = relerr(r_g, r_o), relerr(rf_g, rf_o)
    oracle.surf(prob, qf, rf_o, r_o)
    e = relerr(r_gs, r_o)
    print(f"rel L2 vs oracle: volume {e_vol:.2e}, face terms {e_f:.2e}, vol+surf {e:.2e}",
          flush=True)
    assert e_vol <= 1e-12 and e_f <= 1e-12 and e <= 1e-12


def _oracle_steps(oracle, prob, steps):
    st = prob.copy_state()
    for _ in range(steps):
        for s in range(5):
            oracle.stage(prob, st, P.RK4A[s], P.RK4B[s], prob.dt)
    return st


@pytest.mark.parametrize("n,curved,steps", [(16, False, 100), (32, True, 10)],
                         ids=["affine K=24576 100 steps", "curved K=196608 10 steps"])
def test_full_size_trajectory_against_oracle(oracle, n, curved, steps):
    """The Taylor-Green workloads run for their step counts on the headline
    path (fused stage, graph-replayed steps) against the oracle run through the
    same steps (the reference's rk_run loop, euler.c:3691-3699): the state Q,
    the projected state and the face traces agree to 1e-12 relative L2."""
    from scratch_wadge_amd.device import DeviceSolver
    prob = P.tgv3d_problem(n=n, N=3, curved=curved)
    print(f"case trajectory n={n} curved={curved} K={prob.K} steps={steps}", flush=True)
    dev = DeviceSolver(prob)
    dev.run(steps, fused=True, graph=True)
    g = dev.get_state()          # Qf: the current face-trace buffer
    del dev
    torch.cuda.empty_cache()
    ref = _oracle_steps(oracle, prob, steps)
    errs = {f: relerr(g[f], ref[f]) for f in ("Q", "rhsQ", "Qf")}
    print("rel L2 vs oracle: " + " ".join(f"{f} {e:.2e}" for f, e in errs.items()), flush=True)
    assert max(errs.values()) <= 1e-12
