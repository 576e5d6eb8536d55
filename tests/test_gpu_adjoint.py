"""Discrete adjoint on the GPU: each VJP kernel against the reference gradient
(torch autograd through the torch restatement in torch_ref.py, pinned to the
oracle) and against central finite differences of the matching forward kernel;
the gradient of a density misfit through LSRK45 steps against both."""

import numpy as np
import pytest
import torch

import torch_ref as TR
from helpers import relerr, small_problem

pytestmark = pytest.mark.gpu

EPS = 1e-6
# The update outputs are O(100) (total energy) against O(1e-2) directional
# derivatives, so a smaller step there is dominated by round-off.
EPS_UPDATE = 1e-5


def _solver(prob):
    from scratch_wadge_amd.device import DeviceSolver
    return DeviceSolver(prob)


def _dev(x):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda:0")


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 3), (2, 4)])
def test_rhs_vjp_dot(dim, N):
    from scratch_wadge_amd.device import _ptr
    prob = small_problem(dim, N)
    print(f"case rhs vjp dim={dim} N={N} K={prob.K}", flush=True)
    dev = _solver(prob)
    rng = np.random.default_rng(0)
    # Independent 1% noise on volume and face states: traces jump across every
    # face, so max(lamM, lamP) of the penalty is far from its kink at the
    # finite-difference step (a smooth state has lamM ~ lamP there).
    uq = prob.rhsQ * (1.0 + 1e-2 * rng.standard_normal(prob.rhsQ.shape))
    uf = prob.Qf * (1.0 + 1e-2 * rng.standard_normal(prob.Qf.shape))
    Rb = rng.standard_normal(uq.shape)
    dq = rng.standard_normal(uq.shape) * np.abs(uq)
    df = rng.standard_normal(uf.shape) * np.abs(uf)
    # adjoint side
    c = dev.c
    Rbd, sbar = _dev(Rb), torch.empty_like(dev.s["Qf"])
    uqb, ufb = torch.empty_like(dev.s["rhsQ"]), torch.empty_like(dev.s["Qf"])
    uqd, ufd = _dev(uq), _dev(uf)
    dev.plan.rhs_vjp(prob.K, _ptr(c["vgeo"]), _ptr(c["vfgeo"]), _ptr(c["fgeo"]), _ptr(c["mapPq"]),
                     _ptr(c["nrJ"]), _ptr(c["nsJ"]), _ptr(c["ntJ"]), _ptr(c["Drq"]), _ptr(c["Dsq"]),
                     _ptr(c["Dtq"]), _ptr(c["VqLq"]), _ptr(c["VfPq"]), _ptr(uqd), _ptr(ufd),
                     _ptr(Rbd), _ptr(sbar), _ptr(uqb), _ptr(ufb), dev.stream())
    torch.cuda.synchronize()
    adj = float(np.dot(uqb.cpu().numpy(), dq) + np.dot(ufb.cpu().numpy(), df))
    rp, _ = dev.rhs(uq + EPS * dq, uf + EPS * df)
    rm, _ = dev.rhs(uq - EPS * dq, uf - EPS * df)
    fd = float(np.dot(Rb, (rp - rm) / (2 * EPS)))
    print("rhs vjp: adjoint", adj, "fd", fd, "rel", _rel(adj, fd), flush=True)
    assert _rel(adj, fd) < 1e-6


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 3)])
def test_update_vjp_dot(dim, N):
    from scratch_wadge_amd.device import _ptr
    from scratch_wadge_amd.problem import RK4A, RK4B
    prob = small_problem(dim, N)
    print(f"case update vjp dim={dim} N={N} K={prob.K}", flush=True)
    dev = _solver(prob)
    rng = np.random.default_rng(1)
    Q = prob.Q.copy()
    R = 0.1 * rng.standard_normal(Q.shape)
    res = 1e-3 * rng.standard_normal(Q.shape)
    fa, fb, dt = RK4A[3], RK4B[3], 2e-3
    # output cotangents
    Qb_o = rng.standard_normal(Q.shape)
    rb_o = rng.standard_normal(Q.shape)
    uqb_o = rng.standard_normal(Q.shape)
    ufb_o = rng.standard_normal(prob.Qf.shape)
    # directions for the inputs (Q, res, R)
    dQ = rng.standard_normal(Q.shape) * np.abs(Q) * 1e-3
    dr = rng.standard_normal(Q.shape) * 1e-3
    dR = rng.standard_normal(Q.shape) * 1e-1

    def fwd(Qi, resi, Ri):
        dev.set_state(Q=Qi, resQ=resi, rhsQ=Ri, Qf=prob.Qf)
        dev.update(fa, fb, dt)
        g = dev.get_state()
        return g["Q"], g["resQ"], g["rhsQ"], g["Qf"]

    Qo, ro, uqo, ufo = fwd(Q, res, R)
    c = dev.c
    Qb, rb = _dev(Qb_o), _dev(rb_o)
    Rb = torch.empty_like(Qb)
    Qod, uqbd, ufbd = _dev(Qo), _dev(uqb_o), _dev(ufb_o)    # held until the kernel has run
    dev.plan.update_vjp(prob.K, _ptr(c["Jq"]), _ptr(c["VqPq"]), _ptr(c["VfPq"]), fa, fb, dt,
                        _ptr(Qod), _ptr(Qb), _ptr(rb), _ptr(uqbd), _ptr(ufbd), _ptr(Rb), dev.stream())
    torch.cuda.synchronize()
    adj = float(np.dot(Qb.cpu().numpy(), dQ) + np.dot(rb.cpu().numpy(), dr) + np.dot(Rb.cpu().numpy(), dR))
    e = EPS_UPDATE
    p = fwd(Q + e * dQ, res + e * dr, R + e * dR)
    m = fwd(Q - e * dQ, res - e * dr, R - e * dR)
    cots = (Qb_o, rb_o, uqb_o, ufb_o)
    fd = float(sum(np.dot(cb, (a - b) / (2 * e)) for cb, a, b in zip(cots, p, m)))
    print("update vjp: adjoint", adj, "fd", fd, "rel", _rel(adj, fd), flush=True)
    assert _rel(adj, fd) < 1e-6


@pytest.mark.parametrize("dim,N,steps,ckpt", [(2, 3, 6, 4), (3, 3, 3, 2)])
def test_trajectory_gradient_fd(dim, N, steps, ckpt):
    """Gradient of the density misfit after several LSRK45 steps with respect to
    the initial conservative state, against central differences on seeded
    directions."""
    from scratch_wadge_amd.adjoint import AdjointSolver, RKRollout, density_misfit
    prob = small_problem(dim, N)
    print(f"case trajectory dim={dim} N={N} K={prob.K} steps={steps}", flush=True)
    dev = _solver(prob)
    adj = AdjointSolver(dev, checkpoint_every=ckpt)
    rng = np.random.default_rng(3)
    # Per-node noise keeps the traces discontinuous (see test_rhs_vjp_dot), so
    # that the finite differences, with a step well below the jumps, do not
    # straddle the kink of max(lamM, lamP).
    Q0 = _dev(prob.Q * (1.0 + 1e-2 * rng.standard_normal(prob.Q.shape)))
    target = _dev(prob.Q.reshape(prob.K, prob.nfields, prob.Nq)[:, 0, :].reshape(-1)
                  * (1 + 1e-2 * rng.standard_normal(prob.K * prob.Nq)))
    q = Q0.clone().requires_grad_(True)
    L = density_misfit(prob, RKRollout.apply(q, adj, steps), target)
    L.backward()
    g = q.grad.cpu().numpy()

    def loss(Qi):
        with torch.no_grad():
            return float(density_misfit(prob, adj.forward(Qi, steps), target))

    worst = 0.0
    for k in range(4):
        d = rng.standard_normal(prob.Q.shape) * np.abs(prob.Q)
        eps = 1e-7
        fd = (loss(Q0 + eps * _dev(d)) - loss(Q0 - eps * _dev(d))) / (2 * eps)
        worst = max(worst, _rel(float(g @ d), fd))
        print("trajectory: adjoint", float(g @ d), "fd", fd, flush=True)
    assert worst < 1e-6


# ----------------------------------------------------------------- reference gradient

@pytest.mark.parametrize("dim,N", [(2, 2), (2, 3), (3, 3), (3, 4)])
def test_rhs_vjp_matches_reference_gradient(dim, N):
    from scratch_wadge_amd.device import _ptr
    prob = small_problem(dim, N)
    print(f"case rhs vjp vs autograd dim={dim} N={N} K={prob.K}", flush=True)
    dev = _solver(prob)
    rng = np.random.default_rng(5)
    uq = prob.rhsQ * (1.0 + 1e-2 * rng.standard_normal(prob.rhsQ.shape))
    uf = prob.Qf * (1.0 + 1e-2 * rng.standard_normal(prob.Qf.shape))
    Rb = rng.standard_normal(uq.shape)
    c = dev.c
    uqd, ufd, Rbd = _dev(uq), _dev(uf), _dev(Rb)
    sbar = torch.empty_like(ufd)
    uqb, ufb = torch.empty_like(uqd), torch.empty_like(ufd)
    dev.plan.rhs_vjp(prob.K, _ptr(c["vgeo"]), _ptr(c["vfgeo"]), _ptr(c["fgeo"]), _ptr(c["mapPq"]),
                     _ptr(c["nrJ"]), _ptr(c["nsJ"]), _ptr(c["ntJ"]), _ptr(c["Drq"]), _ptr(c["Dsq"]),
                     _ptr(c["Dtq"]), _ptr(c["VqLq"]), _ptr(c["VfPq"]), _ptr(uqd), _ptr(ufd),
                     _ptr(Rbd), _ptr(sbar), _ptr(uqb), _ptr(ufb), dev.stream())
    ops = TR.Ops(prob, "cuda:0")
    a = uqd.clone().requires_grad_(True)
    b = ufd.clone().requires_grad_(True)
    (TR.rhs(ops, a, b) * Rbd).sum().backward()
    torch.cuda.synchronize()
    e1 = relerr(uqb.cpu().numpy(), a.grad.cpu().numpy())
    e2 = relerr(ufb.cpu().numpy(), b.grad.cpu().numpy())
    print(f"rhs vjp vs autograd: volume {e1:.2e} face {e2:.2e}", flush=True)
    assert e1 < 1e-11 and e2 < 1e-11


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 3)])
def test_update_vjp_matches_reference_gradient(dim, N):
    from scratch_wadge_amd.device import _ptr
    from scratch_wadge_amd.problem import RK4A, RK4B
    prob = small_problem(dim, N)
    print(f"case update vjp vs autograd dim={dim} N={N}", flush=True)
    dev = _solver(prob)
    rng = np.random.default_rng(6)
    Q = prob.Q.copy()
    R = 0.1 * rng.standard_normal(Q.shape)
    res = 1e-3 * rng.standard_normal(Q.shape)
    fa, fb, dt = RK4A[1], RK4B[1], 2e-3
    cots = [rng.standard_normal(Q.shape) for _ in range(3)] + [rng.standard_normal(prob.Qf.shape)]
    dev.set_state(Q=Q, resQ=res, rhsQ=R, Qf=prob.Qf)
    dev.update(fa, fb, dt)
    Qo = dev.s["Q"].clone()
    c = dev.c
    Qb, rb = _dev(cots[0]), _dev(cots[1])
    uqb, ufb = _dev(cots[2]), _dev(cots[3])
    Rb = torch.empty_like(Qb)
    dev.plan.update_vjp(prob.K, _ptr(c["Jq"]), _ptr(c["VqPq"]), _ptr(c["VfPq"]), fa, fb, dt,
                        _ptr(Qo), _ptr(Qb), _ptr(rb), _ptr(uqb), _ptr(ufb), _ptr(Rb), dev.stream())
    ops = TR.Ops(prob, "cuda:0")
    Qt, rt, Rt = (_dev(x).requires_grad_(True) for x in (Q, res, R))
    outs = TR.update(ops, fa, fb, dt, Rt, rt, Qt)
    sum((o * _dev(cb)).sum() for o, cb in zip(outs, cots)).backward()
    torch.cuda.synchronize()
    errs = [relerr(k.cpu().numpy(), t.grad.cpu().numpy()) for k, t in ((Qb, Qt), (rb, rt), (Rb, Rt))]
    print("update vjp vs autograd:", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) < 1e-12


@pytest.mark.parametrize("dim,N,steps,ckpt", [(2, 3, 4, 3), (3, 3, 2, 1)])
def test_trajectory_matches_reference_gradient(dim, N, steps, ckpt):
    """The checkpointed discrete adjoint equals autograd through the torch
    restatement of the same LSRK45 steps."""
    from scratch_wadge_amd.adjoint import AdjointSolver, RKRollout, density_misfit
    from scratch_wadge_amd.problem import RK4A, RK4B
    prob = small_problem(dim, N)
    print(f"case trajectory vs autograd dim={dim} N={N} steps={steps}", flush=True)
    dev = _solver(prob)
    adj = AdjointSolver(dev, checkpoint_every=ckpt)
    rng = np.random.default_rng(7)
    Q0 = _dev(prob.Q * (1.0 + 1e-2 * rng.standard_normal(prob.Q.shape)))
    target = _dev(prob.Q.reshape(prob.K, prob.nfields, prob.Nq)[:, 0, :].reshape(-1))
    q = Q0.clone().requires_grad_(True)
    density_misfit(prob, RKRollout.apply(q, adj, steps), target).backward()
    ops = TR.Ops(prob, "cuda:0")
    qt = Q0.clone().requires_grad_(True)
    density_misfit(prob, TR.run(ops, qt, steps, prob.dt, RK4A, RK4B), target).backward()
    torch.cuda.synchronize()
    e = relerr(q.grad.cpu().numpy(), qt.grad.cpu().numpy())
    print(f"trajectory gradient vs autograd: {e:.2e}", flush=True)
    assert e < 1e-10
