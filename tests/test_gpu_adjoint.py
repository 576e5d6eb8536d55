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

@pytest.mark.parametrize("dim,N", [(2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (3, 1), (3, 2), (3, 3), (3, 4)])
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


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (3, 1), (3, 2), (3, 3), (3, 4)])
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


# ----------------------------------------------------------------- recorded stages, graphs

def test_stage_record_equals_stage_and_stage_vjp_equals_parts():
    """wadges_stage_record (out of place) reproduces wadges_stage bit for bit, and
    wadges_stage_vjp equals wadges_update_vjp followed by wadges_rhs_vjp."""
    from scratch_wadge_amd.device import _ptr
    from scratch_wadge_amd.problem import RK4A, RK4B
    prob = small_problem(3, 3)
    print("case stage record / stage vjp", flush=True)
    dev = _solver(prob)
    c, s = dev.c, dev.s
    rng = np.random.default_rng(8)
    dev.set_state(resQ=1e-3 * rng.standard_normal(prob.Q.shape))
    geo = [_ptr(c[n]) for n in ("vgeo", "vfgeo", "fgeo", "mapPq", "nrJ", "nsJ", "ntJ", "Drq", "Dsq",
                                "Dtq", "VqLq", "VqPq", "VfPq", "Jq")]
    ins = {k: v.clone() for k, v in s.items()}
    outs = {k: torch.empty_like(v) for k, v in s.items()}
    fa, fb, dt = RK4A[2], RK4B[2], prob.dt
    dev.plan.stage_record(prob.K, *geo, fa, fb, dt, _ptr(ins["Qf"]), _ptr(outs["Qf"]),
                          _ptr(ins["rhsQ"]), _ptr(outs["rhsQ"]), _ptr(ins["resQ"]), _ptr(outs["resQ"]),
                          _ptr(ins["Q"]), _ptr(outs["Q"]), dev.stream())
    dev.stage_fused(2, dt)
    torch.cuda.synchronize()
    for k in ("Q", "resQ", "rhsQ", "Qf"):
        assert torch.equal(outs[k], s[k]), k
    # stage VJP against its parts
    cot = {k: _dev(rng.standard_normal(v.numel())) for k, v in s.items()}
    Qb1, rb1 = cot["Q"].clone(), cot["resQ"].clone()
    uqb1, ufb1 = cot["rhsQ"].clone(), cot["Qf"].clone()
    Rb, sbar = torch.empty_like(Qb1), torch.empty_like(ufb1)
    dev.plan.stage_vjp(prob.K, *geo, fa, fb, dt, _ptr(ins["rhsQ"]), _ptr(ins["Qf"]), _ptr(outs["Q"]),
                       _ptr(Qb1), _ptr(rb1), _ptr(uqb1), _ptr(ufb1), _ptr(Rb), _ptr(sbar), _ptr(uqb1),
                       _ptr(ufb1), dev.stream())
    Qb2, rb2 = cot["Q"].clone(), cot["resQ"].clone()
    Rb2, sbar2 = torch.empty_like(Qb2), torch.empty_like(ufb1)
    uqb2, ufb2 = torch.empty_like(uqb1), torch.empty_like(ufb1)
    dev.plan.update_vjp(prob.K, _ptr(c["Jq"]), _ptr(c["VqPq"]), _ptr(c["VfPq"]), fa, fb, dt,
                        _ptr(outs["Q"]), _ptr(Qb2), _ptr(rb2), _ptr(cot["rhsQ"]), _ptr(cot["Qf"]),
                        _ptr(Rb2), dev.stream())
    dev.plan.rhs_vjp(prob.K, *geo[:10], _ptr(c["VqLq"]), _ptr(c["VfPq"]), _ptr(ins["rhsQ"]),
                     _ptr(ins["Qf"]), _ptr(Rb2), _ptr(sbar2), _ptr(uqb2), _ptr(ufb2), dev.stream())
    torch.cuda.synchronize()
    for a, b, name in ((Qb1, Qb2, "Q"), (rb1, rb2, "resQ"), (uqb1, uqb2, "rhsQ"), (ufb1, ufb2, "Qf")):
        e = relerr(a.cpu().numpy(), b.cpu().numpy())
        print("stage vjp vs parts", name, e, flush=True)
        assert e < 1e-14, name


def test_graph_paths_equal_eager():
    """Graph-replayed stepping and the graph-replayed adjoint give the results of
    the eager launches, bit for bit."""
    from scratch_wadge_amd.adjoint import AdjointSolver
    prob = small_problem(2, 3)
    print("case graph vs eager", flush=True)
    a, b = _solver(prob), _solver(prob)
    a.run(5, fused=True, graph=True)
    a.run(3, fused=True, graph=True)
    b.run(8, fused=True)
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    rng = np.random.default_rng(9)
    Q0 = _dev(prob.Q * (1.0 + 1e-3 * rng.standard_normal(prob.Q.shape)))
    Qbar = _dev(rng.standard_normal(prob.Q.shape))
    grads = []
    for use_graph in (True, False):
        adj = AdjointSolver(_solver(prob), checkpoint_every=3, use_graph=use_graph)
        adj.forward(Q0, 7)
        grads.append(adj.backward(Qbar).cpu().numpy())
    assert np.array_equal(grads[0], grads[1])


def test_misfit_gradient_fd_32_directions():
    """Density misfit of a 2D vortex (32x32x2 triangles, N=3) after 400 steps
    against a target grown from Gaussian density bumps of amplitude 1e-2;
    gradient with respect to the initial conservative state by the adjoint with
    checkpoints every 40 steps, against central differences on 32 seeded
    directions."""
    from scratch_wadge_amd import problem as P
    from scratch_wadge_amd.adjoint import AdjointSolver, RKRollout, density_misfit, misfit_setup
    steps = 400
    prob = P.vortex2d_problem(n=32, N=3)
    print(f"case misfit gradient K={prob.K} steps={steps}", flush=True)
    dev = _solver(prob)
    Q0, _, target = misfit_setup(prob, dev, steps)
    adj = AdjointSolver(dev, checkpoint_every=40)
    q = Q0.clone().requires_grad_(True)
    L = density_misfit(prob, RKRollout.apply(q, adj, steps), target)
    L.backward()
    g = q.grad.clone()
    print(f"loss {float(L.detach()):.6e} |grad| {float(torch.linalg.norm(g)):.6e}", flush=True)

    def loss(Qi):
        with torch.no_grad():
            return float(density_misfit(prob, adj.forward(Qi, steps), target))

    rng = np.random.default_rng(11)
    worst = 0.0
    for k in range(32):
        d = _dev(rng.standard_normal(prob.Q.shape) * np.abs(prob.Q))
        eps = 1e-6
        fd = (loss(Q0 + eps * d) - loss(Q0 - eps * d)) / (2 * eps)
        a = float(torch.dot(g, d))
        worst = max(worst, _rel(a, fd))
        print(f"direction {k}: adjoint {a:.12e} fd {fd:.12e} rel {_rel(a, fd):.2e}", flush=True)
    print("worst relative difference", worst, flush=True)
    assert worst < 1e-6
