"""Discrete adjoint of LSRK45 time stepping on the device, with checkpointing.

The forward map is the reference's time loop (euler.c:3672-3699) from a
conservative state Q0 at the volume quadrature points:

    rhsQ, Qf = u(VqPq v(Q0)), u(VfPq v(Q0))      (initial projection, euler.c:4028-4133)
    resQ = 0
    repeat nsteps x 5 stages: vol -> surf -> update

Its reverse sweep chains the hand-written VJP kernels (wadges_update_vjp,
wadges_rhs_vjp, wadges_project_vjp).  Memory follows the usual checkpointing
scheme: the forward pass keeps the full state every `checkpoint_every` steps;
the reverse pass re-runs each segment from its checkpoint, keeping the stage
states of that segment only, then walks the segment's stages backwards.

`RKRollout` wraps the forward map as a torch.autograd.Function whose backward is
this adjoint.
"""

from __future__ import annotations

import torch

from .device import DeviceSolver, _ptr
from .problem import RK4A, RK4B


class AdjointSolver:
    def __init__(self, dev: DeviceSolver, checkpoint_every=40, fused=True):
        self.dev = dev
        self.C = max(1, int(checkpoint_every))
        self.fused = fused
        s = dev.s
        self.sbar = torch.empty_like(s["Qf"])
        self.Rb = torch.empty_like(s["Q"])
        self.stats = {}

    # ------------------------------------------------------------------ forward
    def _project(self, Q0):
        """Initial projection rhsQ = u(VqPq v(Q0)), Qf = u(VfPq v(Q0)) on the host
        side of the driver, like the reference (euler.c:4028-4133)."""
        from .problem import conservative_vars, entropy_vars
        prob = self.dev.prob
        K, nf, Nq = prob.K, prob.nfields, prob.Nq
        U = Q0.detach().reshape(K, nf, Nq).permute(1, 2, 0)          # (nf, Nq, K)
        ops = prob.ops
        VqPq = torch.as_tensor(ops.VqPq, device=U.device)
        VfPq = torch.as_tensor(ops.VfPq, device=U.device)
        V = _entropy_vars_t(U, prob.gamma)
        uq = _conservative_vars_t(torch.einsum("ij,fjk->fik", VqPq, V), prob.gamma)
        uf = _conservative_vars_t(torch.einsum("ij,fjk->fik", VfPq, V), prob.gamma)
        lay = lambda X: X.permute(2, 0, 1).contiguous().reshape(-1)
        return lay(uq), lay(uf)

    def _set_initial(self, Q0):
        d = self.dev
        uq, uf = self._project(Q0)
        d.s["Q"].copy_(Q0.reshape(-1))
        d.s["resQ"].zero_()
        d.s["rhsQ"].copy_(uq)
        d.s["Qf"].copy_(uf)

    def _stage(self, s, dt):
        if self.fused:
            self.dev.stage_fused(s, dt)
        else:
            self.dev.stage(s, dt)

    def _snapshot(self):
        s = self.dev.s
        return {k: s[k].clone() for k in ("Q", "resQ", "rhsQ", "Qf")}

    def _restore(self, snap):
        for k, v in snap.items():
            self.dev.s[k].copy_(v)

    def forward(self, Q0, nsteps, dt=None):
        """Runs nsteps LSRK45 steps from Q0 (device tensor in the Q layout);
        keeps checkpoints; returns the final Q (a copy)."""
        dt = self.dev.prob.dt if dt is None else dt
        self.nsteps, self.dt = nsteps, dt
        self.Q0 = Q0.detach().clone()
        self._set_initial(self.Q0)
        self.ckpt = {}
        for step in range(nsteps):
            if step % self.C == 0:
                self.ckpt[step] = self._snapshot()
            for s in range(5):
                self._stage(s, dt)
        return self.dev.s["Q"].clone()

    # ------------------------------------------------------------------ reverse
    def backward(self, Q_bar_T):
        """Gradient of <Q_bar_T, Q(T)> with respect to Q0."""
        d, c, prob = self.dev, self.dev.c, self.dev.prob
        plan, K, st = d.plan, d.K, d.stream()
        Qb = Q_bar_T.detach().reshape(-1).clone()
        resb = torch.zeros_like(Qb)
        uqb = torch.zeros_like(d.s["rhsQ"])
        ufb = torch.zeros_like(d.s["Qf"])
        uqb_new = torch.zeros_like(uqb)
        ufb_new = torch.zeros_like(ufb)
        dt = self.dt
        for seg in sorted(self.ckpt, reverse=True):
            end = min(seg + self.C, self.nsteps)
            self._restore(self.ckpt[seg])
            states = [self._snapshot()]       # state before each stage, then the end state
            for step in range(seg, end):
                for s in range(5):
                    self._stage(s, dt)
                    states.append(self._snapshot())
            for idx in range(len(states) - 2, -1, -1):
                s = idx % 5
                out, inp = states[idx + 1], states[idx]
                plan.update_vjp(K, _ptr(c["Jq"]), _ptr(c["VqPq"]), _ptr(c["VfPq"]), RK4A[s], RK4B[s],
                                dt, _ptr(out["Q"]), _ptr(Qb), _ptr(resb), _ptr(uqb), _ptr(ufb),
                                _ptr(self.Rb), st)
                plan.rhs_vjp(K, _ptr(c["vgeo"]), _ptr(c["vfgeo"]), _ptr(c["fgeo"]), _ptr(c["mapPq"]),
                             _ptr(c["nrJ"]), _ptr(c["nsJ"]), _ptr(c["ntJ"]), _ptr(c["Drq"]),
                             _ptr(c["Dsq"]), _ptr(c["Dtq"]), _ptr(c["VqLq"]), _ptr(c["VfPq"]),
                             _ptr(inp["rhsQ"]), _ptr(inp["Qf"]), _ptr(self.Rb), _ptr(self.sbar),
                             _ptr(uqb_new), _ptr(ufb_new), st)
                uqb, uqb_new = uqb_new, uqb
                ufb, ufb_new = ufb_new, ufb
            del states
        plan.project_vjp(K, _ptr(c["VqPq"]), _ptr(c["VfPq"]), _ptr(self.Q0), _ptr(uqb), _ptr(ufb),
                         _ptr(Qb), st)
        return Qb


def _entropy_vars_t(U, gamma):
    rho, E = U[0], U[-1]
    m = U[1:-1]
    rhoe = E - 0.5 * (m * m).sum(0) / rho
    s = torch.log((gamma - 1.0) * rhoe) - gamma * torch.log(rho)
    return torch.cat([((-E) / rhoe + (gamma + 1.0 - s))[None], m / rhoe, (-rho / rhoe)[None]], 0)


def _conservative_vars_t(V, gamma):
    vl = V[-1]
    vv = (V[1:-1] ** 2).sum(0)
    h = vv / (2.0 * vl)
    s = gamma - V[0] + h
    rhoe = torch.exp((torch.log(torch.tensor(gamma - 1.0, dtype=V.dtype, device=V.device))
                      - gamma * torch.log(-vl) - s) / (gamma - 1.0))
    return torch.cat([(rhoe * (-vl))[None], rhoe * V[1:-1], (rhoe * (1.0 - h))[None]], 0)


class RKRollout(torch.autograd.Function):
    """Q0 -> Q(T) through nsteps LSRK45 steps of the device solver; the backward
    is the discrete adjoint (checkpointed)."""

    @staticmethod
    def forward(ctx, Q0, adj: AdjointSolver, nsteps: int):
        QT = adj.forward(Q0, nsteps)
        ctx.adj = adj
        return QT

    @staticmethod
    def backward(ctx, gQT):
        return ctx.adj.backward(gQT.contiguous()), None, None


def density_misfit(prob, QT, target_rho):
    """0.5 sum_k sum_i wq_i J_i (rho_i - rho_target_i)^2 (torch, differentiable)."""
    K, nf, Nq = prob.K, prob.nfields, prob.Nq
    wq = torch.as_tensor(prob.wq, device=QT.device)
    J = torch.as_tensor(prob.Jq, device=QT.device).reshape(K, Nq)
    rho = QT.reshape(K, nf, Nq)[:, 0, :]
    return 0.5 * torch.sum(wq[None, :] * J * (rho - target_rho.reshape(K, Nq)) ** 2)


def gaussian_bumps(xq, seed=0, count=4, amp=1e-2, width=1.0, box=5.0):
    """sum of `count` Gaussian bumps of amplitude `amp` and width `width` with
    centres uniform in [-box/2, box/2]^dim, at the points xq (dim x ...)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    dim = len(xq)
    out = np.zeros_like(xq[0])
    for _ in range(count):
        c = rng.uniform(-box / 2, box / 2, size=dim)
        r2 = sum((xq[d] - c[d]) ** 2 for d in range(dim))
        out = out + amp * np.exp(-r2 / (width * width))
    return out


def misfit_setup(prob, dev, nsteps, seed=0, amp=1e-2):
    """Target for the density misfit: the density after nsteps from the initial
    state whose density is multiplied by (1 + Gaussian bumps of amplitude amp),
    at fixed velocity and pressure.  Returns (Q0, Q0_perturbed, target_rho) as
    device tensors in the Q layout (rho layout for the target)."""
    import numpy as np
    from .problem import GAMMA, conservative_vars, entropy_vars  # noqa: F401
    K, nf, Nq = prob.K, prob.nfields, prob.Nq
    U = prob.Q.reshape(K, nf, Nq).copy()
    bump = gaussian_bumps(prob.xq, seed=seed, amp=amp).reshape(Nq, K).T      # (K, Nq)
    rho = U[:, 0, :]
    vel = U[:, 1:-1, :] / rho[:, None, :]
    p = (prob.gamma - 1.0) * (U[:, -1, :] - 0.5 * rho * (vel ** 2).sum(1))
    rho2 = rho * (1.0 + bump)
    U2 = np.concatenate([rho2[:, None], rho2[:, None] * vel,
                         (p / (prob.gamma - 1.0) + 0.5 * rho2 * (vel ** 2).sum(1))[:, None]], 1)
    Q0 = torch.as_tensor(prob.Q, device=dev.device)
    Q0p = torch.as_tensor(U2.reshape(-1), device=dev.device)
    adj = AdjointSolver(dev, checkpoint_every=nsteps + 1)
    QT = adj.forward(Q0p, nsteps)
    target = QT.reshape(K, nf, Nq)[:, 0, :].reshape(-1).clone()
    return Q0, Q0p, target
