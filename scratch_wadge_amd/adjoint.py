"""Discrete adjoint of LSRK45 time stepping on the device, with checkpointing.

The forward map is the reference's time loop (euler.c:3672-3699) from a
conservative state Q0 at the volume quadrature points:

    rhsQ, Qf = u(VqPq v(Q0)), u(VfPq v(Q0))      (initial projection, euler.c:4028-4133)
    resQ = 0
    repeat nsteps x 5 stages: vol -> surf -> update

Its reverse sweep chains the hand-written VJP kernels (wadges_stage_vjp per
stage, wadges_project_vjp at the end).  Memory follows the usual checkpointing
scheme: the forward pass keeps the full state every `checkpoint_every` steps;
the reverse pass re-runs each segment from its checkpoint with
wadges_stage_record, which writes every stage's state into its own slot (no
copies), then walks the segment's stages backwards.  The last segment is
recorded into the slots by the forward pass itself, so the reverse pass
recomputes all segments but that one.

`RKRollout` wraps the forward map as a torch.autograd.Function whose backward is
this adjoint.
"""

from __future__ import annotations

import torch

from .device import DeviceSolver, _ptr
from .problem import RK4A, RK4B


class AdjointSolver:
    """Checkpointed discrete adjoint of LSRK45 on one DeviceSolver.

    Memory: the forward keeps (Q, resQ, rhsQ, Qf) every `checkpoint_every`
    steps; the reverse pass keeps one segment's stage states, (5C + 1) x
    (|Q| + |rhsQ| + |Qf|) doubles for C = checkpoint_every.  With use_graph the
    forward steps and each segment's recomputation + reverse sweep are replayed
    from HIP graphs (captured through torch.cuda.graph on first use)."""

    def __init__(self, dev: DeviceSolver, checkpoint_every=40, fused=True, use_graph=True):
        self.dev = dev
        self.C = max(1, int(checkpoint_every))
        self.fused = fused
        self.use_graph = use_graph
        s = dev.s
        self.sbar = torch.empty_like(s["Qf"])
        self.Rb = torch.empty_like(s["Q"])
        self.Qb = torch.empty_like(s["Q"])
        self.resb = torch.empty_like(s["Q"])
        self.uqb = torch.empty_like(s["rhsQ"])
        self.ufb = torch.empty_like(s["Qf"])
        self.res_work = torch.empty_like(s["Q"])
        self._slots = None
        self._seg_graphs = {}
        self._recorded = None       # first step of the segment the slots hold
        self.stats = {}

    # ------------------------------------------------------------------ forward
    def _project(self, Q0):
        """Initial projection rhsQ = u(VqPq v(Q0)), Qf = u(VfPq v(Q0)) on the host
        side of the driver, like the reference (euler.c:4028-4133)."""
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

    def _snapshot(self):
        s = self.dev.s
        return {k: s[k].clone() for k in ("Q", "resQ", "rhsQ", "Qf")}

    def forward(self, Q0, nsteps, dt=None):
        """Runs nsteps LSRK45 steps from Q0 (device tensor in the Q layout);
        keeps checkpoints; returns the final Q (a copy).  The last segment's
        stages are recorded into the reverse pass's slots as they are computed
        (wadges_stage_record): the reverse pass starts with that segment and
        needs no recomputation for it."""
        dt = self.dev.prob.dt if dt is None else dt
        self.nsteps, self.dt = nsteps, dt
        self.Q0 = Q0.detach().clone()
        self._set_initial(self.Q0)
        self.ckpt = {}
        self._recorded = None
        for seg in range(0, nsteps, self.C):
            self.ckpt[seg] = self._snapshot()
            n = min(self.C, nsteps - seg)
            if seg + self.C >= nsteps and self.fused:
                self._record(5 * n, dt)
                self._recorded = seg
            else:
                self.dev.run(n, dt, fused=self.fused, graph=self.use_graph)
        return self.dev.s["Q"].clone()

    def _record(self, nst, dt):
        """nst stages from the solver's state, each stage's state written to its
        own slot; the solver's state then holds the last one."""
        S = self._ensure_slots()
        s = self.dev.s
        S["Q"][0].copy_(s["Q"])
        S["rhsQ"][0].copy_(s["rhsQ"])
        S["Qf"][0].copy_(s["Qf"])
        self.res_work.copy_(s["resQ"])
        self._run_segment(nst, dt, reverse=False)
        s["Q"].copy_(S["Q"][nst])
        s["rhsQ"].copy_(S["rhsQ"][nst])
        s["Qf"].copy_(S["Qf"][nst])
        s["resQ"].copy_(self.res_work)

    # ------------------------------------------------------------------ reverse
    def _ensure_slots(self):
        if self._slots is None:
            L = 5 * self.C + 1
            s = self.dev.s
            mk = lambda t: torch.empty((L, t.numel()), dtype=t.dtype, device=t.device)
            self._slots = {"Q": mk(s["Q"]), "rhsQ": mk(s["rhsQ"]), "Qf": mk(s["Qf"])}
        return self._slots

    def _segment(self, nst, dt, forward=True, reverse=True):
        """Recompute nst stages from slot 0 into the slots (forward), then walk
        them back (reverse)."""
        d, c = self.dev, self.dev.c
        plan, K, st = d.plan, d.K, d.stream()
        S = self._slots
        P = lambda n: _ptr(c[n])
        geo = (P("vgeo"), P("vfgeo"), P("fgeo"), P("mapPq"), P("nrJ"), P("nsJ"), P("ntJ"), P("Drq"),
               P("Dsq"), P("Dtq"), P("VqLq"), P("VqPq"), P("VfPq"), P("Jq"))
        res = _ptr(self.res_work)
        for idx in range(nst if forward else 0):
            s = idx % 5
            plan.stage_record(K, *geo, RK4A[s], RK4B[s], dt, _ptr(S["Qf"][idx]), _ptr(S["Qf"][idx + 1]),
                              _ptr(S["rhsQ"][idx]), _ptr(S["rhsQ"][idx + 1]), res, res,
                              _ptr(S["Q"][idx]), _ptr(S["Q"][idx + 1]), st)
        uqb, ufb = _ptr(self.uqb), _ptr(self.ufb)
        for idx in range(nst - 1 if reverse else -1, -1, -1):
            s = idx % 5
            plan.stage_vjp(K, *geo, RK4A[s], RK4B[s], dt, _ptr(S["rhsQ"][idx]), _ptr(S["Qf"][idx]),
                           _ptr(S["Q"][idx + 1]), _ptr(self.Qb), _ptr(self.resb), uqb, ufb,
                           _ptr(self.Rb), _ptr(self.sbar), uqb, ufb, st)

    def _run_segment(self, nst, dt, forward=True, reverse=True):
        if not self.use_graph:
            self._segment(nst, dt, forward, reverse)
            return
        key = (nst, dt, forward, reverse)
        g = self._seg_graphs.get(key)
        if g is None:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._segment(nst, dt, forward, reverse)
            self._seg_graphs[key] = g
        g.replay()

    def backward(self, Q_bar_T):
        """Gradient of <Q_bar_T, Q(T)> with respect to Q0."""
        d, c = self.dev, self.dev.c
        S = self._ensure_slots()
        self.Qb.copy_(Q_bar_T.detach().reshape(-1))
        self.resb.zero_()
        self.uqb.zero_()
        self.ufb.zero_()
        for seg in sorted(self.ckpt, reverse=True):
            nst = 5 * (min(seg + self.C, self.nsteps) - seg)
            if seg == self._recorded:
                # the forward recorded this segment's stages into the slots
                self._run_segment(nst, self.dt, forward=False)
                continue
            self._recorded = None          # the slots now hold another segment
            ck = self.ckpt[seg]
            S["Q"][0].copy_(ck["Q"])
            S["rhsQ"][0].copy_(ck["rhsQ"])
            S["Qf"][0].copy_(ck["Qf"])
            self.res_work.copy_(ck["resQ"])
            self._run_segment(nst, self.dt)
        Qb = self.Qb.clone()
        d.plan.project_vjp(d.K, _ptr(c["VqPq"]), _ptr(c["VfPq"]), _ptr(self.Q0), _ptr(self.uqb),
                           _ptr(self.ufb), _ptr(Qb), d.stream())
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
