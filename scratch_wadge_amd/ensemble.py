"""Ensembles of independent problems advanced together (wadges_batch_*).

The reference advances one mesh per process (rk_run, euler.c:3672-3699).  An
ensemble here is a list of problems of one dimension with uneven K and N; each
LSRK stage of all of them is ONE launch of the ensemble kernel, and a whole
LSRK45 step can be replayed from a HIP graph.  The per-problem arithmetic is
that of wadges_stage, so every problem's trajectory is the one its own
DeviceSolver.run(fused=True) produces.

`PerProblem` is the comparison point: the same problems, one wadges_stage
launch per problem per stage.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .device import DeviceSolver
from .problem import vortex2d_problem


def vortex_ensemble(count=64, seed=0, kmin=128, kmax=8192, degrees=(2, 3, 4), strength=(3.0, 6.0),
                    centre=2.0):
    """Independent 2D isentropic vortex problems on [-5,5]^2: K log-uniform in
    [kmin, kmax] rounded to the nearest 2 n^2 (n x n squares, two triangles
    each), N uniform over `degrees`, vortex strength uniform in `strength`,
    centre uniform in [-centre, centre]^2, all drawn from `seed`."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        k = np.exp(rng.uniform(np.log(kmin), np.log(kmax)))
        n = max(1, int(round(np.sqrt(k / 2.0))))
        N = int(rng.choice(degrees))
        beta = float(rng.uniform(*strength))
        x0, y0 = rng.uniform(-centre, centre, size=2)
        out.append(vortex2d_problem(n=n, N=N, beta=beta, x0=float(x0), y0=float(y0),
                                    seed=int(rng.integers(1 << 30))))
    return out


def group_of(sv: DeviceSolver):
    """wadges_group describing one DeviceSolver's arrays (Qf[0] = its current
    face traces, Qf[1] = its second buffer)."""
    p, c, s = sv.prob, sv.c, sv.s
    g = _lib.Group()
    g.N, g.Nq, g.Nfq, g.K, g.dt = p.N, p.Nq, p.Nfq, p.K, p.dt
    for name in ("vgeo", "vfgeo", "fgeo", "mapPq", "nrJ", "nsJ", "ntJ", "Drq", "Dsq", "Dtq",
                 "VqLq", "VqPq", "VfPq", "Jq"):
        setattr(g, name, c[name].data_ptr() if c.get(name) is not None else None)
    for name in ("Q", "rhsQ", "resQ"):
        setattr(g, name, s[name].data_ptr())
    g.Qf[0] = s["Qf"].data_ptr()
    g.Qf[1] = sv.Qf_alt.data_ptr()
    return g


class Ensemble:
    def __init__(self, problems, device="cuda:0"):
        if not problems:
            raise _lib.WadgesError("an ensemble needs at least one problem")
        dims = {p.dim for p in problems}
        if len(dims) != 1:
            raise _lib.WadgesError("all problems of an ensemble must have the same dimension")
        gam = {p.gamma for p in problems}
        tau = {p.tau for p in problems}
        if len(gam) != 1 or len(tau) != 1:
            raise _lib.WadgesError("all problems of an ensemble must share gamma and tau")
        self.dim = dims.pop()
        self.problems = problems
        self.solvers = [DeviceSolver(p, device) for p in problems]
        self.device = self.solvers[0].device
        groups = [group_of(sv) for sv in self.solvers]
        self._qf = [(sv.s["Qf"], sv.Qf_alt) for sv in self.solvers]   # fixed buffer roles
        self.batch = _lib.Batch(self.dim, groups, problems[0].gamma, problems[0].tau, self.stream())
        self.K_total = sum(p.K for p in problems)
        self.dofs = sum(p.K * p.Np * p.nfields for p in problems)   # DOFs incl. fields

    def stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def tune(self, reps=5):
        """Runs the autotuner; returns the mean stage time (ms) of the chosen
        launch shape, or -1 when it came from the cache."""
        return self.batch.tune(reps, self.stream())

    def config(self):
        return self.batch.config()

    def set_config(self, **kw):
        self.batch.set_config(**kw)

    def stage(self, intrk, dt_unused=None):
        from .problem import RK4A, RK4B
        self.batch.stage(RK4A[intrk], RK4B[intrk], self.stream())

    def run(self, nsteps, use_graph=True):
        self.batch.run(nsteps, use_graph, self.stream())

    def state(self, q):
        """Host copy of problem q: Q, rhsQ, resQ and the current face traces Qf."""
        torch.cuda.synchronize(self.device)
        s = self.solvers[q].s
        cur = self.batch.current()
        return {"Q": s["Q"].cpu().numpy(), "rhsQ": s["rhsQ"].cpu().numpy(),
                "resQ": s["resQ"].cpu().numpy(), "Qf": self._qf[q][cur].cpu().numpy()}

    def close(self):
        self.batch.close()


class PerProblem:
    """The same problems advanced one wadges_stage launch per problem per stage."""

    def __init__(self, problems, device="cuda:0"):
        self.problems = problems
        self.solvers = [DeviceSolver(p, device) for p in problems]
        self.device = self.solvers[0].device

    def _steps(self, nsteps):
        for _ in range(nsteps):
            for s in range(5):
                for sv in self.solvers:
                    sv.stage_fused(s, sv.prob.dt)

    def run(self, nsteps, use_graph=False):
        """nsteps LSRK45 steps.  use_graph replays a torch-captured HIP graph of
        two steps (ten stages bring every face-trace buffer back to its role);
        an odd remainder step runs eagerly."""
        if not use_graph:
            self._steps(nsteps)
            return
        if not hasattr(self, "_graph"):
            side = torch.cuda.Stream(self.device)
            side.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(side):
                self._steps(2)          # warm-up outside capture (prepares tables)
            torch.cuda.current_stream(self.device).wait_stream(side)
            torch.cuda.synchronize(self.device)
            self._graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self._graph):
                self._steps(2)
            torch.cuda.synchronize(self.device)
            self._warm_steps = 2
        for _ in range(nsteps // 2):
            self._graph.replay()
        if nsteps % 2:
            self._steps(1)

    def state(self, q):
        torch.cuda.synchronize(self.device)
        return {k: v.cpu().numpy() for k, v in self.solvers[q].s.items()}
