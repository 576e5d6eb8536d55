"""Device-resident problem and the RK driver over the C ABI.

The reference's rk_step / rk_run (euler.c:3610-3738) as host code: constant
arrays are uploaded once, the state lives in HBM, and each LSRK stage is
vol -> surf -> update launched asynchronously on one HIP stream.  PyTorch is
used only to allocate device memory and to get the current stream; all
arithmetic runs in libwadges.so.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .problem import RK4A, RK4B, Problem


def _dev(a, device, dtype=torch.float64):
    if a is None:
        return None
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dtype)


def _ptr(t):
    return None if t is None else t.data_ptr()


class DeviceSolver:
    CONST = ("vgeo", "vfgeo", "fgeo", "Jq", "Drq", "Dsq", "Dtq", "VqLq", "VqPq", "VfPq", "nrJ",
             "nsJ", "ntJ")
    STATE = ("Q", "rhsQ", "Qf", "resQ")

    def __init__(self, prob: Problem, device="cuda:0"):
        if not torch.cuda.is_available():
            raise _lib.WadgesError("the WADGES kernels need a HIP device (torch.cuda is unavailable)")
        self.prob = prob
        self.device = torch.device(device)
        self.K = prob.K
        nq_tot = prob.K * prob.Nq * prob.nfields
        nf_tot = prob.K * prob.Nf * prob.nfields
        sizes = {"vgeo": prob.K * prob.Nq * prob.dim ** 2, "vfgeo": prob.K * prob.Nf * prob.dim ** 2,
                 "fgeo": prob.K * prob.Nf * (prob.dim + 1), "Jq": prob.K * prob.Nq,
                 "Q": nq_tot, "rhsQ": nq_tot, "resQ": nq_tot, "Qf": nf_tot}
        for name, n in sizes.items():
            if getattr(prob, name).size != n:
                raise _lib.WadgesError(f"{name} has {getattr(prob, name).size} entries, expected {n}")
        m = np.asarray(prob.mapPq)
        if m.size != prob.K * prob.Nf or (m.size and int(m.max()) + (prob.nfields - 1) * prob.Nf >= nf_tot):
            raise _lib.WadgesError("mapPq does not index inside Qf")
        self.plan = _lib.Plan(prob.dim, prob.N, prob.Nq, prob.Nfq, prob.gamma, prob.tau)
        self.c = {k: _dev(getattr(prob, k), self.device) for k in self.CONST}
        self.c["mapPq"] = torch.from_numpy(prob.mapPq.astype(np.int32)).to(self.device)
        self.s = {k: _dev(getattr(prob, k), self.device) for k in self.STATE}
        self.rhsQf = torch.zeros_like(self.s["Qf"])
        self.Qf_alt = torch.zeros_like(self.s["Qf"])   # second face-trace buffer (fused stage)

    # -------------------------------------------------------------- state
    def set_state(self, **arrays):
        for k, v in arrays.items():
            self.s[k].copy_(torch.as_tensor(np.ascontiguousarray(v), dtype=torch.float64))

    def get_state(self):
        torch.cuda.synchronize(self.device)
        return {k: v.cpu().numpy() for k, v in self.s.items()}

    def stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    # -------------------------------------------------------------- kernels
    def vol(self, rhsQ=None, Qf=None, rhsQf=None):
        c = self.c
        rhsQ = self.s["rhsQ"] if rhsQ is None else rhsQ
        Qf = self.s["Qf"] if Qf is None else Qf
        rhsQf = self.rhsQf if rhsQf is None else rhsQf
        self.plan.vol(self.K, _ptr(c["vgeo"]), _ptr(c["vfgeo"]), _ptr(c["nrJ"]), _ptr(c["nsJ"]),
                      _ptr(c["ntJ"]), _ptr(c["Drq"]), _ptr(c["Dsq"]), _ptr(c["Dtq"]), None,
                      _ptr(c["VqLq"]), _ptr(c["VfPq"]), _ptr(self.s["Q"]), _ptr(Qf), _ptr(rhsQ),
                      _ptr(rhsQf), self.stream())

    def surf(self, rhsQ=None, Qf=None, rhsQf=None):
        c = self.c
        rhsQ = self.s["rhsQ"] if rhsQ is None else rhsQ
        Qf = self.s["Qf"] if Qf is None else Qf
        rhsQf = self.rhsQf if rhsQf is None else rhsQf
        self.plan.surf(self.K, _ptr(c["fgeo"]), _ptr(c["mapPq"]), _ptr(c["VqLq"]), _ptr(Qf),
                       _ptr(rhsQf), _ptr(rhsQ), self.stream())

    def update(self, fa, fb, dt):
        c, s = self.c, self.s
        self.plan.update(self.K, _ptr(c["Jq"]), _ptr(c["VqPq"]), _ptr(c["VfPq"]), fa, fb, dt,
                         _ptr(s["rhsQ"]), _ptr(s["resQ"]), _ptr(s["Q"]), _ptr(s["Qf"]), self.stream())

    def stage(self, intrk, dt):
        """One LSRK stage as the reference runs it: vol -> surf -> update."""
        self.vol()
        self.surf()
        self.update(RK4A[intrk], RK4B[intrk], dt)

    def stage_fused(self, intrk, dt):
        """One LSRK stage in one launch (wadges_stage); swaps the face-trace buffers."""
        c, s = self.c, self.s
        self.plan.stage(self.K, _ptr(c["vgeo"]), _ptr(c["vfgeo"]), _ptr(c["fgeo"]), _ptr(c["mapPq"]),
                        _ptr(c["nrJ"]), _ptr(c["nsJ"]), _ptr(c["ntJ"]), _ptr(c["Drq"]), _ptr(c["Dsq"]),
                        _ptr(c["Dtq"]), _ptr(c["VqLq"]), _ptr(c["VqPq"]), _ptr(c["VfPq"]), _ptr(c["Jq"]),
                        RK4A[intrk], RK4B[intrk], dt, _ptr(s["Qf"]), _ptr(self.Qf_alt), _ptr(s["rhsQ"]),
                        _ptr(s["resQ"]), _ptr(s["Q"]), self.stream())
        s["Qf"], self.Qf_alt = self.Qf_alt, s["Qf"]

    def step(self, dt, fused=False):
        for intrk in range(5):
            (self.stage_fused if fused else self.stage)(intrk, dt)

    def run(self, nsteps, dt=None, fused=False, graph=False):
        """nsteps LSRK45 steps.  graph=True replays a HIP graph of two steps
        (captured through torch.cuda.graph at first use; ten stages bring the
        double-buffered face traces back to their roles), an odd last step runs
        directly.  The graph bakes in dt and the state buffers."""
        dt = self.prob.dt if dt is None else dt
        if not graph:
            for _ in range(nsteps):
                self.step(dt, fused)
            return
        if not hasattr(self, "_graphs"):
            self._graphs = {}
        key = (float(dt), bool(fused), self.s["Qf"].data_ptr())   # which face buffer is current
        if nsteps >= 2 and key not in self._graphs:
            self.step(dt, fused)              # tables prepared outside the capture
            self.step(dt, fused)
            nsteps -= 2
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self.step(dt, fused)
                self.step(dt, fused)
            self._graphs[key] = g
        for _ in range(nsteps // 2):
            self._graphs[key].replay()
        if nsteps % 2:
            self.step(dt, fused)

    def diagnostics(self, Qexact=None):
        """(kinetic energy, entropy, L2 error, mass) of the current state, reduced on
        the device; Qexact in the Q layout (host or device array) for the error."""
        if not hasattr(self, "_wq"):
            self._wq = _dev(self.prob.wq, self.device)
            self._aux = torch.empty(self.K * 4, dtype=torch.float64, device=self.device)
            self._tot = torch.empty(4, dtype=torch.float64, device=self.device)
        qex = None if Qexact is None else torch.as_tensor(Qexact, dtype=torch.float64).to(self.device)
        self.plan.diagnostics(self.K, _ptr(self._wq), _ptr(self.c["Jq"]), _ptr(self.s["Q"]),
                              _ptr(qex), _ptr(self._aux), _ptr(self._tot), self.stream())
        ke, ent, err2, mass = self._tot.cpu().tolist()
        return {"kinetic_energy": ke, "entropy": ent, "l2_error": err2 ** 0.5, "mass": mass}

    def rhs(self, rhsQ_in, Qf):
        """Volume + surface RHS of a given projected state; returns host arrays."""
        r = _dev(rhsQ_in, self.device)
        qf = _dev(Qf, self.device)
        rf = torch.zeros_like(qf)
        self.vol(r, qf, rf)
        self.surf(r, qf, rf)
        torch.cuda.synchronize(self.device)
        return r.cpu().numpy(), rf.cpu().numpy()

    def vol_only(self, rhsQ_in, Qf):
        r = _dev(rhsQ_in, self.device)
        qf = _dev(Qf, self.device)
        rf = torch.zeros_like(qf)
        self.vol(r, qf, rf)
        torch.cuda.synchronize(self.device)
        return r.cpu().numpy(), rf.cpu().numpy()
