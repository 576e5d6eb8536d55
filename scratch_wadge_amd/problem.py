"""Problem setup: meshes + operators + geometry + initial state, in kernel layouts.

Mirrors the host side of the reference driver (euler.c:app_new :2691-3002,
initial condition and projection :3926-4133, time step :4180-4193) for the
workloads this package runs:

* 2D isentropic vortex on a periodic square (euler_vortex, euler.c:3283-3322),
* 3D Taylor-Green vortex on a periodic cube (euler_Taylor_Green, euler.c:3381-3411),
  on affine or sinusoidally warped tetrahedra.

All arrays are float64 numpy in the layouts of the reference kernels.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import mesh as meshmod
from .operators import RefOps, build_ref_ops, colmajor

GAMMA = 1.4

# Carpenter-Kennedy 5-stage 4th-order low-storage RK (euler.c:2779-2796).
RK4A = (0.0,
        -567301805773.0 / 1357537059087.0,
        -2404267990393.0 / 2016746695238.0,
        -3550918686646.0 / 2091501179385.0,
        -1275806237668.0 / 842570457699.0)
RK4B = (1432997174477.0 / 9575080441755.0,
        5161836677717.0 / 13612068292357.0,
        1720146321549.0 / 2090206949498.0,
        3134564353537.0 / 4481467310338.0,
        2277821191437.0 / 14882151754819.0)


@dataclass
class Problem:
    dim: int
    N: int
    K: int
    ops: RefOps
    gamma: float
    tau: float
    # constant inputs (kernel layouts)
    vgeo: np.ndarray
    vfgeo: np.ndarray
    fgeo: np.ndarray
    Jq: np.ndarray
    mapPq: np.ndarray
    # constant operators, column-major flattened
    Drq: np.ndarray
    Dsq: np.ndarray
    Dtq: np.ndarray | None
    VqLq: np.ndarray
    VqPq: np.ndarray
    VfPq: np.ndarray
    nrJ: np.ndarray
    nsJ: np.ndarray
    ntJ: np.ndarray | None
    wq: np.ndarray
    # state
    Q: np.ndarray
    rhsQ: np.ndarray
    Qf: np.ndarray
    resQ: np.ndarray
    xq: np.ndarray = None
    dt: float = 0.0
    hmin: float = 0.0
    meta: dict = field(default_factory=dict)

    @property
    def Nq(self):
        return self.ops.Nq

    @property
    def Nfq(self):
        return self.ops.Nfq

    @property
    def Nf(self):
        return self.ops.Nf

    @property
    def nfields(self):
        return self.dim + 2

    @property
    def Np(self):
        return self.ops.Np

    def copy_state(self):
        return {k: getattr(self, k).copy() for k in ("Q", "rhsQ", "Qf", "resQ")}


# ------------------------------------------------------------------ pointwise maps (vectorised)

def entropy_vars(U, gamma=GAMMA):
    """Conservative -> entropy variables along axis 0 (euler.c:3216-3248)."""
    rho, E = U[0], U[-1]
    m = U[1:-1]
    rhoe = E - 0.5 * np.sum(m * m, 0) / rho
    s = np.log((gamma - 1.0) * rhoe / rho ** gamma)
    V = np.empty_like(U)
    V[0] = (-E) / rhoe + (gamma + 1.0 - s)
    V[1:-1] = m / rhoe
    V[-1] = -rho / rhoe
    return V


def conservative_vars(V, gamma=GAMMA):
    """Entropy -> conservative variables along axis 0 (euler.c:3250-3278)."""
    vl = V[-1]
    vv = np.sum(V[1:-1] ** 2, 0)
    s = gamma - V[0] + vv / (2.0 * vl)
    rhoe = ((gamma - 1.0) / (-vl) ** gamma) ** (1.0 / (gamma - 1.0)) * np.exp(-s / (gamma - 1.0))
    U = np.empty_like(V)
    U[0] = rhoe * (-vl)
    U[1:-1] = rhoe * V[1:-1]
    U[-1] = rhoe * (1.0 - vv / (2.0 * vl))
    return U


def vortex_2d(x, y, t=0.0, x0=0.0, y0=0.0, beta=5.0, u0=1.0, v0=1.0, gamma=GAMMA):
    """Isentropic vortex (euler.c:3291-3305) with centre (x0,y0), drift (u0,v0)."""
    xt = x - x0 - u0 * t
    yt = y - y0 - v0 * t
    r2 = xt * xt + yt * yt
    ex = np.exp(1.0 - r2)
    u = u0 - beta * ex * yt / (2.0 * math.pi)
    v = v0 + beta * ex * xt / (2.0 * math.pi)
    rho0 = 1.0 - 0.5 * (gamma - 1.0) * (1.0 / (8.0 * gamma * math.pi ** 2)) * (beta * ex) ** 2
    rho = rho0 ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], 0)


def vortex_3d(x, y, z, t=0.0, gamma=GAMMA):
    """The reference's 3D isentropic vortex on [0,10] x [0,20] x [0,10]
    (euler_vortex, euler.c:3324-3351): the 2D vortex extruded along z, centred
    at (5, 5), moving with unit speed along y; Pi_max = 0.4, p0 = 1/gamma.  The
    energy is written as in euler.c:3345-3346, p0/(gamma-1) (1 + tmp^gamma) +
    |m|^2/(2 rho): the pressure p0 (1 + tmp^gamma) is the vortex's p0 tmp^gamma
    plus a constant p0, which leaves it an exact solution (a uniform pressure
    changes no force)."""
    gm1 = gamma - 1.0
    rx = -(y - 5.0 - t)
    ry = x - 5.0
    r2 = rx * rx + ry * ry
    p0 = 1.0 / gamma
    L = 0.4 * np.exp(0.5 * (1.0 - r2))
    tmp = (1.0 - 0.5 * gm1 * L * L) ** (1.0 / gm1)
    rho = tmp
    u = rx * L
    v = 1.0 + ry * L
    w = 0.0 * x
    E = p0 / gm1 * (1.0 + tmp ** gamma) + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, E], 0)


def vortex_warp_2d(x, y, a=0.5):
    """Curvilinear warp of [0,20] x [-5,5] (operators.cpp:374-383)."""
    xx, yy = x / 20.0, (y + 5.0) / 10.0
    return (x + 2.0 * a * np.sin(np.pi * xx) * np.sin(2.0 * np.pi * yy),
            y - a * np.sin(2.0 * np.pi * xx) * np.sin(np.pi * yy))


def vortex_warp_3d(x, y, z, a=0.5):
    """Curvilinear warp of [0,10] x [0,20] x [0,10] (operators.cpp:504-515)."""
    xx, yy, zz = x / 10.0, y / 20.0, z / 10.0
    sa = np.sin(np.pi * xx) * np.sin(2.0 * np.pi * yy) * np.sin(np.pi * zz)
    return (x + a * sa,
            y - 2.0 * a * np.sin(2.0 * np.pi * xx) * np.sin(np.pi * yy) * np.sin(2.0 * np.pi * zz),
            z + a * sa)


def taylor_green(x, y, z, gamma=GAMMA):
    """Taylor-Green vortex, Mach 0.1 (euler.c:3384-3411)."""
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = 0.0 * x
    p = 100.0 / gamma + (1.0 / 16.0) * (np.cos(2.0 * x) + np.cos(2.0 * y)) * (2.0 + np.cos(2.0 * z))
    rho = 1.0 + 0.0 * x
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, E], 0)


# ------------------------------------------------------------------ setup

def _to_layout(U):
    """(nfields, Nq, K) -> [K][fld][Nq] flattened."""
    return np.ascontiguousarray(U.transpose(2, 0, 1)).reshape(-1)


def project_initial_state(ops: RefOps, Uq, gamma=GAMMA):
    """Projection of the initial state (euler.c:3958-4133).

    Q = VqPq U;  rhsQ = u(VqPq v(Q));  Qf = u(VfPq v(Q)).  Uq: (nfields, Nq, K)."""
    Q = np.einsum("ij,fjk->fik", ops.VqPq, Uq)
    V = entropy_vars(Q, gamma)
    Vq = np.einsum("ij,fjk->fik", ops.VqPq, V)
    Vf = np.einsum("ij,fjk->fik", ops.VfPq, V)
    return Q, conservative_vars(Vq, gamma), conservative_vars(Vf, gamma)


def time_step(dim, N, Jq, fgeo, cfl):
    """dt = CFL * hmin / CN with hmin = min_k min(J)/max(sJ) (euler.c:3177-3213, :4180-4193)."""
    sJ = fgeo[:, dim, :]
    hmin = float(np.min(np.min(Jq, axis=1) / np.max(sJ, axis=1)))
    CN = (N + 1) * (N + 2) / 2.0 if dim == 2 else (N + 1) * (N + 3) / 3.0
    return cfl * hmin / CN, hmin


def make_problem(dim, N, mesh, ic, warp=None, gamma=GAMMA, tau=1.0, cfl=0.5, state=None):
    """Assemble a Problem.  ic(xq) -> (nfields, Nq, K) conservative state at quadrature points."""
    ops = build_ref_ops(dim, N)
    geo = meshmod.build_geometry(mesh, ops, warp=warp)
    mapPq, nmatched = meshmod.build_mapPq(mesh, ops, xf=geo.xf)
    Uq = ic(*geo.xq) if state is None else state
    Q, Uvq, Uvf = project_initial_state(ops, Uq, gamma)
    dt, hmin = time_step(dim, N, geo.Jq, geo.fgeo, cfl)
    nref = ops.nref
    return Problem(
        dim=dim, N=N, K=mesh.K, ops=ops, gamma=gamma, tau=tau,
        vgeo=geo.vgeo.reshape(-1), vfgeo=geo.vfgeo.reshape(-1), fgeo=geo.fgeo.reshape(-1),
        Jq=geo.Jq.reshape(-1), mapPq=mapPq.reshape(-1),
        Drq=colmajor(ops.Drst[0]), Dsq=colmajor(ops.Drst[1]),
        Dtq=colmajor(ops.Drst[2]) if dim == 3 else None,
        VqLq=colmajor(ops.VqLq), VqPq=colmajor(ops.VqPq), VfPq=colmajor(ops.VfPq),
        nrJ=np.ascontiguousarray(nref[:, 0]), nsJ=np.ascontiguousarray(nref[:, 1]),
        ntJ=np.ascontiguousarray(nref[:, 2]) if dim == 3 else None,
        wq=ops.wq.copy(),
        Q=_to_layout(Q), rhsQ=_to_layout(Uvq), Qf=_to_layout(Uvf),
        resQ=np.zeros(mesh.K * ops.Nq * (dim + 2)),
        xq=geo.xq, dt=dt, hmin=hmin,
        meta={"matched_face_points": nmatched, "geometry": geo, "mesh": mesh})


def varying_state(prob: Problem, seed=0, amp=0.3):
    """A smooth but strongly varying state on prob's mesh: density and pressure
    1 + amp sin(k.x + phase) and velocities of 0.5 (wave numbers and phases
    from seed), projected like the initial condition.  Neighbouring nodes
    differ by more than the log mean's series threshold, so both branches of
    the log mean are taken.  Returns (Q, rhsQ, Qf) in kernel layout."""
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
    Q, Uq, Uf = project_initial_state(prob.ops, U, prob.gamma)
    if not (np.isfinite(Uq).all() and np.isfinite(Uf).all()):
        raise ValueError("the varying state left the admissible set")
    return _to_layout(Q), _to_layout(Uq), _to_layout(Uf)


# ------------------------------------------------------------------ named workloads

def vortex2d_problem(n=16, N=3, perturb=0.05, seed=1, beta=5.0, x0=0.0, y0=0.0, tau=1.0,
                     cfl=0.5, u0=1.0, v0=1.0):
    """2D isentropic vortex on [-5,5]^2, n x n squares x 2 triangles, perturbed vertices."""
    m = meshmod.periodic_box_tri(n, -5.0, 5.0, perturb=perturb, seed=seed)
    ic = lambda x, y: vortex_2d(x, y, 0.0, x0, y0, beta, u0, v0)
    return make_problem(2, N, m, ic, tau=tau, cfl=cfl)


def rk_steps(dt, final_time):
    """Step count and step of rk_run (euler.c:3678-3683): ceil(T/dt) steps of T/steps."""
    n = int(math.ceil(final_time / dt))
    return n, final_time / n


def vortex_reference_problem(mesh_path, N, curved=False, cfl=0.5, tau=1.0):
    """The reference's vortex convergence study on one of its gmsh meshes
    (docs/main.tex:1622-1632, :1766-1795): 2D, the vortex of euler.c:3290-3305
    (centre (5, 0), beta = 5, unit speed along x) on [0,20] x [-5,5]; 3D, the
    vortex of euler.c:3324-3351 on [0,10] x [0,20] x [0,10]; curved=True
    applies the warp of operators.cpp:374-383 / :504-515 (a = 1/2)."""
    import os
    dim = 3 if "periodicCube" in os.path.basename(mesh_path) else 2
    m = meshmod.mesh_from_gmsh(mesh_path, dim)
    if dim == 2:
        ic = lambda x, y: vortex_2d(x, y, 0.0, 5.0, 0.0, 5.0, 1.0, 0.0)
        warp = vortex_warp_2d if curved else None
    else:
        ic = lambda x, y, z: vortex_3d(x, y, z, 0.0)
        warp = vortex_warp_3d if curved else None
    return make_problem(dim, N, m, ic, warp=warp, tau=tau, cfl=cfl)


def vortex_reference_exact(prob, t):
    """The exact vortex of vortex_reference_problem at time t at the volume
    quadrature points, in the Q layout."""
    if prob.dim == 2:
        x, y = prob.xq
        return _to_layout(vortex_2d(x, y, t, 5.0, 0.0, 5.0, 1.0, 0.0))
    x, y, z = prob.xq
    return _to_layout(vortex_3d(x, y, z, t))


def tgv3d_problem(n=16, N=3, curved=False, warp_amp=0.1, tau=1.0, cfl=0.05):
    """3D Taylor-Green vortex on [-pi,pi]^3, n^3 cubes x 6 tets, optionally warped.

    The reference's step rule dt = CFL hmin / CN (euler.c:4185-4192) has no wave
    speed in it; with the Mach-0.1 Taylor-Green state (sound speed ~10) the
    prefs default CFL 0.25 is unstable within a few steps, so the default here
    is 0.05."""
    m = meshmod.periodic_box_tet(n)
    warp = (lambda x, y, z: meshmod.tg_warp(x, y, z, warp_amp)) if curved else None
    return make_problem(3, N, m, taylor_green, warp=warp, tau=tau, cfl=cfl)
