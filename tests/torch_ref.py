"""Differentiable restatement of the WADGES stage in torch float64 (test
infrastructure only: the reference gradient the hand-written VJP kernels are
checked against).

It follows the same reference code as the C++ oracle (oracle/wadges_oracle.cpp),
vectorised over elements and node pairs so that torch.autograd can
differentiate it:

  logmean        okl/Euler3D.okl:57-68 (log branch as intended), okl/Euler2D.okl:54-66
  ec_flux        okl/Euler3D.okl:128-162, okl/Euler2D.okl:139-161
  volume term    okl/Euler3D.okl:2123-2523, okl/Euler2D.okl:1112-1350
  surface term   okl/Euler3D.okl:3242-3473, okl/Euler2D.okl:1357-1495
  update         okl/Euler3D.okl:2527-2884, okl/Euler2D.okl:1729-2282

Its forward values are pinned against the oracle by tests/test_torch_ref.py.
Both branches of the log mean are evaluated on safe arguments and selected with
torch.where, so the branch not taken contributes neither NaN nor gradient.
"""

from __future__ import annotations

import torch


def _t(x, like):
    return torch.as_tensor(x, dtype=torch.float64, device=like.device)


def logmean(dim, a, b):
    xi = a / b
    f = (xi - 1.0) / (xi + 1.0)
    u = f * f
    u2 = u * u
    if dim == 3:
        small = u.abs() < 1e-4
        F = 105.0 + 35.0 * u + 21.0 * u2 + 15.0 * u2 * u
        ser = 52.5 * (a + b) / F
        fs = torch.where(small, torch.ones_like(f), f)
        xis = torch.where(small, torch.full_like(xi, 2.0), xi)
        lg = (a + b) * fs / torch.log(xis)
        return torch.where(small, ser, lg)
    small = u.abs() < 1e-3
    Fser = 1.0 + u / 3.0 + u2 / 5.0 + u2 * u / 7.0 + u2 * u2 / 9.0
    fs = torch.where(small, torch.ones_like(f), f)
    xis = torch.where(small, torch.full_like(xi, 2.0), xi)
    Flog = torch.log(xis) / 2.0 / fs
    return (a + b) / (2.0 * torch.where(small, Fser, Flog))


def prim(dim, U, gamma):
    """U (..., NF) -> rho, vel (..., dim), beta."""
    rho = U[..., 0]
    vel = U[..., 1:1 + dim] / rho[..., None]
    E = U[..., dim + 1]
    uu = (vel * vel).sum(-1)
    p = (gamma - 1.0) * (E - 0.5 * rho * uu)
    beta = rho / (2.0 * p)
    return rho, vel, beta, p, uu


def ec_flux(dim, A, B, gamma):
    """Two-point flux of broadcastable primitive tuples; returns (..., dim, NF)."""
    ra, va, ba = A[0], A[1], A[2]
    rb, vb, bb = B[0], B[1], B[2]
    rholog = logmean(dim, ra, rb)
    betalog = logmean(dim, ba, bb)
    rhoavg = 0.5 * (ra + rb)
    vavg = 0.5 * (va + vb)
    vnavg = 2.0 * (vavg * vavg).sum(-1) - 0.5 * ((va * va).sum(-1) + (vb * vb).sum(-1))
    pa = rhoavg / (ba + bb)
    f4 = rholog / (2.0 * (gamma - 1.0) * betalog) + pa + 0.5 * rholog * vnavg
    cols = []
    for d in range(dim):
        m = rholog * vavg[..., d]
        row = [m] + [m * vavg[..., e] + (pa if e == d else 0.0) for e in range(dim)] + [f4 * vavg[..., d]]
        cols.append(torch.stack(row, -1))
    return torch.stack(cols, -2)


class Ops:
    """Device tensors of one problem for the torch restatement."""

    def __init__(self, prob, device):
        dim, K, Nq, Nf, Nfq = prob.dim, prob.K, prob.Nq, prob.Nf, prob.Nfq
        T = lambda a: torch.as_tensor(a, dtype=torch.float64, device=device)
        o = prob.ops
        self.dim, self.K, self.Nq, self.Nf, self.Nfq, self.nf = dim, K, Nq, Nf, Nfq, dim + 2
        self.gamma, self.tau = prob.gamma, prob.tau
        self.D = torch.stack([T(D) for D in o.Drst], 0)                               # (dim,Nq,Nq)
        self.VqLq, self.VqPq, self.VfPq = T(o.VqLq), T(o.VqPq), T(o.VfPq)
        face = torch.arange(Nf, device=device) // Nfq
        self.nref = T(o.nref)[face * Nfq]                                             # (Nf, dim)
        self.vgeo = T(prob.vgeo).reshape(K, dim, dim, Nq).permute(0, 3, 1, 2)         # (K,Nq,r,d)
        self.vfgeo = T(prob.vfgeo).reshape(K, dim, dim, Nf).permute(0, 3, 1, 2)       # (K,Nf,r,d)
        fg = T(prob.fgeo).reshape(K, dim + 1, Nf)
        self.nJ, self.sJ = fg[:, :dim, :].permute(0, 2, 1), fg[:, dim, :]             # (K,Nf,dim),(K,Nf)
        self.mapPq = torch.as_tensor(prob.mapPq.astype("int64"), device=device).reshape(K, Nf)
        self.J = T(prob.Jq).reshape(K, 1, Nq)


def vol(ops, uq, uf):
    """Volume term: (rhsQ (K,NF,Nq), rhsQf (K,NF,Nf)) from projected states."""
    dim, g = ops.dim, ops.gamma
    Uq = uq.reshape(ops.K, ops.nf, ops.Nq).transpose(1, 2)          # (K,Nq,NF)
    Uf = uf.reshape(ops.K, ops.nf, ops.Nf).transpose(1, 2)
    pq = prim(dim, Uq, g)[:3]
    pf = prim(dim, Uf, g)[:3]
    ex = lambda P, ax: tuple(x.unsqueeze(ax) for x in P)
    # volume-volume pairs
    F = ec_flux(dim, ex(pq, 2), ex(pq, 1), g)                        # (K,Nq,Nq,dim,NF)
    Ga = 0.5 * (ops.vgeo[:, :, None] + ops.vgeo[:, None, :])         # (K,Nq,Nq,r,d)
    Dr = ops.D.permute(1, 2, 0)                                       # (Nq,Nq,r)
    coef = torch.einsum("ijr,kijrd->kijd", Dr, Ga)
    rq = torch.einsum("kijd,kijdf->kif", coef, F)
    # volume-face pairs, both blocks
    Fvf = ec_flux(dim, ex(pq, 2), ex(pf, 1), g)                      # (K,Nq,Nf,dim,NF)
    Gvf = 0.5 * (ops.vgeo[:, :, None] + ops.vfgeo[:, None, :])       # (K,Nq,Nf,r,d)
    nvf = torch.einsum("jr,kijrd->kijd", ops.nref, Gvf)              # face j's reference normal
    rq = rq + torch.einsum("ij,kijd,kijdf->kif", 0.5 * ops.VqLq, nvf, Fvf)
    nfv = torch.einsum("ir,kjird->kjid", ops.nref, Gvf)              # (K,Nq,Nf,d) normal of face i
    rf = torch.einsum("ij,kjid,kjidf->kif", -0.5 * ops.VfPq, nfv, Fvf)
    return rq.transpose(1, 2), rf.transpose(1, 2)


def surf(ops, uf, rf):
    """Surface term lifted to the volume rows: (K,NF,Nq)."""
    dim, g = ops.dim, ops.gamma
    UfK = uf.reshape(ops.K, ops.nf, ops.Nf)
    flat = uf.reshape(-1)
    idx = ops.mapPq[:, None, :] + ops.Nf * torch.arange(ops.nf, device=uf.device)[None, :, None]
    UPK = flat[idx]                                                  # (K,NF,Nf)
    UM, UP = UfK.transpose(1, 2), UPK.transpose(1, 2)                # (K,Nf,NF)
    rM, vM, bM, pM, uuM = prim(dim, UM, g)
    rP, vP, bP, pP, uuP = prim(dim, UP, g)
    lamM = torch.sqrt(uuM) + torch.sqrt(g * pM / rM)
    lamP = torch.sqrt(uuP) + torch.sqrt(g * pP / rP)
    F = ec_flux(dim, (rM, vM, bM), (rP, vP, bP), g)                  # (K,Nf,dim,NF)
    Fn = torch.einsum("kid,kidf->kif", ops.nJ, F)
    Lfc = 0.25 * ops.tau * torch.maximum(lamM, lamP) * ops.sJ
    sfS = 0.5 * Fn + rf.transpose(1, 2) - Lfc[..., None] * (UP - UM)
    return torch.einsum("ij,kjf->kfi", ops.VqLq, sfS)


def rhs(ops, uq, uf):
    rq, rf = vol(ops, uq, uf)
    return (rq + surf(ops, uf, rf)).reshape(-1)


def entropy_vars(dim, U, gamma):
    """U (..., NF) -> V (..., NF) (okl/Euler3D.okl:91-98, okl/Euler2D.okl:73-79)."""
    rho, E = U[..., 0], U[..., -1]
    m = U[..., 1:-1]
    rhoe = E - 0.5 * (m * m).sum(-1) / rho
    s = torch.log((gamma - 1.0) * rhoe) - gamma * torch.log(rho)
    return torch.cat([((-E) / rhoe + (gamma + 1.0 - s))[..., None], m / rhoe[..., None],
                      (-rho / rhoe)[..., None]], -1)


def conservative_vars(dim, V, gamma):
    """V (..., NF) -> U (..., NF) (okl/Euler3D.okl:110-120, okl/Euler2D.okl:89-98)."""
    vl = V[..., -1]
    vv = (V[..., 1:-1] ** 2).sum(-1)
    h = vv / (2.0 * vl)
    sV = gamma - V[..., 0] + h
    lg = torch.log(torch.tensor(gamma - 1.0, dtype=V.dtype, device=V.device))
    rhoe = torch.exp((lg - gamma * torch.log(-vl) - sV) / (gamma - 1.0))
    return torch.cat([(rhoe * (-vl))[..., None], rhoe[..., None] * V[..., 1:-1],
                      (rhoe * (1.0 - h))[..., None]], -1)


def update(ops, fa, fb, dt, R, res, Q):
    """Returns (Q', res', rhsQ' = u(VqPq v~), Qf' = u(VfPq v~)), flattened."""
    K, nf, Nq, g = ops.K, ops.nf, ops.Nq, ops.gamma
    P = ops.VqPq
    Rk = R.reshape(K, nf, Nq)
    a = torch.einsum("ij,kfj->kfi", -2.0 * P, Rk) / ops.J
    r = torch.einsum("ij,kfj->kfi", P, a)
    res2 = fa * res.reshape(K, nf, Nq) + dt * r
    U = Q.reshape(K, nf, Nq) + fb * res2
    V = entropy_vars(ops.dim, U.transpose(1, 2), g).transpose(1, 2) * ops.J
    a2 = torch.einsum("ij,kfj->kfi", P, V) / ops.J
    vt = torch.einsum("ij,kfj->kfi", P, a2)
    uq = conservative_vars(ops.dim, torch.einsum("ij,kfj->kif", P, vt), g).transpose(1, 2)
    uf = conservative_vars(ops.dim, torch.einsum("ij,kfj->kif", ops.VfPq, vt), g).transpose(1, 2)
    return U.reshape(-1), res2.reshape(-1), uq.reshape(-1), uf.reshape(-1)


def project(ops, Q):
    """Initial projection rhsQ = u(VqPq v(Q)), Qf = u(VfPq v(Q)) (euler.c:4028-4133)."""
    K, nf, Nq, g = ops.K, ops.nf, ops.Nq, ops.gamma
    V = entropy_vars(ops.dim, Q.reshape(K, nf, Nq).transpose(1, 2), g)      # (K,Nq,NF)
    uq = conservative_vars(ops.dim, torch.einsum("ij,kjf->kif", ops.VqPq, V), g)
    uf = conservative_vars(ops.dim, torch.einsum("ij,kjf->kif", ops.VfPq, V), g)
    return uq.transpose(1, 2).reshape(-1), uf.transpose(1, 2).reshape(-1)


def run(ops, Q0, nsteps, dt, rk4a, rk4b):
    """LSRK45 from Q0 with the initial projection; returns the final Q."""
    uq, uf = project(ops, Q0)
    Q, res = Q0, torch.zeros_like(Q0)
    for _ in range(nsteps):
        for s in range(5):
            R = rhs(ops, uq, uf)
            Q, res, uq, uf = update(ops, rk4a[s], rk4b[s], dt, R, res, Q)
    return Q
