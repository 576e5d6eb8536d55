"""Orthonormal polynomial bases and Gauss rules on the reference simplices.

Reference triangle: vertices (-1,-1), (1,-1), (-1,1)  (area 2).
Reference tetrahedron: vertices (-1,-1,-1), (1,-1,-1), (-1,1,-1), (-1,-1,1)
(volume 4/3).

These are the textbook Jacobi/Dubiner constructions (Hesthaven & Warburton,
"Nodal Discontinuous Galerkin Methods", 2008, sections 3.1, 6.1 and 10.1),
the same mathematical objects the reference builds with Eigen in
operators.cpp (JacobiP, JacobiGQ :1717, Simplex2DP, Simplex3DP, Vandermonde2D/3D).
They are setup-time host code; nothing here runs on the hot path.
"""

from __future__ import annotations

import math

import numpy as np


def jacobi_p(x, alpha, beta, n):
    """Orthonormal Jacobi polynomial P_n^{(alpha,beta)} evaluated at x."""
    x = np.asarray(x, dtype=np.float64)
    pl = np.zeros((n + 1,) + x.shape)
    gamma0 = (2.0 ** (alpha + beta + 1) / (alpha + beta + 1)
              * math.gamma(alpha + 1) * math.gamma(beta + 1)
              / math.gamma(alpha + beta + 1))
    pl[0] = 1.0 / math.sqrt(gamma0)
    if n == 0:
        return pl[0]
    gamma1 = (alpha + 1) * (beta + 1) / (alpha + beta + 3) * gamma0
    pl[1] = ((alpha + beta + 2) * x / 2 + (alpha - beta) / 2) / math.sqrt(gamma1)
    if n == 1:
        return pl[1]
    aold = 2.0 / (2 + alpha + beta) * math.sqrt(
        (alpha + 1) * (beta + 1) / (alpha + beta + 3))
    for i in range(1, n):
        h1 = 2 * i + alpha + beta
        anew = 2.0 / (h1 + 2) * math.sqrt(
            (i + 1) * (i + 1 + alpha + beta) * (i + 1 + alpha) * (i + 1 + beta)
            / (h1 + 1) / (h1 + 3))
        bnew = -(alpha * alpha - beta * beta) / h1 / (h1 + 2)
        pl[i + 1] = 1.0 / anew * (-aold * pl[i - 1] + (x - bnew) * pl[i])
        aold = anew
    return pl[n]


def grad_jacobi_p(x, alpha, beta, n):
    x = np.asarray(x, dtype=np.float64)
    if n == 0:
        return np.zeros_like(x)
    return math.sqrt(n * (n + alpha + beta + 1)) * jacobi_p(x, alpha + 1, beta + 1, n - 1)


def jacobi_gq(n, alpha=0.0, beta=0.0):
    """Gauss-Jacobi nodes and weights with n+1 points (Golub-Welsch)."""
    if n == 0:
        return np.array([-(alpha - beta) / (alpha + beta + 2.0)]), np.array([2.0])
    h1 = 2.0 * np.arange(n + 1) + alpha + beta
    J = np.diag(-0.5 * (alpha ** 2 - beta ** 2) / (h1 + 2) / np.where(h1 == 0, 1, h1))
    i = np.arange(1, n + 1, dtype=np.float64)
    off = 2.0 / (h1[:-1] + 2) * np.sqrt(
        i * (i + alpha + beta) * (i + alpha) * (i + beta) / (h1[:-1] + 1) / (h1[:-1] + 3))
    J = J + np.diag(off, 1)
    if alpha + beta < 1e-14:
        J[0, 0] = 0.0
    J = J + J.T
    lam, V = np.linalg.eigh(J)
    w = V[0, :] ** 2 * 2 ** (alpha + beta + 1) / (alpha + beta + 1) * math.gamma(
        alpha + 1) * math.gamma(beta + 1) / math.gamma(alpha + beta + 1)
    return lam, w


# ---------------------------------------------------------------- triangle

def rs_to_ab(r, s):
    r = np.asarray(r, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    a = np.where(np.abs(s - 1) > 1e-14, 2 * (1 + r) / np.where(np.abs(s - 1) > 1e-14, 1 - s, 1) - 1, -1.0)
    return a, s


def simplex2d_p(a, b, i, j):
    h1 = jacobi_p(a, 0, 0, i)
    h2 = jacobi_p(b, 2 * i + 1, 0, j)
    return math.sqrt(2.0) * h1 * h2 * (1 - b) ** i


def grad_simplex2d_p(a, b, i, j):
    fa = jacobi_p(a, 0, 0, i)
    dfa = grad_jacobi_p(a, 0, 0, i)
    gb = jacobi_p(b, 2 * i + 1, 0, j)
    dgb = grad_jacobi_p(b, 2 * i + 1, 0, j)
    dmodr = dfa * gb
    if i > 0:
        dmodr = dmodr * (0.5 * (1 - b)) ** (i - 1)
    dmods = dfa * (gb * (0.5 * (1 + a)))
    if i > 0:
        dmods = dmods * (0.5 * (1 - b)) ** (i - 1)
    tmp = dgb * (0.5 * (1 - b)) ** i
    if i > 0:
        tmp = tmp - 0.5 * i * gb * (0.5 * (1 - b)) ** (i - 1)
    dmods = dmods + fa * tmp
    return 2 ** (i + 0.5) * dmodr, 2 ** (i + 0.5) * dmods


def vandermonde_2d(N, r, s):
    a, b = rs_to_ab(r, s)
    cols = []
    for i in range(N + 1):
        for j in range(N - i + 1):
            cols.append(simplex2d_p(a, b, i, j))
    return np.stack(cols, axis=-1)


def grad_vandermonde_2d(N, r, s):
    a, b = rs_to_ab(r, s)
    vr, vs = [], []
    for i in range(N + 1):
        for j in range(N - i + 1):
            dr, ds = grad_simplex2d_p(a, b, i, j)
            vr.append(dr)
            vs.append(ds)
    return np.stack(vr, axis=-1), np.stack(vs, axis=-1)


# ---------------------------------------------------------------- tetrahedron

def rst_to_abc(r, s, t):
    r = np.asarray(r, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    den_a = -s - t
    a = np.where(np.abs(den_a) > 1e-14, 2 * (1 + r) / np.where(np.abs(den_a) > 1e-14, den_a, 1) - 1, -1.0)
    den_b = 1 - t
    b = np.where(np.abs(t - 1) > 1e-14, 2 * (1 + s) / np.where(np.abs(t - 1) > 1e-14, den_b, 1) - 1, -1.0)
    return a, b, t


def simplex3d_p(a, b, c, i, j, k):
    h1 = jacobi_p(a, 0, 0, i)
    h2 = jacobi_p(b, 2 * i + 1, 0, j)
    h3 = jacobi_p(c, 2 * (i + j) + 2, 0, k)
    return 2 * math.sqrt(2.0) * h1 * h2 * ((1 - b) ** i) * h3 * ((1 - c) ** (i + j))


def grad_simplex3d_p(a, b, c, i, j, k):
    fa = jacobi_p(a, 0, 0, i)
    dfa = grad_jacobi_p(a, 0, 0, i)
    gb = jacobi_p(b, 2 * i + 1, 0, j)
    dgb = grad_jacobi_p(b, 2 * i + 1, 0, j)
    hc = jacobi_p(c, 2 * (i + j) + 2, 0, k)
    dhc = grad_jacobi_p(c, 2 * (i + j) + 2, 0, k)

    vr = dfa * (gb * hc)
    if i > 0:
        vr = vr * (0.5 * (1 - b)) ** (i - 1)
    if i + j > 0:
        vr = vr * (0.5 * (1 - c)) ** (i + j - 1)

    vs = 0.5 * (1 + a) * vr
    tmp = dgb * (0.5 * (1 - b)) ** i
    if i > 0:
        tmp = tmp + (-0.5 * i) * (gb * (0.5 * (1 - b)) ** (i - 1))
    if i + j > 0:
        tmp = tmp * (0.5 * (1 - c)) ** (i + j - 1)
    tmp = fa * (tmp * hc)
    vs = vs + tmp

    vt = 0.5 * (1 + a) * vr + 0.5 * (1 + b) * tmp
    tmp = dhc * (0.5 * (1 - c)) ** (i + j)
    if i + j > 0:
        tmp = tmp - 0.5 * (i + j) * (hc * (0.5 * (1 - c)) ** (i + j - 1))
    tmp = fa * (gb * tmp)
    tmp = tmp * (0.5 * (1 - b)) ** i
    vt = vt + tmp

    sc = 2 ** (2 * i + j + 1.5)
    return vr * sc, vs * sc, vt * sc


def vandermonde_3d(N, r, s, t):
    a, b, c = rst_to_abc(r, s, t)
    cols = []
    for i in range(N + 1):
        for j in range(N - i + 1):
            for k in range(N - i - j + 1):
                cols.append(simplex3d_p(a, b, c, i, j, k))
    return np.stack(cols, axis=-1)


def grad_vandermonde_3d(N, r, s, t):
    a, b, c = rst_to_abc(r, s, t)
    vr, vs, vt = [], [], []
    for i in range(N + 1):
        for j in range(N - i + 1):
            for k in range(N - i - j + 1):
                dr, ds, dt = grad_simplex3d_p(a, b, c, i, j, k)
                vr.append(dr)
                vs.append(ds)
                vt.append(dt)
    return np.stack(vr, axis=-1), np.stack(vs, axis=-1), np.stack(vt, axis=-1)


def vandermonde_1d(N, r):
    return np.stack([jacobi_p(r, 0, 0, i) for i in range(N + 1)], axis=-1)


# ---------------------------------------------------------------- nodes

def equi_nodes_2d(N):
    """Equispaced nodes on the reference triangle, ordered so the s=-1 edge comes first."""
    pts = []
    for j in range(N + 1):
        for i in range(N + 1 - j):
            pts.append((-1 + 2.0 * i / N, -1 + 2.0 * j / N) if N > 0 else (-1 / 3, -1 / 3))
    p = np.array(pts)
    return p[:, 0], p[:, 1]


def equi_nodes_3d(N):
    """Equispaced nodes on the reference tet, ordered so the t=-1 face comes first."""
    pts = []
    for k in range(N + 1):
        for j in range(N + 1 - k):
            for i in range(N + 1 - k - j):
                if N > 0:
                    pts.append((-1 + 2.0 * i / N, -1 + 2.0 * j / N, -1 + 2.0 * k / N))
                else:
                    pts.append((-0.5, -0.5, -0.5))
    p = np.array(pts)
    return p[:, 0], p[:, 1], p[:, 2]
