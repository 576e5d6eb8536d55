"""Generator for the fully symmetric simplex quadrature rules this package uses.

The rules are found from scratch by solving the moment equations
    sum_q w_q phi_k(x_q) = integral of phi_k     for every orthonormal phi_k of
                                                  degree <= d
over symmetry-orbit parameterisations (barycentric orbits S3/S21/S111 on the
triangle; S4/S31/S22/S211/S1111 on the tetrahedron) with a Levenberg-Marquardt
least-squares solve and random restarts, keeping only rules with positive
weights and interior points.  The point counts are chosen to equal the ones the
reference reads from its tabulated rules for M = 2N (tet_cubature /
tri_cubature, operators.cpp:1870, :2023): triangle 3/6/12/16/25 points for
degree 2/4/6/8/10 and tetrahedron 4/11/23/44 points for degree 2/4/6/8, so a
degree-N element has the same Nq and Nfq as in the reference.

Run as ``python -m scratch_wadge_amd.quadrature_gen`` to regenerate the
``<key>_gen`` entries of ``quadrature_rules.json``; the tests re-check every
stored rule against the exact monomial integrals.  The operators are built
from the reference's tabulated rules (the ``<key>`` entries, from its
node_data.h); the generated ones are kept as an independent check: the
triangle rules of degree 2-8 are the reference's point sets, the others are
other rules of the same sizes.
"""

from __future__ import annotations

import itertools
import json
import math
import os
import sys

import numpy as np
from scipy.optimize import least_squares


TRI_V = np.array([[-1.0, -1.0], [1.0, -1.0], [-1.0, 1.0]])
TET_V = np.array([[-1.0, -1.0, -1.0], [1.0, -1.0, -1.0], [-1.0, 1.0, -1.0], [-1.0, -1.0, 1.0]])
TRI_AREA = 2.0
TET_VOL = 4.0 / 3.0

# orbit name -> (number of free barycentric parameters, number of points)
TRI_ORBITS = {"S3": (0, 1), "S21": (1, 3), "S111": (2, 6)}
TET_ORBITS = {"S4": (0, 1), "S31": (1, 4), "S22": (1, 6), "S211": (2, 12), "S1111": (3, 24)}


# Orbit generators: base barycentric tuple as a function of the free parameters,
# with role labels (equal labels = equal coordinates) that fix the orbit size.
def _orbit_base(dim, kind, p):
    if dim == 2:
        if kind == "S3":
            return [1 / 3] * 3, (0, 0, 0)
        if kind == "S21":
            return [p[0], p[0], 1 - 2 * p[0]], (0, 0, 1)
        return [p[0], p[1], 1 - p[0] - p[1]], (0, 1, 2)
    if kind == "S4":
        return [0.25] * 4, (0, 0, 0, 0)
    if kind == "S31":
        return [p[0]] * 3 + [1 - 3 * p[0]], (0, 0, 0, 1)
    if kind == "S22":
        return [p[0], p[0], 0.5 - p[0], 0.5 - p[0]], (0, 0, 1, 1)
    if kind == "S211":
        return [p[0], p[0], p[1], 1 - 2 * p[0] - p[1]], (0, 0, 1, 2)
    return [p[0], p[1], p[2], 1 - p[0] - p[1] - p[2]], (0, 1, 2, 3)


_PERM_CACHE = {}


def _orbit_perms(dim, kind):
    key = (dim, kind)
    if key not in _PERM_CACHE:
        _, roles = _orbit_base(dim, kind, [0.1, 0.2, 0.3])
        n = len(roles)
        out, keys = [], set()
        for perm in itertools.permutations(range(n)):
            k = tuple(roles[i] for i in perm)
            if k in keys:
                continue
            keys.add(k)
            out.append(perm)
        orbits = TRI_ORBITS if dim == 2 else TET_ORBITS
        assert len(out) == orbits[kind][1]
        _PERM_CACHE[key] = np.array(out)
    return _PERM_CACHE[key]


def _orbit_points(dim, kind, p):
    base, _ = _orbit_base(dim, kind, p)
    return np.asarray(base, dtype=np.float64)[_orbit_perms(dim, kind)]


def _monomials(nbar, deg):
    """Exponent tuples of all homogeneous barycentric monomials of total degree deg."""
    out = []
    for c in itertools.combinations_with_replacement(range(nbar), deg):
        e = [0] * nbar
        for i in c:
            e[i] += 1
        out.append(tuple(e))
    return np.array(sorted(set(out)))


def _exact_moments(nbar, exps):
    """Integral of prod(lambda_i^a_i) over the simplex, divided by its volume:
    (nbar-1)! prod(a_i!) / (nbar-1+sum a_i)!  (Dirichlet integral)."""
    d = nbar - 1
    return np.array([math.factorial(d) * math.prod(math.factorial(a) for a in e)
                     / math.factorial(d + sum(e)) for e in exps])


def _assemble(dim, structure, params):
    """Return (barycentric points, per-point orbit index, number of position parameters)."""
    orbits = TRI_ORBITS if dim == 2 else TET_ORBITS
    pts, owner = [], []
    ofs = 0
    for oi, kind in enumerate(structure):
        npar = orbits[kind][0]
        lam = _orbit_points(dim, kind, params[ofs:ofs + npar])
        ofs += npar
        pts.append(lam)
        owner += [oi] * len(lam)
    return np.vstack(pts), np.array(owner), ofs


_MOM_CACHE = {}


def _residual(z, dim, deg, structure):
    key = (dim, deg)
    if key not in _MOM_CACHE:
        exps = _monomials(dim + 1, deg)
        _MOM_CACHE[key] = (exps, _exact_moments(dim + 1, exps))
    exps, mom = _MOM_CACHE[key]
    lam, owner, ofs = _assemble(dim, structure, z)
    w = z[ofs:][owner]
    vol = TRI_AREA if dim == 2 else TET_VOL
    P = np.prod(lam[:, None, :] ** exps[None, :, :], axis=2)   # (npts, nmono)
    return (w @ P) / vol - mom


def _bounds(dim, structure):
    orbits = TRI_ORBITS if dim == 2 else TET_ORBITS
    lo, hi = [], []
    for kind in structure:
        npar = orbits[kind][0]
        if kind in ("S21",):
            lo += [0.0]; hi += [0.5]
        elif kind == "S31":
            lo += [0.0]; hi += [1.0 / 3.0]
        elif kind == "S22":
            lo += [0.0]; hi += [0.25]
        else:
            lo += [0.0] * npar; hi += [1.0 / (dim + 1) + 0.5] * npar
    return lo, hi


def solve_rule(dim, deg, structure, tries=400, seed=1):
    orbits = TRI_ORBITS if dim == 2 else TET_ORBITS
    nbar = dim + 1
    npar = sum(orbits[k][0] for k in structure)
    npts = sum(orbits[k][1] for k in structure)
    vol = TRI_AREA if dim == 2 else TET_VOL
    rng = np.random.default_rng(seed)
    plo, phi = _bounds(dim, structure)
    lo = np.array(plo + [0.0] * len(structure))
    hi = np.array(phi + [vol] * len(structure))
    for _ in range(tries):
        z0 = np.concatenate([rng.uniform(0.0, 1.0, npar) * (np.array(phi) - np.array(plo)) + np.array(plo),
                             rng.uniform(0.2, 1.0, len(structure)) * vol / npts])
        z0 = np.clip(z0, lo + 1e-6, hi - 1e-6)
        try:
            sol = least_squares(_residual, z0, args=(dim, deg, structure), method="trf",
                                bounds=(lo, hi), xtol=1e-15, ftol=1e-15, gtol=1e-15,
                                max_nfev=300)
            if np.max(np.abs(sol.fun)) < 1e-6:
                sol = least_squares(_residual, sol.x, args=(dim, deg, structure), method="lm",
                                    xtol=1e-15, ftol=1e-15, gtol=1e-15, max_nfev=2000)
        except Exception:
            continue
        if np.max(np.abs(sol.fun)) > 1e-13:
            continue
        lam, owner, ofs = _assemble(dim, structure, sol.x)
        w = sol.x[ofs:][owner]
        if np.min(lam) < 1e-9 or np.min(w) <= 0:
            continue
        V = TRI_V if dim == 2 else TET_V
        return lam @ V, w
    return None


def _free_points(z, npts):
    """Collapsed (Duffy) coordinates in [0,1]^3 -> reference tetrahedron points."""
    a, b, c = z[:npts], z[npts:2 * npts], z[2 * npts:3 * npts]
    zz = c
    yy = b * (1 - c)
    xx = a * (1 - b) * (1 - c)
    return np.stack([-1 + 2 * xx, -1 + 2 * yy, -1 + 2 * zz], 1)


def _free_residual(z, deg, npts):
    from . import refelem
    x = _free_points(z, npts)
    w = z[3 * npts:]
    B = refelem.vandermonde_3d(deg, x[:, 0], x[:, 1], x[:, 2])
    target = np.zeros(B.shape[1])
    target[0] = math.sqrt(TET_VOL)
    return B.T @ w - target


def _free_jac(z, deg, npts):
    from . import refelem
    a, b, c = z[:npts], z[npts:2 * npts], z[2 * npts:3 * npts]
    w = z[3 * npts:]
    x = _free_points(z, npts)
    B = refelem.vandermonde_3d(deg, x[:, 0], x[:, 1], x[:, 2])          # (npts, nb)
    Br, Bs, Bt = refelem.grad_vandermonde_3d(deg, x[:, 0], x[:, 1], x[:, 2])
    # r = -1 + 2 a (1-b)(1-c), s = -1 + 2 b (1-c), t = -1 + 2 c
    dr_da, dr_db, dr_dc = 2 * (1 - b) * (1 - c), -2 * a * (1 - c), -2 * a * (1 - b)
    ds_db, ds_dc = 2 * (1 - c), -2 * b
    dt_dc = 2.0 + 0 * c
    Ja = (Br * dr_da[:, None]) * w[:, None]
    Jb = (Br * dr_db[:, None] + Bs * ds_db[:, None]) * w[:, None]
    Jc = (Br * dr_dc[:, None] + Bs * ds_dc[:, None] + Bt * dt_dc[:, None]) * w[:, None]
    return np.concatenate([Ja, Jb, Jc, B], 0).T                           # (nb, 4 npts)


def solve_free_tet_rule(deg, npts, tries=200, seed=5):
    """Tetrahedron rule without imposed symmetry (positive weights, interior points).

    With 4*npts unknowns against (deg+1)(deg+2)(deg+3)/6 moment equations the
    system is underdetermined for the point counts used here, and bounded
    least squares from random interior starts converges readily."""
    rng = np.random.default_rng(seed)
    lo = np.concatenate([np.full(3 * npts, 0.005), np.zeros(npts)])
    hi = np.concatenate([np.full(3 * npts, 0.995), np.full(npts, TET_VOL)])
    for _ in range(tries):
        z0 = np.concatenate([rng.uniform(0.05, 0.95, 3 * npts), np.full(npts, TET_VOL / npts)])
        sol = least_squares(_free_residual, z0, jac=_free_jac, args=(deg, npts), method="trf",
                            bounds=(lo, hi), xtol=1e-15, ftol=1e-15, gtol=1e-15, max_nfev=400)
        if np.max(np.abs(sol.fun)) > 1e-13:
            continue
        x = _free_points(sol.x, npts)
        w = sol.x[3 * npts:]
        lam = np.concatenate([(1 + x) / 2, -(1 + x.sum(1, keepdims=True)) / 2], 1)
        if np.min(lam) < 1e-3 or np.min(w) < 1e-4 * TET_VOL:
            continue
        return x, w
    return None


# (dim, degree) -> orbit structures to try, in order; the point count equals the
# reference's rule size for that degree.
TARGETS = {
    (2, 2): [("S21",)],
    (2, 4): [("S21", "S21")],
    (2, 6): [("S21", "S21", "S111")],
    (2, 8): [("S3", "S21", "S21", "S21", "S111")],
    (2, 10): [("S3", "S21", "S21", "S111", "S111", "S111")],
    (3, 2): [("S31",)],
    (3, 4): [("S4", "S31", "S22"), ("S31", "S31", "S4", "S22")],
    (3, 6): [("S4", "S31", "S31", "S31", "S31", "S22"), ("S4", "S31", "S22", "S211"),
             ("S4", "S31", "S22", "S22", "S22")],
    (3, 8): [("S31", "S31", "S31", "S31", "S31", "S211", "S211"),
             ("S31", "S31", "S22", "S22", "S22", "S22", "S211"),
             ("S31", "S31", "S31", "S22", "S22", "S1111")],
}

EXPECTED_COUNTS = {(2, 2): 3, (2, 4): 6, (2, 6): 12, (2, 8): 16, (2, 10): 25,
                   (3, 2): 4, (3, 4): 11, (3, 6): 23, (3, 8): 44}


def _store(rules, dim, deg, x, w, st):
    order = np.lexsort(tuple(x[:, d] for d in reversed(range(dim))))
    rules[f"{'tri' if dim == 2 else 'tet'}{deg}_gen"] = {
        "degree": deg, "points": x[order].tolist(), "weights": w[order].tolist(),
        "structure": list(st), "source": "quadrature_gen"}


def main(keys=None, out_path=None, tries=20000):
    """Solve the rules named in keys (default: all TARGETS) and merge them into the JSON file."""
    out_path = out_path or os.path.join(os.path.dirname(__file__), "quadrature_rules.json")
    rules = {}
    if os.path.exists(out_path):
        with open(out_path) as f:
            rules = json.load(f)
    for (dim, deg) in (keys or list(TARGETS)):
        found = None
        if dim == 3 and deg >= 4:
            # the tetrahedron rules with these point counts are not fully symmetric
            st = ("free",)
            found = solve_free_tet_rule(deg, EXPECTED_COUNTS[(dim, deg)], tries=tries)
        else:
            for st in TARGETS[(dim, deg)]:
                found = solve_rule(dim, deg, st, tries=tries)
                if found is not None:
                    break
        if found is None:
            print(f"no rule found for dim={dim} degree={deg}", file=sys.stderr)
            continue
        x, w = found
        assert len(w) == EXPECTED_COUNTS[(dim, deg)]
        _store(rules, dim, deg, x, w, st)
        print(f"dim={dim} degree={deg}: {len(w)} points ({'+'.join(st)})")
    with open(out_path, "w") as f:
        json.dump(rules, f, indent=1)


if __name__ == "__main__":
    ks = [tuple(int(v) for v in a.split(",")) for a in sys.argv[1:]] or None
    main(ks)
