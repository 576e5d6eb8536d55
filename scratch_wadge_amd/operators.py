"""Reference-element operators of the hybridized entropy-stable DG scheme.

Host-side setup (runs once per problem, never on the hot path).  The formulas are
the reference's:

* reference operators (operators.cpp: build_ref_ops_2D :104-212,
  build_ref_ops_3D :214-341): Vq, Pq = M^{-1} Vq^T W, face quadrature and the
  scaled reference normals nrJ/nsJ/ntJ;
* hybridized operators (operators_c.cpp:230-241 for 3D, :62-69 for 2D):
      Lq   = M^{-1} Vfq^T Wf
      VqLq = Vq Lq,  VqPq = Vq Pq,  VfPq = Vfq Pq
      Drq  = Vq Dr Pq - 1/2 Vq Lq diag(nrJ) Vfq Pq   (likewise Dsq, Dtq).

The polynomial space is spanned with the orthonormal Dubiner basis directly
(the reference goes through a nodal basis first); every operator above maps
quadrature values to quadrature values, so the choice of basis only changes
round-off.  The quadrature rules are the reference's tabulated ones (the
tables of its node_data.h that tet_cubature / tri_cubature load,
operators.cpp:1870-2100: the same points, weights, order and digits), stored
in quadrature_rules.json, so that volume points, face points and operators are
the reference's.  The same file keeps this package's own generated rules of the
same sizes under "<key>_gen" (quadrature_gen.py); use_node_data() reads the
rules of another node_data.h.
"""

from __future__ import annotations

import functools
import json
import os
from dataclasses import dataclass

import numpy as np

from . import refelem

_RULES_PATH = os.path.join(os.path.dirname(__file__), "quadrature_rules.json")

# Rules read from a reference node_data.h (use_node_data); they take precedence
# over the stored ones for the degrees they cover.
_OVERRIDE: dict = {}


@functools.lru_cache(maxsize=None)
def _rules():
    with open(_RULES_PATH) as f:
        return json.load(f)


def parse_node_data(path):
    """Quadrature rules tabulated in a reference node_data.h (the tables that
    tet_cubature / tri_cubature load, operators.cpp:1870-2100): for each degree M
    present, 'tet{M}' from p_rq_N{M}, p_sq_N{M}, p_tq_N{M}, p_wq_N{M} and 'tri{M}'
    from p_rfq_N{M}, p_sfq_N{M}, p_wfq_N{M}, on the biunit reference elements
    (the same as this package's); and for each degree N the tetrahedron's
    interpolation nodes 'nodes{N}' from p_r_N{N}, p_s_N{N}, p_t_N{N} (Nodes3D,
    operators.cpp:1791-1840).  Returns {key: {"points": [[..]], "weights": [..]}}
    (no weights for the node sets)."""
    import re
    with open(path) as f:
        src = f.read()
    arrays = {}
    for m in re.finditer(r"double\s+(p_\w+?)\s*\[\s*\d+\s*\]\s*=\s*\{([^}]*)\}", src):
        arrays[m.group(1)] = [float(x) for x in m.group(2).split(",") if x.strip()]
    rules = {}
    for name in arrays:
        m = re.fullmatch(r"p_r_N(\d+)", name)
        if m and all(f"p_{c}_N{m.group(1)}" in arrays for c in "st"):
            n = m.group(1)
            pts = np.stack([arrays[f"p_{c}_N{n}"] for c in "rst"], 1)
            rules[f"nodes{n}"] = {"points": pts.tolist()}
        m = re.fullmatch(r"p_wq_N(\d+)", name)
        if m and all(f"p_{c}q_N{m.group(1)}" in arrays for c in "rst"):
            M = m.group(1)
            pts = np.stack([arrays[f"p_{c}q_N{M}"] for c in "rst"], 1)
            rules[f"tet{M}"] = {"points": pts.tolist(), "weights": arrays[name]}
        m = re.fullmatch(r"p_wfq_N(\d+)", name)
        if m and all(f"p_{c}fq_N{m.group(1)}" in arrays for c in "rs"):
            M = m.group(1)
            pts = np.stack([arrays[f"p_{c}fq_N{M}"] for c in "rs"], 1)
            rules[f"tri{M}"] = {"points": pts.tolist(), "weights": arrays[name]}
    return rules


def use_node_data(path=None):
    """Build operators from the quadrature rules of the reference's own
    node_data.h at `path` (a user's copy of the reference; nothing of it ships
    with this package), so that the operators equal the reference's up to
    round-off.  path=None goes back to this package's stored rules.  The
    environment variable WADGES_NODE_DATA does the same at import.  Returns the
    rule keys now taken from the file."""
    _OVERRIDE.clear()
    if path:
        _OVERRIDE.update(parse_node_data(path))
    if "build_ref_ops" in globals():
        build_ref_ops.cache_clear()   # operators built from the previous rules
    return sorted(_OVERRIDE)


def _rule(key):
    return _OVERRIDE.get(key) or _rules()[key]


def tri_rule(degree):
    r = _rule(f"tri{degree}")
    x = np.array(r["points"])
    return x[:, 0], x[:, 1], np.array(r["weights"])


def ref_nodes(dim, N):
    """Interpolation nodes of degree N on the reference element, the
    reference's tabulated Nodes3D set (operators.cpp:1791-1840); in 2D its
    first (N+1)(N+2)/2 nodes, which lie on the face t = -1 (Nodes2D,
    operators.cpp:1780-1789).  Returns (r, s[, t])."""
    x = np.array(_rule(f"nodes{N}")["points"])
    if dim == 2:
        npt = (N + 1) * (N + 2) // 2
        if not np.all(x[:npt, 2] == -1.0):
            raise ValueError(f"the first {npt} nodes of degree {N} are not on the face t = -1")
        return x[:npt, 0], x[:npt, 1]
    return x[:, 0], x[:, 1], x[:, 2]


def tet_rule(degree):
    r = _rule(f"tet{degree}")
    x = np.array(r["points"])
    return x[:, 0], x[:, 1], x[:, 2], np.array(r["weights"])


if os.environ.get("WADGES_NODE_DATA"):
    use_node_data(os.environ["WADGES_NODE_DATA"])


def available_orders(dim):
    """Polynomial degrees N for which a degree-2N volume (and face) rule is stored."""
    keys = _rules().keys()
    out = []
    for N in range(1, 8):
        if dim == 2 and f"tri{2 * N}" in keys:
            out.append(N)
        if dim == 3 and f"tet{2 * N}" in keys and f"tri{2 * N}" in keys:
            out.append(N)
    return out


@dataclass
class RefOps:
    dim: int
    N: int
    Np: int
    Nq: int
    Nfq: int
    Nfaces: int
    # quadrature
    rq: np.ndarray          # (Nq, dim) reference coordinates of volume points
    wq: np.ndarray          # (Nq,)
    rf: np.ndarray          # (Nf, dim) reference coordinates of face points, face-major
    wf: np.ndarray          # (Nf,)
    nref: np.ndarray        # (Nf, dim) scaled reference normals nrJ, nsJ(, ntJ)
    # modal operators
    Vq: np.ndarray          # (Nq, Np)
    Vf: np.ndarray          # (Nf, Np)
    Pq: np.ndarray          # (Np, Nq)
    Dmodal: list            # dim x (Np, Np) modal derivative matrices
    # hybridized operators the kernels use (column-major when flattened with order='F')
    Drst: list              # dim x (Nq, Nq)
    VqLq: np.ndarray        # (Nq, Nf)
    VqPq: np.ndarray        # (Nq, Nq)
    VfPq: np.ndarray        # (Nf, Nq)

    @property
    def Nf(self):
        return self.Nfq * self.Nfaces


def _modal_derivs(dim, N):
    """Modal derivative matrices D_r with  d/dr (sum_k c_k phi_k) = sum_k (D_r c)_k phi_k.

    Evaluated through a unisolvent point set: D = V^{-1} V_r.
    """
    if dim == 2:
        r, s = refelem.equi_nodes_2d(N)
        V = refelem.vandermonde_2d(N, r, s)
        Vr, Vs = refelem.grad_vandermonde_2d(N, r, s)
        Vi = np.linalg.inv(V)
        return [Vi @ Vr, Vi @ Vs]
    r, s, t = refelem.equi_nodes_3d(N)
    V = refelem.vandermonde_3d(N, r, s, t)
    Vr, Vs, Vt = refelem.grad_vandermonde_3d(N, r, s, t)
    Vi = np.linalg.inv(V)
    return [Vi @ Vr, Vi @ Vs, Vi @ Vt]


def face_points_2d(N):
    """Gauss points on the three edges, edge-major (operators.cpp:135-165)."""
    r1, w1 = refelem.jacobi_gq(N)      # JacobiGQ(ceil(M/2)) with M = 2N: N+1 points
    one = np.ones_like(r1)
    rf = np.concatenate([r1, -r1, -one])
    sf = np.concatenate([-one, r1, -r1])
    nr = np.concatenate([0 * one, one, -one])
    ns = np.concatenate([-one, one, 0 * one])
    wf = np.concatenate([w1, w1, w1])
    return np.stack([rf, sf], 1), wf, np.stack([nr, ns], 1), len(r1)


def face_points_3d(N):
    """Triangle rule on the four faces, face-major (operators.cpp:244-291)."""
    a, b, w = tri_rule(2 * N)
    one = np.ones_like(a)
    r = np.concatenate([a, a, -(one + a + b), -one])
    s = np.concatenate([b, -one, a, a])
    t = np.concatenate([-one, b, b, b])
    z = 0 * one
    nr = np.concatenate([z, z, one, -one])
    ns = np.concatenate([z, -one, one, z])
    nt = np.concatenate([-one, z, one, z])
    wf = np.concatenate([w, w, w, w])
    return np.stack([r, s, t], 1), wf, np.stack([nr, ns, nt], 1), len(a)


@functools.lru_cache(maxsize=None)
def build_ref_ops(dim, N):
    if dim == 2:
        r, s, wq = tri_rule(2 * N)
        rq = np.stack([r, s], 1)
        Vq = refelem.vandermonde_2d(N, r, s)
        rf, wf, nref, nfq = face_points_2d(N)
        Vf = refelem.vandermonde_2d(N, rf[:, 0], rf[:, 1])
        nfaces = 3
    else:
        r, s, t, wq = tet_rule(2 * N)
        rq = np.stack([r, s, t], 1)
        Vq = refelem.vandermonde_3d(N, r, s, t)
        rf, wf, nref, nfq = face_points_3d(N)
        Vf = refelem.vandermonde_3d(N, rf[:, 0], rf[:, 1], rf[:, 2])
        nfaces = 4
    M = Vq.T @ (wq[:, None] * Vq)
    Pq = np.linalg.solve(M, Vq.T * wq[None, :])
    Lq = np.linalg.solve(M, Vf.T * wf[None, :])
    Dm = _modal_derivs(dim, N)
    VqLq = Vq @ Lq
    VqPq = Vq @ Pq
    VfPq = Vf @ Pq
    Drst = [Vq @ Dm[d] @ Pq - 0.5 * VqLq @ (nref[:, d][:, None] * VfPq) for d in range(dim)]
    return RefOps(dim=dim, N=N, Np=Vq.shape[1], Nq=len(wq), Nfq=nfq, Nfaces=nfaces,
                  rq=rq, wq=wq, rf=rf, wf=wf, nref=nref, Vq=Vq, Vf=Vf, Pq=Pq, Dmodal=Dm,
                  Drst=Drst, VqLq=VqLq, VqPq=VqPq, VfPq=VfPq)


def colmajor(A):
    """Flatten an operator the way the kernels index it: A[i + j*rows]."""
    return np.ascontiguousarray(np.asarray(A, dtype=np.float64).flatten(order="F"))
