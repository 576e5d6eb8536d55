"""Structured periodic simplex meshes, curved geometry and face connectivity.

Host-side setup that produces the kernels' constant inputs in the reference's
layouts (operators_c.cpp:135-177 for 2D, :322-389 for 3D; euler.c:3158-3170 for
the field-strided neighbour map):

    vgeo  [K][Nvgeo][Nq]   geometric factors (rxJ, ryJ, (rzJ), sxJ, ...) at volume points
    vfgeo [K][Nvgeo][Nf]   the same at face points
    fgeo  [K][Nfgeo][Nf]   scaled normals nxJ, nyJ, (nzJ) and surface Jacobian sJ
    Jq    [K][Nq]          VqPq * J (projected Jacobian, operators.cpp:758-759)
    mapPq [K][Nf]          e_nbr*Nf*Nfields + (face node of the neighbour)

The reference reads gmsh files and matches periodic faces by coordinates
(euler.c:556-833, :3004-3175).  Here the meshes are generated: a periodic box of
n^d cells, squares split into 2 triangles and cubes into 6 Kuhn tetrahedra,
with vertex ids taken modulo n so periodic faces share vertex ids.  Face points
are matched through their barycentric weights on the shared vertex ids, which
is exact for the fully symmetric face rules used here.

Geometric factors follow build_geofacs_2D/3D (operators.cpp:343-456, :458-763):
the mapping is the degree-N interpolant of the (optionally warped) nodal
positions; 3D uses Kopriva's conservative curl form at degree N+1
(operators.cpp:617-719), interpolated back to degree N.
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

from . import refelem
from .operators import build_ref_ops, RefOps

# Kuhn split of the unit cube into 6 tetrahedra along the (0,0,0)-(1,1,1) diagonal.
_KUHN = [list(p) for p in itertools.permutations(range(3))]


@dataclass
class SimplexMesh:
    dim: int
    K: int
    EToV: np.ndarray        # (K, dim+1) periodic vertex ids
    VX: np.ndarray          # (K, dim+1, dim) unwrapped vertex coordinates per element
    period: np.ndarray      # (dim,) box lengths
    lo: np.ndarray          # (dim,) box lower corner
    # True: periodic faces do not share vertex ids (meshes read from files); the
    # face points left unmatched by the vertex ids are matched by coordinates
    # modulo the box (modify_mapP, euler.c:3004-3175).
    periodic_points: bool = False


def periodic_box_tri(n, lo=-5.0, hi=5.0, perturb=0.0, seed=0):
    """[lo,hi]^2 split into n x n squares, each cut into two right triangles.

    Interior vertices are moved by U(-perturb*h, perturb*h) per coordinate
    (boundary vertices stay, so periodic images still coincide)."""
    assert n >= 3
    h = (hi - lo) / n
    g = np.linspace(lo, hi, n + 1)
    X, Y = np.meshgrid(g, g, indexing="ij")
    pos = np.stack([X, Y], -1)  # (n+1, n+1, 2)
    if perturb > 0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-perturb * h, perturb * h, size=pos.shape)
        d[0, :, :] = d[-1, :, :] = 0
        d[:, 0, :] = d[:, -1, :] = 0
        pos = pos + d
    ev, vx = [], []
    for i in range(n):
        for j in range(n):
            c = [(i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)]
            for tri in ([0, 1, 2], [0, 2, 3]):
                ids = [c[t] for t in tri]
                ev.append([(a % n) * n + (b % n) for a, b in ids])
                vx.append([pos[a, b] for a, b in ids])
    ev = np.array(ev, dtype=np.int64)
    vx = np.array(vx)
    return _orient(SimplexMesh(2, len(ev), ev, vx, np.array([hi - lo] * 2), np.array([lo] * 2)))


def periodic_box_tet(n, lo=-np.pi, hi=np.pi):
    """[lo,hi]^3 split into n^3 cubes, each cut into 6 Kuhn tetrahedra."""
    assert n >= 3
    h = (hi - lo) / n
    # Kuhn paths: corner offsets of the 6 tets of the unit cube, (6, 4, 3)
    paths = []
    for perm in _KUHN:
        path = [np.zeros(3, dtype=np.int64)]
        for ax in perm:
            nxt = path[-1].copy()
            nxt[ax] += 1
            path.append(nxt)
        paths.append(path)
    paths = np.array(paths)
    I, J, Kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    cells = np.stack([I.ravel(), J.ravel(), Kk.ravel()], 1)            # (n^3, 3)
    corners = cells[:, None, None, :] + paths[None, :, :, :]           # (n^3, 6, 4, 3)
    corners = corners.reshape(-1, 4, 3)
    w = corners % n
    ev = (w[..., 0] * n + w[..., 1]) * n + w[..., 2]
    vx = lo + h * corners.astype(np.float64)
    return _orient(SimplexMesh(3, len(ev), ev.astype(np.int64), vx, np.array([hi - lo] * 3),
                               np.array([lo] * 3)))


def _orient(m):
    """Swap two vertices of negatively oriented elements."""
    if m.dim == 2:
        a = m.VX[:, 1] - m.VX[:, 0]
        b = m.VX[:, 2] - m.VX[:, 0]
        det = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    else:
        a = m.VX[:, 1] - m.VX[:, 0]
        b = m.VX[:, 2] - m.VX[:, 0]
        c = m.VX[:, 3] - m.VX[:, 0]
        det = np.einsum("ki,ki->k", a, np.cross(b, c))
    neg = det < 0
    m.EToV[neg] = m.EToV[neg][:, [0, 2, 1] + list(range(3, m.dim + 1))]
    m.VX[neg] = m.VX[neg][:, [0, 2, 1] + list(range(3, m.dim + 1))]
    return m


# ------------------------------------------------------------------ geometry

def tg_warp(x, y, z, a=0.1):
    """Cyclic sinusoidal warp (operators.cpp:517-529 with amplitude a)."""
    x = x + a * np.sin(x) * np.sin(y) * np.sin(z)
    y = y + a * np.sin(x) * np.sin(y) * np.sin(z)
    z = z + a * np.sin(x) * np.sin(y) * np.sin(z)
    return x, y, z


def _nodes(dim, N):
    """The degree-N nodes the geometry is interpolated on: the reference's
    (Nodes2D / Nodes3D, operators.cpp:1780-1840), as the reference maps each
    element (operators.cpp:357-370, :485-499) and builds the degree-(N+1)
    Kopriva geofacs (operators.cpp:619-636)."""
    from .operators import ref_nodes
    return ref_nodes(dim, N)


def _vdm(dim, N, pts):
    return refelem.vandermonde_2d(N, *pts) if dim == 2 else refelem.vandermonde_3d(N, *pts)


def _gvdm(dim, N, pts):
    return refelem.grad_vandermonde_2d(N, *pts) if dim == 2 else refelem.grad_vandermonde_3d(N, *pts)


def nodal_positions(mesh, N, warp=None):
    """Physical coordinates of the degree-N nodes, (dim, Np, K)."""
    r = _nodes(mesh.dim, N)
    if mesh.dim == 2:
        lam = [-(r[0] + r[1]) / 2, (1 + r[0]) / 2, (1 + r[1]) / 2]
    else:
        lam = [-(1 + r[0] + r[1] + r[2]) / 2, (1 + r[0]) / 2, (1 + r[1]) / 2, (1 + r[2]) / 2]
    lam = np.stack(lam, 1)  # (Np, nv)
    X = np.einsum("nv,kvd->dnk", lam, mesh.VX)
    if warp is not None:
        X = np.stack(warp(*X), 0)
    return X


@dataclass
class Geometry:
    vgeo: np.ndarray
    vfgeo: np.ndarray
    fgeo: np.ndarray
    Jq: np.ndarray
    xq: np.ndarray      # (dim, Nq, K) physical volume quadrature points
    xf: np.ndarray      # (dim, Nf, K)
    J_raw: np.ndarray   # (Nq, K) unprojected Jacobian at quadrature points


def build_geometry(mesh, ops: RefOps, warp=None) -> Geometry:
    dim, N = mesh.dim, ops.N
    rn = _nodes(dim, N)
    V = _vdm(dim, N, rn)
    Vi = np.linalg.inv(V)
    X = nodal_positions(mesh, N, warp)                      # (dim, Np, K)
    Vqn = ops.Vq @ Vi                                        # nodal -> quadrature
    Vfn = ops.Vf @ Vi                                        # nodal -> face points
    G = _gvdm(dim, N, rn)
    Dn = [Gd @ Vi for Gd in G]                               # nodal derivative matrices
    K = mesh.K
    if dim == 2:
        x, y = X
        xr, xs = Vqn @ (Dn[0] @ x), Vqn @ (Dn[1] @ x)
        yr, ys = Vqn @ (Dn[0] @ y), Vqn @ (Dn[1] @ y)
        geo_q = [ys, -xs, -yr, xr]                           # rxJ, ryJ, sxJ, syJ
        J = xr * ys - yr * xs
        xrf, xsf = Vfn @ (Dn[0] @ x), Vfn @ (Dn[1] @ x)
        yrf, ysf = Vfn @ (Dn[0] @ y), Vfn @ (Dn[1] @ y)
        geo_f = [ysf, -xsf, -yrf, xrf]
        nr, ns = ops.nref[:, 0][:, None], ops.nref[:, 1][:, None]
        nxJ = geo_f[0] * nr + geo_f[2] * ns
        nyJ = geo_f[1] * nr + geo_f[3] * ns
        sJ = np.sqrt(nxJ ** 2 + nyJ ** 2)
        fgeo = [nxJ, nyJ, sJ]
    else:
        x, y, z = X
        dq = [[Vqn @ (Dn[r] @ c) for r in range(3)] for c in (x, y, z)]   # dq[c][r]
        (xr, xs, xt), (yr, ys, yt), (zr, zs, zt) = dq
        J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt)
        # original (cross-product) face factors give sJ (operators.cpp:562-615)
        df = [[Vfn @ (Dn[r] @ c) for r in range(3)] for c in (x, y, z)]
        (xrf, xsf, xtf), (yrf, ysf, ytf), (zrf, zsf, ztf) = df
        rxJf = ysf * ztf - zsf * ytf
        sxJf = ytf * zrf - ztf * yrf
        txJf = yrf * zsf - zrf * ysf
        ryJf = xtf * zsf - xsf * ztf
        syJf = xrf * ztf - xtf * zrf
        tyJf = xsf * zrf - zsf * xrf
        rzJf = xsf * ytf - xtf * ysf
        szJf = xtf * yrf - xrf * ytf
        tzJf = xrf * ysf - xsf * yrf
        nr, ns, nt = (ops.nref[:, d][:, None] for d in range(3))
        nx0 = rxJf * nr + sxJf * ns + txJf * nt
        ny0 = ryJf * nr + syJf * ns + tyJf * nt
        nz0 = rzJf * nr + szJf * ns + tzJf * nt
        sJ = np.sqrt(nx0 ** 2 + ny0 ** 2 + nz0 ** 2)
        # Kopriva conservative curl form at degree N+1 (operators.cpp:617-719)
        N2 = N + 1
        r2 = _nodes(3, N2)
        V2 = _vdm(3, N2, r2)
        V2i = np.linalg.inv(V2)
        D2 = [Gd @ V2i for Gd in _gvdm(3, N2, r2)]
        V12 = _vdm(3, N, r2) @ Vi                          # degree-N nodal -> degree-(N+1) nodes
        V21 = _vdm(3, N2, rn) @ V2i                         # degree-(N+1) nodal -> degree-N nodes
        x2, y2, z2 = V12 @ x, V12 @ y, V12 @ z
        Dr, Ds, Dt = D2
        xr2, yr2 = Dr @ x2, Dr @ y2
        xs2, ys2 = Ds @ x2, Ds @ y2
        xt2, yt2 = Dt @ x2, Dt @ y2
        rxJ = Dt @ (ys2 * z2) - Ds @ (yt2 * z2)
        sxJ = Dr @ (yt2 * z2) - Dt @ (yr2 * z2)
        txJ = Ds @ (yr2 * z2) - Dr @ (ys2 * z2)
        ryJ = -(Dt @ (xs2 * z2) - Ds @ (xt2 * z2))
        syJ = -(Dr @ (xt2 * z2) - Dt @ (xr2 * z2))
        tyJ = -(Ds @ (xr2 * z2) - Dr @ (xs2 * z2))
        rzJ = -(Dt @ (ys2 * x2) - Ds @ (yt2 * x2))
        szJ = -(Dr @ (yt2 * x2) - Dt @ (yr2 * x2))
        tzJ = -(Ds @ (yr2 * x2) - Dr @ (ys2 * x2))
        nodal = [V21 @ a for a in (rxJ, ryJ, rzJ, sxJ, syJ, szJ, txJ, tyJ, tzJ)]
        geo_q = [Vqn @ a for a in nodal]
        geo_f = [Vfn @ a for a in nodal]
        nxJ = geo_f[0] * nr + geo_f[3] * ns + geo_f[6] * nt
        nyJ = geo_f[1] * nr + geo_f[4] * ns + geo_f[7] * nt
        nzJ = geo_f[2] * nr + geo_f[5] * ns + geo_f[8] * nt
        fgeo = [nxJ, nyJ, nzJ, sJ]
    Jq = ops.VqPq @ J
    vgeo = np.ascontiguousarray(np.stack(geo_q, 0).transpose(2, 0, 1))    # (K, Nvgeo, Nq)
    vfgeo = np.ascontiguousarray(np.stack(geo_f, 0).transpose(2, 0, 1))
    fg = np.ascontiguousarray(np.stack(fgeo, 0).transpose(2, 0, 1))
    xq = np.stack([Vqn @ c for c in X], 0)
    xf = np.stack([Vfn @ c for c in X], 0)
    return Geometry(vgeo=vgeo, vfgeo=vfgeo, fgeo=fg, Jq=np.ascontiguousarray(Jq.T),
                    xq=xq, xf=xf, J_raw=J)


# ------------------------------------------------------------------ connectivity

# Reference-element faces as local vertex triples, in face order
# (3D: t=-1, s=-1, r+s+t=-1, r=-1; 2D: s=-1, r+s=0, r=-1).
FACE_VERTS = {2: [[0, 1], [1, 2], [0, 2]], 3: [[0, 1, 2], [0, 1, 3], [1, 2, 3], [0, 2, 3]]}


def build_mapPq(mesh, ops: RefOps, nfields=None, xf=None):
    """Field-strided neighbour map mapPq[k][Nf] (euler.c:3158-3170).

    Face points are matched through the shared vertex ids and their barycentric
    weights; for meshes with mesh.periodic_points the points left over are then
    matched by their coordinates xf (dim, Nf, K) modulo the box.  Unmatched face
    points map to themselves, which the surface kernel treats as a zero jump."""
    dim = mesh.dim
    nfields = nfields or dim + 2
    Nfq, Nfaces = ops.Nfq, ops.Nfaces
    Nf = Nfq * Nfaces
    K = mesh.K
    rf = ops.rf
    if dim == 2:
        lam = np.stack([-(rf[:, 0] + rf[:, 1]) / 2, (1 + rf[:, 0]) / 2, (1 + rf[:, 1]) / 2], 1)
    else:
        lam = np.stack([-(1 + rf.sum(1)) / 2, (1 + rf[:, 0]) / 2, (1 + rf[:, 1]) / 2,
                        (1 + rf[:, 2]) / 2], 1)
    fv = np.array(FACE_VERTS[dim])                          # (Nfaces, dim)
    # per face point: global ids of the face vertices and the barycentric weights on them
    face_of = np.repeat(np.arange(Nfaces), Nfq)             # (Nf,)
    loc = fv[face_of]                                       # (Nf, dim) local vertex index
    w = np.take_along_axis(lam, loc, 1)                     # (Nf, dim)
    gid = mesh.EToV[:, loc]                                 # (K, Nf, dim)
    order = np.argsort(gid, axis=2, kind="stable")
    gid_s = np.take_along_axis(gid, order, 2)
    w_s = np.take_along_axis(np.broadcast_to(w, gid.shape), order, 2)
    wq = np.round(w_s[:, :, :dim - 1] * 2 ** 20).astype(np.int64)  # last weight is implied
    keys = np.concatenate([gid_s, wq], 2).reshape(K * Nf, -1)
    elem = np.repeat(np.arange(K), Nf)
    node = np.tile(np.arange(Nf), K)
    srt = np.lexsort(keys.T[::-1])
    ks = keys[srt]
    same = np.all(ks[1:] == ks[:-1], axis=1)
    if np.any(same[1:] & same[:-1]):
        raise ValueError("face point shared by more than two elements")
    mapP = node + elem * Nf                                 # self by default
    a = srt[:-1][same]
    b = srt[1:][same]
    mapP = mapP.copy()
    mapP[a] = node[b] + elem[b] * Nf
    mapP[b] = node[a] + elem[a] * Nf
    nmatched = int(np.sum(same))
    if mesh.periodic_points:
        if xf is None:
            raise ValueError("periodic matching by coordinates needs the face points xf")
        X = np.transpose(xf, (0, 2, 1)).reshape(dim, -1)          # index k*Nf + node
        nmatched += _match_periodic_points(mapP, X, mesh.lo, mesh.period)
    en = mapP // Nf
    mapPq = (mapP - en * Nf) + en * Nf * nfields
    return np.ascontiguousarray(mapPq.reshape(K, Nf).astype(np.uint32)), nmatched


# ------------------------------------------------------------------ gmsh ingest and connectivity

# Face-to-vertex and orientation tables of the reference's connectivity
# (euler.c:429-510): face f of an element has vertices FToFV[f]; a face seen
# from the neighbour is the same vertex list permuted by OToFV[o].
FTOFV = {2: np.array([[0, 1], [1, 2], [2, 0]]),
         3: np.array([[0, 1, 2], [0, 1, 3], [1, 2, 3], [0, 2, 3]])}
OTOFV = {2: np.array([[0, 1], [1, 0]]),
         3: np.array([[0, 1, 2], [2, 0, 1], [1, 2, 0], [1, 0, 2], [2, 1, 0], [0, 2, 1]])}
GMSH_TYPE = {2: 2, 3: 4}     # gmsh element type of a triangle / tetrahedron


def read_gmsh(path, dim):
    """gmsh 2.2 ASCII reader in the manner of host_mesh_read_msh (euler.c:556-701):
    $Nodes (id x y z), $Elements of the triangle (2D) or tetrahedron (3D) type
    with their tags skipped; other element types and the $Periodic section are
    ignored.  Returns (VX (Nv, dim) vertex coordinates, EToV (K, dim+1)
    zero-based vertex ids in file order)."""
    want = GMSH_TYPE[dim]
    nodes, elems = {}, []
    section = None
    with open(path) as f:
        lines = iter(f.read().splitlines())
        for line in lines:
            s = line.strip()
            if s.startswith("$"):
                section = s[1:] if not s.startswith("$End") else None
                if section in ("Nodes", "Elements"):
                    next(lines)                      # count line
                continue
            if not s:
                continue
            if section == "Nodes":
                w = s.split()
                nodes[int(w[0])] = [float(x) for x in w[1:1 + dim]]
            elif section == "Elements":
                w = [int(x) for x in s.split()]
                if w[1] == want:
                    ntags = w[2]
                    elems.append(w[3 + ntags:3 + ntags + dim + 1])
    if not nodes or not elems:
        raise ValueError(f"{path}: no $Nodes or no dimension-{dim} simplices")
    ids = np.array(sorted(nodes))
    VX = np.array([nodes[i] for i in ids], dtype=np.float64)
    remap = np.full(ids.max() + 1, -1, dtype=np.int64)
    remap[ids] = np.arange(len(ids))
    EToV = remap[np.array(elems, dtype=np.int64)]
    if (EToV < 0).any():
        raise ValueError(f"{path}: element refers to a missing node")
    return VX, EToV


def connect(EToV, dim):
    """Face connectivity with the reference's conventions (host_mesh_connect,
    euler.c:939-1612, single rank): faces are matched through their sorted vertex
    ids; EToE[e][f], EToF[e][f] name the neighbour element and face, EToO[e][f] the
    orientation o with EToV[e][FToFV[f][OToFV[o][v]]] == EToV[e1][FToFV[f1][v]]
    (the relation euler.c:1663-1705 checks).  Unmatched faces point to
    themselves with orientation 0 (euler.c:1565-1573)."""
    K = EToV.shape[0]
    nfc = dim + 1
    fv = EToV[:, FTOFV[dim]]                                 # (K, nfc, nfv)
    keys = np.sort(fv, axis=2).reshape(K * nfc, -1)
    srt = np.lexsort(keys.T[::-1])
    ks = keys[srt]
    same = np.all(ks[1:] == ks[:-1], axis=1)
    if np.any(same[1:] & same[:-1]):
        raise ValueError("face shared by more than two elements")
    EToE = np.repeat(np.arange(K)[:, None], nfc, 1)
    EToF = np.repeat(np.arange(nfc)[None, :], K, 0)
    EToO = np.zeros((K, nfc), dtype=np.int64)
    a, b = srt[:-1][same], srt[1:][same]
    ea, fa, eb, fb = a // nfc, a % nfc, b // nfc, b % nfc
    EToE[ea, fa], EToF[ea, fa] = eb, fb
    EToE[eb, fb], EToF[eb, fb] = ea, fa
    perms = OTOFV[dim]
    for (e0, f0, e1, f1) in ((ea, fa, eb, fb), (eb, fb, ea, fa)):
        v0, v1 = fv[e0, f0], fv[e1, f1]                      # (m, nfv)
        hit = np.all(v0[:, perms] == v1[:, None, :], axis=2)  # (m, norien)
        if not hit.any(1).all():
            raise ValueError("matched faces with no orientation relation")
        EToO[e0, f0] = hit.argmax(1)
    return EToE, EToF, EToO


def _match_periodic_points(mapP, X, lo, period, tol=1e-7):
    """Pairs the face points that map to themselves by their coordinates modulo
    the box (modify_mapP, euler.c:3004-3175, which compares all pairs; here the
    wrapped coordinates are hashed).  Updates mapP in place; returns the number
    of pairs."""
    free = np.nonzero(mapP == np.arange(mapP.size))[0]
    if free.size == 0:
        return 0
    Y = (X[:, free].T - lo) / period                       # (m, dim) in box units
    Y = Y - np.floor(Y)
    Y[Y > 1.0 - tol] -= 1.0                                # the upper face wraps to the lower
    q = np.round(Y / (10 * tol)).astype(np.int64)
    srt = np.lexsort(q.T[::-1])
    qs = q[srt]
    same = np.all(qs[1:] == qs[:-1], axis=1)
    if np.any(same[1:] & same[:-1]):
        raise ValueError("more than two face points coincide modulo the box")
    a, b = free[srt[:-1][same]], free[srt[1:][same]]
    d = X[:, a] - X[:, b]
    d = d - np.round(d / period[:, None]) * period[:, None]
    if np.abs(d).max(initial=0.0) > 1e3 * tol * period.max():
        raise ValueError("periodic face points do not coincide")
    mapP[a], mapP[b] = b, a
    return int(a.size)


def mesh_from_gmsh(path, dim, periodic=True):
    """SimplexMesh from a gmsh 2.2 file.  Interior faces are matched through the
    file's vertex ids; with periodic=True the faces on opposite sides of the
    bounding box are matched point by point by coordinates (like modify_mapP),
    not through vertex ids, which coarse periodic meshes make ambiguous."""
    VX, EToV = read_gmsh(path, dim)
    lo, period = VX.min(0), VX.max(0) - VX.min(0)
    m = SimplexMesh(dim, EToV.shape[0], EToV.copy(), VX[EToV].copy(), period, lo,
                    periodic_points=periodic)
    return _orient(m)
