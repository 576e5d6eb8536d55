"""Mesh ingest (gmsh 2.2 reader), face connectivity with the reference's
orientation conventions, and periodic face-point matching, on the reference's
own mesh files (tests/golden)."""

import os

import numpy as np
import pytest

from helpers import has_rule
from scratch_wadge_amd import mesh as M
from scratch_wadge_amd import problem as P
from scratch_wadge_amd.operators import build_ref_ops

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_conn_tet_connectivity_all_orientations():
    VX, EToV = M.read_gmsh(os.path.join(GOLD, "test_conn_tet.msh"), 3)
    assert VX.shape == (720, 3) and EToV.shape == (288, 4)
    EToE, EToF, EToO = M.connect(EToV, 3)
    K = EToV.shape[0]
    matched = EToE != np.arange(K)[:, None]
    # each pair of tets (2p, 2p+1) shares exactly one face
    assert matched.sum() == K
    assert np.all(EToE[matched] // 2 == np.nonzero(matched)[0] // 2)
    # the relation the reference checks (euler.c:1663-1705), for every face
    fv, ov = M.FTOFV[3], M.OTOFV[3]
    for e in range(K):
        for f in range(4):
            e1, f1, o = EToE[e, f], EToF[e, f], EToO[e, f]
            assert np.array_equal(EToV[e, fv[f][ov[o]]], EToV[e1, fv[f1]])
    # neighbour of the neighbour is the element itself
    assert np.array_equal(EToE[EToE, EToF], np.repeat(np.arange(K)[:, None], 4, 1))
    # all six orientations occur, equally often
    assert np.array_equal(np.bincount(EToO[matched], minlength=6), [48] * 6)


@pytest.mark.parametrize("N", [1, 3, 4])
def test_conn_tet_face_points_coincide(N):
    if not has_rule(3, N):
        pytest.skip("no quadrature rule")
    m = M.mesh_from_gmsh(os.path.join(GOLD, "test_conn_tet.msh"), 3, periodic=False)
    ops = build_ref_ops(3, N)
    geo = M.build_geometry(m, ops)
    mapPq, nmatched = M.build_mapPq(m, ops)
    Nf = ops.Nfq * ops.Nfaces
    assert nmatched == m.K * ops.Nfq // 2            # matched pairs: one face per element
    mp = mapPq.reshape(-1).astype(np.int64)
    e_n = mp // (Nf * 5)
    node = mp - e_n * Nf * 5
    xf = geo.xf.transpose(0, 2, 1).reshape(3, -1)   # (3, K*Nf) physical face points
    own = np.arange(m.K * Nf)
    other = e_n * Nf + node
    moved = other != own
    assert moved.sum() == m.K * ops.Nfq             # one shared face per element
    assert np.abs(xf[:, own[moved]] - xf[:, other[moved]]).max() < 1e-14 * np.abs(xf).max()


@pytest.mark.parametrize("name,dim,N", [("periodicSquare1.msh", 2, 3), ("periodicCubeCoarse.msh", 3, 3)])
def test_periodic_mesh_all_points_matched_and_free_stream(oracle, name, dim, N):
    m = M.mesh_from_gmsh(os.path.join(GOLD, name), dim)
    ops = build_ref_ops(dim, N)
    Nf = ops.Nfq * ops.Nfaces
    geo = M.build_geometry(m, ops)
    mapPq, nmatched = M.build_mapPq(m, ops, xf=geo.xf)
    assert nmatched == m.K * Nf // 2                # every face point has a partner
    mp = mapPq.reshape(-1).astype(np.int64)
    e_n = mp // (Nf * (dim + 2))
    other = e_n * Nf + (mp - e_n * Nf * (dim + 2))
    xf = geo.xf.transpose(0, 2, 1).reshape(dim, -1)
    d = xf[:, other] - xf[:, np.arange(m.K * Nf)]
    d = d - np.round(d / m.period[:, None]) * m.period[:, None]
    assert np.abs(d).max() < 1e-10                  # partners coincide modulo the period
    # area / volume of the box
    wJ = ops.wq[None, :] * geo.Jq.reshape(m.K, -1)
    assert abs(wJ.sum() - np.prod(m.period)) < 1e-10 * np.prod(m.period)
    # a constant state is preserved: the oracle's RHS vanishes
    const = lambda *x: np.stack([np.ones_like(x[0])] + [0.3 * np.ones_like(x[0])] * dim +
                                [np.full_like(x[0], 2.5)], 0)
    prob = P.make_problem(dim, N, m, const)
    r, _ = oracle.rhs(prob, prob.rhsQ.copy(), prob.Qf.copy())
    assert np.abs(r).max() < 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize("name,dim,N", [("periodicSquare1.msh", 2, 3), ("periodicCubeCoarse.msh", 3, 3)])
def test_gpu_rhs_on_reference_mesh(oracle, name, dim, N):
    """The HIP RHS on a mesh read from the reference's gmsh files agrees with the
    oracle."""
    from helpers import relerr
    from scratch_wadge_amd.device import DeviceSolver
    m = M.mesh_from_gmsh(os.path.join(GOLD, name), dim)

    def ic(*x):
        L = m.period
        ph = sum(np.sin(2 * np.pi * (x[d] - m.lo[d]) / L[d]) for d in range(dim))
        rho = 1.0 + 0.2 * ph
        vel = [0.3 * np.cos(2 * np.pi * (x[(d + 1) % dim] - m.lo[(d + 1) % dim]) / L[(d + 1) % dim])
               for d in range(dim)]
        p = 1.0 + 0.1 * ph
        E = p / 0.4 + 0.5 * rho * sum(v * v for v in vel)
        return np.stack([rho] + [rho * v for v in vel] + [E], 0)

    prob = P.make_problem(dim, N, m, ic)
    print(f"case gmsh {name} K={prob.K}", flush=True)
    dev = DeviceSolver(prob)
    r_g, _ = dev.rhs(prob.rhsQ, prob.Qf)
    r_o, _ = oracle.rhs(prob, prob.rhsQ.copy(), prob.Qf.copy())
    e = relerr(r_g, r_o)
    print(f"gmsh {name}: rel L2 {e:.2e}", flush=True)
    assert e < 1e-12
