"""Host-side setup: quadrature rules, reference operators, meshes, geometry, maps."""

import itertools
import math

import numpy as np
import pytest

from scratch_wadge_amd import mesh as M
from scratch_wadge_amd import operators as O
from scratch_wadge_amd import problem as P
from scratch_wadge_amd import refelem


def simplex_monomial(dim, exps):
    """Exact integral of prod x_d^{e_d} over the reference simplex (vertices at -1, 1),
    by x = -1 + 2 xi on the unit simplex and the Dirichlet integral
    int xi^a = prod a_d! / (dim + sum a)!."""
    tot = 0.0
    for ks in itertools.product(*[range(e + 1) for e in exps]):
        c = 1.0
        for e, k in zip(exps, ks):
            c *= math.comb(e, k) * 2.0 ** k * (-1.0) ** (e - k)
        num = math.prod(math.factorial(k) for k in ks)
        tot += c * num / math.factorial(dim + sum(ks))
    return tot * 2.0 ** dim


@pytest.mark.parametrize("key", ["tri2", "tri4", "tri6", "tri8", "tri10", "tet2", "tet4", "tet6", "tet8"])
def test_rule_exactness(key):
    rules = O._rules()
    r = rules[key]
    x = np.array(r["points"])
    w = np.array(r["weights"])
    dim = x.shape[1]
    deg = r["degree"]
    assert np.all(w > 0)
    if dim == 2:
        lam = np.stack([-(x[:, 0] + x[:, 1]) / 2, (1 + x[:, 0]) / 2, (1 + x[:, 1]) / 2], 1)
    else:
        lam = np.stack([-(1 + x.sum(1)) / 2, (1 + x[:, 0]) / 2, (1 + x[:, 1]) / 2, (1 + x[:, 2]) / 2], 1)
    assert lam.min() > 0
    for exps in itertools.product(range(deg + 1), repeat=dim):
        if sum(exps) > deg:
            continue
        q = np.sum(w * np.prod(x ** np.array(exps), 1))
        assert q == pytest.approx(simplex_monomial(dim, exps), abs=1e-13), exps
    expected = {"tri2": 3, "tri4": 6, "tri6": 12, "tri8": 16, "tri10": 25, "tet2": 4, "tet4": 11,
                "tet6": 23, "tet8": 44}
    assert len(w) == expected[key]


@pytest.mark.parametrize("deg", [2, 4, 6, 8, 10])
def test_triangle_rules_fully_symmetric(deg):
    """Face rules must be invariant under the 6 vertex permutations so that face
    points of neighbouring elements coincide for every orientation."""
    a, b, w = O.tri_rule(deg)
    lam = np.stack([-(a + b) / 2, (1 + a) / 2, (1 + b) / 2], 1)
    key = lambda L, W: sorted(zip(np.round(L[:, 0], 12), np.round(L[:, 1], 12), np.round(W, 12)))
    base = key(lam, w)
    for perm in itertools.permutations(range(3)):
        assert key(lam[:, perm], w) == base


@pytest.mark.parametrize("dim,N", [(2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (3, 1), (3, 2), (3, 3), (3, 4)])
def test_hybridized_operators(dim, N):
    ops = O.build_ref_ops(dim, N)
    Np, Nq = ops.Np, ops.Nq
    np.testing.assert_allclose(ops.Pq @ ops.Vq, np.eye(Np), atol=1e-12)
    np.testing.assert_allclose(ops.VqPq @ ops.VqPq, ops.VqPq, atol=1e-12)
    W = np.diag(ops.wq)
    for d in range(dim):
        D = ops.Drst[d]
        # skew-symmetry of W Drq (hybridized SBP: the boundary term sits in the face blocks)
        S = W @ D
        np.testing.assert_allclose(S + S.T, 0.0, atol=1e-12 * np.abs(S).max())
        # free-stream: Drq 1 + 1/2 VqLq nref = 0
        np.testing.assert_allclose(D @ np.ones(Nq) + 0.5 * ops.VqLq @ ops.nref[:, d], 0.0, atol=1e-12)
    # lift identity: Vq^T W VqLq = Vf^T Wf
    np.testing.assert_allclose(ops.Vq.T @ W @ ops.VqLq, ops.Vf.T * ops.wf[None, :], atol=1e-12)
    # modal derivative is exact on degree-N polynomials: d/dr of phi_k
    r = ops.rq
    if dim == 2:
        V = refelem.vandermonde_2d(N, r[:, 0], r[:, 1])
        Vr = refelem.grad_vandermonde_2d(N, r[:, 0], r[:, 1])
    else:
        V = refelem.vandermonde_3d(N, r[:, 0], r[:, 1], r[:, 2])
        Vr = refelem.grad_vandermonde_3d(N, r[:, 0], r[:, 1], r[:, 2])
    for d in range(dim):
        np.testing.assert_allclose(V @ ops.Dmodal[d], Vr[d], atol=1e-11)


def test_grad_vandermonde_fd():
    rng = np.random.default_rng(0)
    p = rng.uniform(-0.8, -0.1, (5, 3))
    h = 1e-6
    G = refelem.grad_vandermonde_3d(3, p[:, 0], p[:, 1], p[:, 2])
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        fd = (refelem.vandermonde_3d(3, *(p + e).T) - refelem.vandermonde_3d(3, *(p - e).T)) / (2 * h)
        np.testing.assert_allclose(G[d], fd, atol=1e-7)


@pytest.mark.parametrize("dim,curved", [(2, False), (3, False), (3, True)])
def test_mesh_and_maps(dim, curved):
    if dim == 2:
        prob = P.vortex2d_problem(n=5, N=3, perturb=0.05, seed=2)
        vol = 100.0
    else:
        prob = P.tgv3d_problem(n=3, N=3, curved=curved)
        vol = (2 * np.pi) ** 3
    K, Nf, nf = prob.K, prob.Nf, prob.nfields
    assert prob.meta["matched_face_points"] * 2 == K * Nf
    m = prob.mapPq.astype(np.int64)
    e = m // (Nf * nf)
    node = m - e * Nf * nf
    flat = node + e * Nf
    own = np.arange(K * Nf)
    assert np.all(flat != own)
    assert np.array_equal(flat[flat], own)  # involution
    geo = prob.meta["geometry"]
    xf = np.concatenate([geo.xf[d].T.reshape(-1)[None] for d in range(dim)], 0)
    diff = xf[:, own] - xf[:, flat]
    period = 10.0 if dim == 2 else 2 * np.pi
    diff = diff - period * np.round(diff / period)
    assert np.abs(diff).max() < 1e-12
    # scaled normals are opposite across a face
    fg = prob.fgeo.reshape(K, dim + 1, Nf)
    nJ = np.concatenate([fg[:, d, :].reshape(-1)[None] for d in range(dim)], 0)
    np.testing.assert_allclose(nJ[:, own], -nJ[:, flat], atol=1e-10 * np.abs(nJ).max())
    # Jacobian positive, integrates to the box volume
    J = prob.Jq.reshape(K, prob.Nq)
    assert J.min() > 0
    assert np.einsum("i,ki->", prob.wq, J) == pytest.approx(vol, rel=1e-10)


def test_geometric_conservation_law():
    """Discrete GCL of the Kopriva geofacs: sum_r D_r (r_d J) = 0 at the degree-N
    nodes (the check operators.cpp:683-687 prints)."""
    ops = O.build_ref_ops(3, 3)
    m = M.periodic_box_tet(3)
    warp = lambda x, y, z: M.tg_warp(x, y, z, 0.1)
    geo = M.build_geometry(m, ops, warp=warp)
    # geofacs at quadrature points are Vq * nodal values; recover modal coefficients by Pq
    G = geo.vgeo  # (K, 9, Nq)
    div = np.zeros((m.K, 3, ops.Np))
    for d in range(3):
        for r in range(3):
            c = np.einsum("mi,ki->km", ops.Pq, G[:, r * 3 + d, :])
            div[:, d, :] += c @ ops.Dmodal[r].T
    assert np.abs(div).max() < 1e-12 * np.abs(G).max()


def test_vortex_ensemble_generator():
    """The ensemble workload: K log-uniform in [128, 8192] rounded to 2 n^2,
    N in {2, 3, 4}, reproducible from the seed."""
    import numpy as np
    from scratch_wadge_amd.ensemble import vortex_ensemble
    a = vortex_ensemble(12, seed=5)
    b = vortex_ensemble(12, seed=5)
    assert [(p.K, p.N) for p in a] == [(p.K, p.N) for p in b]
    for p in a:
        n = int(round(np.sqrt(p.K / 2)))
        assert p.K == 2 * n * n and 64 <= p.K <= 2 * 64 * 64
        assert p.N in (2, 3, 4) and p.dim == 2 and p.dt > 0
    assert np.array_equal(a[0].Q, b[0].Q)


def test_node_data_parser(tmp_path):
    """operators.parse_node_data / use_node_data on a header in the reference's
    table format (C arrays p_rq_N{M}, ... p_wfq_N{M}), written here from this
    package's own degree-2 rules."""
    tri = O._rules()["tri2"]
    tet = O._rules()["tet2"]
    x, w = np.array(tri["points"]), np.array(tri["weights"])
    X, W = np.array(tet["points"]), np.array(tet["weights"])

    def arr(name, v):
        return f"double {name}[{len(v)}] = {{" + " , ".join(f"{a:.17g}" for a in v) + "};\n"

    src = "int p_Nq_N2 = 4;\n" + arr("p_rq_N2", X[:, 0]) + arr("p_sq_N2", X[:, 1]) + \
        arr("p_tq_N2", X[:, 2]) + arr("p_wq_N2", W) + arr("p_rfq_N2", x[:, 0]) + \
        arr("p_sfq_N2", x[:, 1]) + arr("p_wfq_N2", w) + arr("p_r_N2", X[:, 0])
    path = tmp_path / "node_data.h"
    path.write_text(src)
    rules = O.parse_node_data(str(path))
    assert sorted(rules) == ["tet2", "tri2"]
    np.testing.assert_array_equal(np.array(rules["tet2"]["points"]), X)
    np.testing.assert_array_equal(np.array(rules["tri2"]["weights"]), w)
    try:
        assert O.use_node_data(str(path)) == ["tet2", "tri2"]
        ops = O.build_ref_ops(3, 1)
        np.testing.assert_array_equal(ops.wq, W)
    finally:
        O.use_node_data(None)
    assert O.build_ref_ops(3, 1).Nq == 4
