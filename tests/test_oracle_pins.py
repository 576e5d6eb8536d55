"""Pins of the CPU oracle against what the mathematics fixes.

The reference has no RHS-level golden vectors (its only self-checks print), so
each oracle routine is pinned here by closed forms and identities chosen so
that a plausible slip fails: the log mean against (a-b)/(ln a - ln b), the
flux against the Euler flux and Tadmor's entropy-conservation identity, the
variable maps against each other and against the gradient of the entropy, and
the assembled RHS against free-stream preservation, conservation and discrete
entropy conservation on periodic meshes.  Planted mistakes (the reference's
as-shipped 3D log branch) are checked to be rejected.
"""

import math

import numpy as np
import pytest

from helpers import small_problem, rough_state

GAMMA = 1.4


def closed_logmean(a, b):
    return (a - b) / (math.log(a) - math.log(b))


@pytest.mark.parametrize("dim", [2, 3])
def test_logmean_closed_form(oracle, dim):
    rng = np.random.default_rng(0)
    for _ in range(2000):
        a = math.exp(rng.uniform(-3, 3))
        b = a * math.exp(rng.choice([rng.uniform(-2, 2), rng.uniform(-0.05, 0.05)]))
        if a == b:
            continue
        ref = closed_logmean(a, b)
        got = oracle.logmean(dim, a, b)
        # the closed form loses digits to cancellation as b -> a
        tol = 1e-13 + 1e-16 / max(abs(math.log(a / b)), 1e-300)
        assert abs(got - ref) <= tol * abs(ref) * 10, (a, b)


@pytest.mark.parametrize("dim", [2, 3])
def test_logmean_limits_and_symmetry(oracle, dim):
    for a in (1e-3, 0.7, 1.0, 3.0, 250.0):
        assert oracle.logmean(dim, a, a) == pytest.approx(a, rel=1e-15)
        for r in (1.0 + 1e-9, 1.001, 1.02, 1.5, 7.0):
            assert oracle.logmean(dim, a, a * r) == pytest.approx(oracle.logmean(dim, a * r, a), rel=1e-14)
            # the log mean lies between the geometric and arithmetic means
            lm = oracle.logmean(dim, a, a * r)
            assert math.sqrt(a * a * r) * (1 - 1e-14) <= lm <= 0.5 * (a + a * r) * (1 + 1e-14)


def test_logmean_branch_continuity(oracle):
    # both branches agree where they meet (3D: u = f^2 = 1e-4, 2D: 1e-3)
    for dim, th in ((3, 1e-4), (2, 1e-3)):
        f = math.sqrt(th)
        xi = (1 + f) / (1 - f)
        lo = oracle.logmean(dim, xi * (1 - 1e-9), 1.0)
        hi = oracle.logmean(dim, xi * (1 + 1e-9), 1.0)
        assert abs(lo - hi) <= 1e-8 * lo


def test_planted_logmean_bug_is_rejected(oracle):
    """The as-shipped 3D log branch (okl/Euler3D.okl:68) must fail the closed-form pin."""
    a, b = 1.0, 1.5
    oracle.set_logmean_variant(1)
    try:
        bad = oracle.logmean(3, a, b)
    finally:
        oracle.set_logmean_variant(0)
    assert abs(bad - closed_logmean(a, b)) > 1e-3
    assert oracle.logmean(3, a, b) == pytest.approx(closed_logmean(a, b), rel=1e-14)


def _state(dim, rng, amp=0.5):
    rho = 1.0 + amp * rng.uniform(-0.9, 1.0)
    u = rng.uniform(-1, 1, dim)
    p = 1.0 + amp * rng.uniform(-0.9, 1.0)
    E = p / (GAMMA - 1) + 0.5 * rho * u @ u
    return np.concatenate([[rho], rho * u, [E]])


def euler_flux(dim, U):
    rho, m, E = U[0], U[1:-1], U[-1]
    u = m / rho
    p = (GAMMA - 1) * (E - 0.5 * rho * u @ u)
    F = np.zeros((dim, dim + 2))
    for d in range(dim):
        F[d, 0] = m[d]
        F[d, 1:-1] = m[d] * u
        F[d, 1 + d] += p
        F[d, -1] = u[d] * (E + p)
    return F


@pytest.mark.parametrize("dim", [2, 3])
def test_flux_consistency(oracle, dim):
    rng = np.random.default_rng(1)
    for _ in range(200):
        U = _state(dim, rng)
        np.testing.assert_allclose(oracle.ec_flux(dim, GAMMA, U, U), euler_flux(dim, U),
                                   rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_flux_symmetry(oracle, dim):
    rng = np.random.default_rng(2)
    for _ in range(200):
        UL, UR = _state(dim, rng), _state(dim, rng)
        np.testing.assert_allclose(oracle.ec_flux(dim, GAMMA, UL, UR),
                                   oracle.ec_flux(dim, GAMMA, UR, UL), rtol=1e-13, atol=1e-13)


def _tadmor_residual(oracle, dim, UL, UR):
    """(v_L - v_R) . F_d - (psi_L - psi_R)_d with the reference's entropy
    variables v = dU/du for U = -rho s, whose potential is psi_d = (gamma-1) rho u_d."""
    F = oracle.ec_flux(dim, GAMMA, UL, UR)
    vL, vR = oracle.VU(dim, GAMMA, UL), oracle.VU(dim, GAMMA, UR)
    psiL = (GAMMA - 1) * UL[1:-1]
    psiR = (GAMMA - 1) * UR[1:-1]
    return F @ (vL - vR) - (psiL - psiR), np.abs(F).max() * np.abs(vL - vR).max()


@pytest.mark.parametrize("dim", [2, 3])
def test_tadmor_entropy_conservation(oracle, dim):
    rng = np.random.default_rng(3)
    for _ in range(300):
        UL, UR = _state(dim, rng), _state(dim, rng)
        res, scale = _tadmor_residual(oracle, dim, UL, UR)
        assert np.abs(res).max() <= 1e-12 * max(scale, 1.0)


def test_planted_flux_bug_is_rejected(oracle):
    """With the as-shipped log branch the EC identity fails by O(1)."""
    rng = np.random.default_rng(4)
    UL, UR = _state(3, rng, amp=0.9), _state(3, rng, amp=0.9)
    UR[0] *= 1.6
    oracle.set_logmean_variant(1)
    try:
        res, scale = _tadmor_residual(oracle, 3, UL, UR)
    finally:
        oracle.set_logmean_variant(0)
    assert np.abs(res).max() > 1e-3 * scale


@pytest.mark.parametrize("dim", [2, 3])
def test_entropy_variables(oracle, dim):
    rng = np.random.default_rng(5)
    for _ in range(100):
        U = _state(dim, rng)
        V = oracle.VU(dim, GAMMA, U)
        np.testing.assert_allclose(oracle.UV(dim, GAMMA, V), U, rtol=1e-13)

        # v = dU/du for the entropy U(u) = -rho ln(p / rho^gamma)
        def ent(W):
            rho, m, E = W[0], W[1:-1], W[-1]
            p = (GAMMA - 1) * (E - 0.5 * m @ m / rho)
            return -rho * math.log(p / rho ** GAMMA)

        g = np.zeros(dim + 2)
        for c in range(dim + 2):
            h = 1e-6 * max(1.0, abs(U[c]))
            e = np.zeros(dim + 2)
            e[c] = h
            g[c] = (ent(U + e) - ent(U - e)) / (2 * h)
        np.testing.assert_allclose(V, g, rtol=1e-7, atol=1e-7)


# ------------------------------------------------------------------ assembled RHS invariants

CASES = [(2, 1, True), (2, 3, True), (2, 5, True), (3, 1, False), (3, 3, False), (3, 3, True)]


@pytest.mark.parametrize("dim,N,curved", CASES)
def test_free_stream(oracle, dim, N, curved):
    prob = small_problem(dim, N, curved=curved)
    U0 = np.array([1.2, 0.3, -0.4, 0.2, 3.0]) if dim == 3 else np.array([1.2, 0.3, -0.4, 3.0])
    rq = np.tile(np.repeat(U0, prob.Nq), prob.K)
    qf = np.tile(np.repeat(U0, prob.Nf), prob.K)
    r, _ = oracle.rhs(prob, rq, qf)
    assert np.abs(r).max() < 1e-12 * 10 * np.abs(U0).max() ** 2


@pytest.mark.parametrize("dim,N,curved", CASES)
def test_conservation(oracle, dim, N, curved):
    """sum_k sum_i wq_i rhs_i = 0 per field on periodic meshes (the WADG update
    then conserves sum wq J u)."""
    prob = small_problem(dim, N, curved=curved)
    _, rq, qf = rough_state(prob, seed=11)
    r, _ = oracle.rhs(prob, rq, qf)
    R = r.reshape(prob.K, prob.nfields, prob.Nq)
    tot = np.einsum("i,kfi->f", prob.wq, R)
    scale = np.einsum("i,kfi->f", prob.wq, np.abs(R))
    assert np.all(np.abs(tot) <= 1e-13 * scale.max() * 10)


@pytest.mark.parametrize("dim,N,curved", CASES)
def test_entropy_conservation_tau0(oracle, dim, N, curved):
    """With tau = 0 the discrete entropy production sum wq v(u~).rhs vanishes."""
    from scratch_wadge_amd import problem as P
    prob = small_problem(dim, N, curved=curved, tau=0.0)
    _, rq, qf = rough_state(prob, seed=12)
    r, _ = oracle.rhs(prob, rq, qf)
    nf = prob.nfields
    U = rq.reshape(prob.K, nf, prob.Nq).transpose(1, 2, 0)
    V = P.entropy_vars(U)
    terms = prob.wq[:, None] * np.einsum("fik,fik->ik", V, r.reshape(prob.K, nf, prob.Nq).transpose(1, 2, 0))
    assert abs(terms.sum()) <= 1e-13 * np.abs(terms).sum() * 10
    # with the penalty the scheme dissipates entropy (du/dt = -2 M^-1 rhs)
    prob1 = small_problem(dim, N, curved=curved, tau=1.0)
    r1, _ = oracle.rhs(prob1, rq, qf)
    t1 = prob.wq[:, None] * np.einsum("fik,fik->ik", V, r1.reshape(prob.K, nf, prob.Nq).transpose(1, 2, 0))
    assert t1.sum() > 1e3 * abs(terms.sum())


def test_entropy_planted_bug_breaks_conservation(oracle):
    from scratch_wadge_amd import problem as P
    prob = small_problem(3, 3, curved=True, tau=0.0)
    _, rq, qf = rough_state(prob, seed=12)
    oracle.set_logmean_variant(1)
    try:
        r, _ = oracle.rhs(prob, rq, qf)
    finally:
        oracle.set_logmean_variant(0)
    nf = prob.nfields
    U = rq.reshape(prob.K, nf, prob.Nq).transpose(1, 2, 0)
    V = P.entropy_vars(U)
    terms = prob.wq[:, None] * np.einsum("fik,fik->ik", V, r.reshape(prob.K, nf, prob.Nq).transpose(1, 2, 0))
    assert abs(terms.sum()) > 1e-6 * np.abs(terms).sum()


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 3)])
def test_update_identities(oracle, dim, N):
    """Update: the increment is linear in rhs, and the entropy projection is idempotent."""
    prob = small_problem(dim, N, curved=False)
    Q, rq, qf = rough_state(prob, seed=13)
    rng = np.random.default_rng(0)
    r1 = rng.standard_normal(rq.shape)
    r2 = rng.standard_normal(rq.shape)

    def incr(r):
        st = {"Q": Q.copy(), "rhsQ": r.copy(), "Qf": qf.copy(), "resQ": np.zeros_like(Q)}
        oracle.update(prob, 0.0, 1.0, 1.0, st["rhsQ"], st["resQ"], st["Q"], st["Qf"])
        return st["resQ"], st

    a, _ = incr(r1)
    b, _ = incr(r2)
    c, _ = incr(2.0 * r1 + 0.5 * r2)
    np.testing.assert_allclose(c, 2.0 * a + 0.5 * b, rtol=1e-12, atol=1e-12 * np.abs(c).max())
    # affine mesh: J is constant per element, so the WADG inverse mass is the
    # exact one, -2 VqPq rhs / J
    ops = prob.ops
    R = r1.reshape(prob.K, prob.nfields, prob.Nq)
    J = prob.Jq.reshape(prob.K, prob.Nq)
    expect = -2.0 * np.einsum("ij,kfj->kfi", ops.VqPq, R) / J[:, None, :]
    np.testing.assert_allclose(a.reshape(R.shape), expect, rtol=1e-11, atol=1e-12 * np.abs(expect).max())
    # entropy projection: running the update with a zero increment maps (Q, u~)
    # to itself when u~ already is u(P v(Q)) of the same Q
    _, st = incr(np.zeros_like(rq))
    st2 = {k: v.copy() for k, v in st.items()}
    st2["rhsQ"][:] = 0.0
    oracle.update(prob, 0.0, 1.0, 1.0, st2["rhsQ"], st2["resQ"], st2["Q"], st2["Qf"])
    np.testing.assert_allclose(st2["rhsQ"], st["rhsQ"], rtol=1e-12)
    np.testing.assert_allclose(st2["Qf"], st["Qf"], rtol=1e-12)


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 3)])
def test_curved_update_projects_before_and_after_the_weight(oracle, dim, N):
    """The weight-adjusted inverse mass is VqPq diag(1/J) VqPq (okl/Euler3D.okl:
    2586-2703): on a curved mesh (J varying inside elements) an RHS orthogonal
    to the polynomial space (VqPq rhs = 0) gives a zero increment, and every
    increment lies in that space (VqPq r = r).  Dropping the first projection, a
    slip that changes nothing on affine meshes, is planted in the oracle and
    rejected."""
    import os
    from scratch_wadge_amd import problem as P
    if dim == 2:   # the reference's curved vortex mesh (warp a = 1/2)
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "periodicSquare1.msh")
        prob = P.vortex_reference_problem(path, N, curved=True)
    else:
        prob = small_problem(dim, N, curved=True)
    Q, rq, qf = rough_state(prob, seed=21)
    K, nf, Nq = prob.K, prob.nfields, prob.Nq
    VqPq = prob.ops.VqPq
    J = prob.Jq.reshape(K, Nq)
    assert np.ptp(J, axis=1).max() > 1e-3 * np.abs(J).max()      # J varies inside elements
    rng = np.random.default_rng(3)
    g = rng.standard_normal((K, nf, Nq))
    perp = g - np.einsum("ij,kfj->kfi", VqPq, g)                # VqPq perp = 0

    def incr(r):
        st = {"Q": Q.copy(), "rhsQ": r.reshape(-1).copy(), "Qf": qf.copy(), "resQ": np.zeros_like(Q)}
        oracle.update(prob, 0.0, 1.0, 1.0, st["rhsQ"], st["resQ"], st["Q"], st["Qf"])
        return st["resQ"].reshape(K, nf, Nq)

    scale = np.abs(incr(g)).max()
    assert np.abs(incr(perp)).max() <= 1e-12 * scale
    r = incr(g)
    np.testing.assert_allclose(np.einsum("ij,kfj->kfi", VqPq, r), r, rtol=0, atol=1e-12 * scale)
    try:
        oracle.set_update_variant(1)
        planted = incr(perp)
    finally:
        oracle.set_update_variant(0)
    assert np.abs(planted).max() > 1e-6 * scale


def test_penalty_constant_pinned_by_published_errors(oracle):
    """The Lax-Friedrichs penalty tau/4 max(lamM, lamP) sJ (okl/Euler3D.okl:
    3316-3330; okl/Euler2D.okl has tau = 1): the oracle's isentropic-vortex L2
    error on the reference's periodicSquare2.msh at N = 3, T = 5 equals the
    reference's published 0.05160533 (docs/figs/errors_3d.m:6) to 2e-4, while
    the penalty doubled (tau = 2, a constant of 1/2) or left out (tau = 0)
    misses it by more than 20 times that."""
    import os
    from scratch_wadge_amd import problem as P
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "periodicSquare2.msh")
    pub = 0.05160533

    def err(tau):
        prob = P.vortex_reference_problem(path, 3, curved=False, cfl=0.5, tau=tau)
        nsteps, dt = P.rk_steps(prob.dt, 5.0)
        st = prob.copy_state()
        for _ in range(nsteps):
            for s in range(5):
                oracle.stage(prob, st, P.RK4A[s], P.RK4B[s], dt)
        K, nf, Nq = prob.K, prob.nfields, prob.Nq
        wJ = prob.wq[None, :] * prob.Jq.reshape(K, Nq)
        d = (st["Q"] - P.vortex_reference_exact(prob, 5.0)).reshape(K, nf, Nq)
        return float(np.sqrt(np.sum(wJ[:, None, :] * d * d)))

    assert abs(err(1.0) - pub) <= 2e-4 * pub
    for tau in (0.0, 2.0):
        assert abs(err(tau) - pub) > 4e-3 * pub
