"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

Same inputs on both sides; the bar is the path's RHS tolerance: relative
L2 difference <= 1e-12 for the volume, surface and update outputs, stage by
stage.
"""

import numpy as np
import pytest

from helpers import has_rule, relerr, rough_state, small_problem

pytestmark = pytest.mark.gpu

TOL = 1e-12

CASES = [(2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (3, 1), (3, 2), (3, 3), (3, 4)]


def _solver(prob):
    from scratch_wadge_amd.device import DeviceSolver
    return DeviceSolver(prob)


@pytest.mark.parametrize("dim,N", CASES)
@pytest.mark.parametrize("state", ["ic", "rough"])
def test_vol_surf_parity(oracle, dim, N, state):
    if not has_rule(dim, N):
        pytest.skip("no quadrature rule for this degree")
    prob = small_problem(dim, N)
    if state == "rough":
        Q, rq, qf = rough_state(prob, seed=dim * 10 + N)
    else:
        rq, qf = prob.rhsQ, prob.Qf
    print(f"case dim={dim} N={N} state={state} K={prob.K}", flush=True)
    dev = _solver(prob)
    # volume kernel alone
    r_gpu, rf_gpu = dev.vol_only(rq, qf)
    r_ref = rq.copy()
    rf_ref = np.zeros_like(qf)
    oracle.vol(prob, r_ref, qf, rf_ref)
    e_vol, e_volf = relerr(r_gpu, r_ref), relerr(rf_gpu, rf_ref)
    # volume + surface
    r_gpu, _ = dev.rhs(rq, qf)
    r_ref, _ = oracle.rhs(prob, rq, qf)
    e_rhs = relerr(r_gpu, r_ref)
    print(f"rel L2: vol {e_vol:.2e} volf {e_volf:.2e} rhs {e_rhs:.2e}", flush=True)
    assert e_vol < TOL and e_volf < TOL and e_rhs < TOL


@pytest.mark.parametrize("dim,N", CASES)
def test_update_parity(oracle, dim, N):
    if not has_rule(dim, N):
        pytest.skip("no quadrature rule for this degree")
    prob = small_problem(dim, N)
    Q, rq, qf = rough_state(prob, seed=7)
    rng = np.random.default_rng(1)
    rhs = 0.1 * rng.standard_normal(rq.shape)
    res = 0.01 * rng.standard_normal(rq.shape)
    print(f"case update dim={dim} N={N} K={prob.K}", flush=True)
    dev = _solver(prob)
    dev.set_state(Q=Q, rhsQ=rhs, Qf=qf, resQ=res)
    fa, fb, dt = -0.41, 0.37, 1e-3
    dev.update(fa, fb, dt)
    g = dev.get_state()
    ref = {"Q": Q.copy(), "rhsQ": rhs.copy(), "Qf": qf.copy(), "resQ": res.copy()}
    oracle.update(prob, fa, fb, dt, ref["rhsQ"], ref["resQ"], ref["Q"], ref["Qf"])
    for k in ("Q", "resQ", "rhsQ", "Qf"):
        assert relerr(g[k], ref[k]) < TOL, k


def _oracle_run(oracle, prob, steps, noise=0.0, seed=0):
    """LSRK45 in the oracle; noise > 0 multiplies every RHS by (1 + noise*N(0,1))
    before the update, which measures how round-off in the RHS propagates."""
    from scratch_wadge_amd.problem import RK4A, RK4B
    rng = np.random.default_rng(seed)
    st = prob.copy_state()
    for _ in range(steps):
        for s in range(5):
            rf = np.zeros_like(st["Qf"])
            oracle.vol(prob, st["rhsQ"], st["Qf"], rf)
            oracle.surf(prob, st["Qf"], rf, st["rhsQ"])
            if noise:
                st["rhsQ"] *= 1.0 + noise * rng.standard_normal(st["rhsQ"].shape)
            oracle.update(prob, RK4A[s], RK4B[s], prob.dt, st["rhsQ"], st["resQ"], st["Q"], st["Qf"])
    return st


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 3)])
def test_rk_steps_parity(oracle, dim, N):
    """Five full LSRK45 steps (25 stages) from the projected initial state.

    The solution, the projected state and the face traces must agree to 1e-12.
    The RK register resQ is a small difference of large stage increments: its
    round-off floor is measured by rerunning the oracle with 1e-15 relative
    noise on every RHS, and the GPU must stay within 10x of that floor."""
    prob = small_problem(dim, N)
    print(f"case rk dim={dim} N={N} K={prob.K}", flush=True)
    dev = _solver(prob)
    dev.run(5)
    g = dev.get_state()
    st = _oracle_run(oracle, prob, 5)
    floor = _oracle_run(oracle, prob, 5, noise=1e-15)
    errs = {k: relerr(g[k], st[k]) for k in ("Q", "resQ", "rhsQ", "Qf")}
    floors = {k: relerr(floor[k], st[k]) for k in errs}
    print("relative L2 after 25 stages:", errs, "round-off floor:", floors, flush=True)
    for k in ("Q", "rhsQ", "Qf"):
        assert errs[k] < 1e-12, k
    assert errs["resQ"] < max(10 * floors["resQ"], 1e-12)


@pytest.mark.parametrize("dim,N", [(3, 3), (3, 4), (2, 1), (2, 2)])
def test_deterministic(oracle, dim, N):
    """Two RHS evaluations of the same state give the same bits: every LDS
    accumulation has a program-fixed order, including the shapes whose lane
    groups add into shared addresses in separate instructions (3D N = 4 face-point
    walk, 2D N = 1, 2 face-by-face walk)."""
    if not has_rule(dim, N):
        pytest.skip("no quadrature rule for this degree")
    prob = small_problem(dim, N)
    dev = _solver(prob)
    a, af = dev.rhs(prob.rhsQ, prob.Qf)
    b, bf = dev.rhs(prob.rhsQ, prob.Qf)
    assert np.array_equal(a, b) and np.array_equal(af, bf)


@pytest.mark.parametrize("dim,N", CASES)
def test_fused_stage_parity(oracle, dim, N):
    """wadges_stage (vol + surf + update in one launch) against the oracle's
    vol -> surf -> update, one stage from a strongly varying state."""
    if not has_rule(dim, N):
        pytest.skip("no quadrature rule for this degree")
    prob = small_problem(dim, N)
    # the projected initial state (2D: vortex, both log-mean branches; 3D: TGV);
    # rough_state is not used here: on the 3-cell curved 3D mesh its face-point
    # entropy extrapolation is admissible but extreme, and one update step from
    # it leaves the admissible set in the oracle as well
    Q, rq, qf = prob.Q, prob.rhsQ, prob.Qf
    res = 1e-4 * np.random.default_rng(2).standard_normal(rq.shape)
    print(f"case fused stage dim={dim} N={N} K={prob.K}", flush=True)
    dev = _solver(prob)
    dev.set_state(Q=Q, rhsQ=rq, Qf=qf, resQ=res)
    dev.stage_fused(2, 1e-3)
    g = dev.get_state()
    ref = {"Q": Q.copy(), "rhsQ": rq.copy(), "Qf": qf.copy(), "resQ": res.copy()}
    from scratch_wadge_amd.problem import RK4A, RK4B
    oracle.stage(prob, ref, RK4A[2], RK4B[2], 1e-3)
    assert all(np.isfinite(v).all() for v in ref.values()), "test state left the admissible set"
    errs = {k: relerr(g[k], ref[k]) for k in ref}
    print("fused stage rel L2:", errs, flush=True)
    for k, e in errs.items():
        assert e < TOL, k


@pytest.mark.parametrize("dim,N", [(2, 3), (3, 3)])
def test_fused_run_matches_split_run(dim, N):
    """Five LSRK45 steps with the fused stage and with the three-kernel stage."""
    prob = small_problem(dim, N)
    a = _solver(prob)
    a.run(5)
    b = _solver(prob)
    b.run(5, fused=True)
    ga, gb = a.get_state(), b.get_state()
    for k in ("Q", "rhsQ", "Qf"):
        assert relerr(gb[k], ga[k]) < 1e-12, k
