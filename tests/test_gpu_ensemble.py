"""Ensembles (wadges_batch_*): one launch per stage for many problems of uneven
K and N.  Every problem's trajectory must be the one its own single-problem
fused stage produces (that path is checked against the oracle in
test_gpu_parity), for every launch shape, with and without the HIP graph, and
the autotuner must leave the state untouched."""

import numpy as np
import pytest

from helpers import relerr
from scratch_wadge_amd import problem as P

pytestmark = pytest.mark.gpu

FIELDS = ("Q", "rhsQ", "resQ", "Qf")


def mixed_2d():
    # Weak vortices: strong ones are not resolved on these 3x3 .. 6x6 meshes and
    # blow up within a few steps (in the oracle as well).
    specs = [(3, 1), (3, 2), (4, 3), (3, 4), (3, 5), (5, 3), (6, 2), (4, 4)]
    return [P.vortex2d_problem(n=n, N=N, beta=1.0 + 0.1 * q, x0=0.3 * q - 1.0, y0=0.1 * q,
                               seed=q + 1) for q, (n, N) in enumerate(specs)]


def mixed_3d():
    return [P.tgv3d_problem(n=3, N=3, curved=True), P.tgv3d_problem(n=3, N=4, curved=False),
            P.tgv3d_problem(n=3, N=1, curved=True), P.tgv3d_problem(n=4, N=3, curved=False),
            P.tgv3d_problem(n=3, N=2, curved=True)]


def compare(ens, ref, label):
    worst, bitwise = 0.0, True
    for q in range(len(ens.problems)):
        a, b = ens.state(q), ref.state(q)
        for f in FIELDS:
            assert np.isfinite(a[f]).all(), (label, q, f)
            worst = max(worst, relerr(a[f], b[f]))
            bitwise &= bool(np.array_equal(a[f], b[f]))
    print(f"{label}: worst rel L2 {worst:.3e} bitwise {bitwise}", flush=True)
    assert worst <= 1e-14, label


CONFIGS = [dict(grid=0), dict(grid=0, order=1), dict(grid=0, waves_per_group=2),
           dict(grid=0, waves_per_group=4, order=1), dict(grid=3, tiles_per_item=1),
           dict(grid=5, tiles_per_item=2, dynamic=1), dict(grid=7, tiles_per_item=4, dynamic=1, order=1),
           dict(grid=2, tiles_per_item=2, order=1), dict(grid=1000, tiles_per_item=1, dynamic=1)]


@pytest.mark.parametrize("cfg", CONFIGS, ids=[str(c) for c in CONFIGS])
def test_ensemble_2d_matches_single_problem(cfg):
    from scratch_wadge_amd.ensemble import Ensemble, PerProblem
    probs = mixed_2d()
    print("case ensemble 2d", cfg, [(p.N, p.K) for p in probs], flush=True)
    ens, ref = Ensemble(probs), PerProblem(probs)
    ens.set_config(**cfg)
    assert {k: ens.config()[k] for k in cfg} == cfg
    ens.run(3, use_graph=False)
    ref.run(3)
    compare(ens, ref, f"2d {cfg}")


def test_ensemble_graph_both_parities():
    """Graph replays starting from either face-trace buffer, mixed with eager steps."""
    from scratch_wadge_amd.ensemble import Ensemble, PerProblem
    probs = mixed_2d()
    print("case ensemble graph", flush=True)
    ens, ref = Ensemble(probs), PerProblem(probs)
    ens.set_config(grid=4, tiles_per_item=2, dynamic=1)
    ens.run(3, use_graph=True)
    ens.run(1, use_graph=False)
    ens.run(2, use_graph=True)
    ref.run(6)
    compare(ens, ref, "2d graph")


def test_ensemble_3d_matches_single_problem():
    from scratch_wadge_amd.ensemble import Ensemble, PerProblem
    probs = mixed_3d()
    print("case ensemble 3d", [(p.N, p.K) for p in probs], flush=True)
    for cfg in (dict(grid=0), dict(grid=0, waves_per_group=2),
                dict(grid=3, tiles_per_item=2, dynamic=1, order=1)):
        ens, ref = Ensemble(probs), PerProblem(probs)
        ens.set_config(**cfg)
        ens.run(2, use_graph=True)
        ref.run(2)
        compare(ens, ref, f"3d {cfg}")


def test_tune_preserves_state_and_caches():
    from scratch_wadge_amd.ensemble import Ensemble, PerProblem
    probs = mixed_2d()
    print("case ensemble tune", flush=True)
    ens, ref = Ensemble(probs), PerProblem(probs)
    ens.run(1, use_graph=False)
    ref.run(1)
    before = [ens.state(q) for q in range(len(probs))]
    ms = ens.tune(reps=2)
    assert ms > 0
    cfg = ens.config()
    print("tuned", cfg, ms, flush=True)
    for q, b in enumerate(before):
        a = ens.state(q)
        for f in FIELDS:
            assert np.array_equal(a[f], b[f]), (q, f)
    ens.run(2, use_graph=True)
    ref.run(2)
    compare(ens, ref, "2d after tune")
    ens2 = Ensemble(probs)
    assert ens2.tune(reps=2) == -1.0           # same signature: cached
    assert ens2.config() == cfg


def test_plan_invalidate_rereads_operators():
    """Tables prepared from operator arrays are reused until the pointers change
    or the plan is invalidated."""
    import torch
    from scratch_wadge_amd.device import DeviceSolver
    prob = P.vortex2d_problem(n=3, N=3)
    print("case plan invalidate", flush=True)
    a = DeviceSolver(prob)
    r0, _ = a.rhs(prob.rhsQ, prob.Qf)
    a.c["Drq"].mul_(2.0)                        # in place: same pointer
    r_stale, _ = a.rhs(prob.rhsQ, prob.Qf)
    a.plan.invalidate()
    r1, _ = a.rhs(prob.rhsQ, prob.Qf)
    b = DeviceSolver(prob)
    b.c["Drq"].mul_(2.0)
    r_ref, _ = b.rhs(prob.rhsQ, prob.Qf)
    torch.cuda.synchronize()
    assert np.array_equal(r_stale, r0)
    assert np.array_equal(r1, r_ref)
    assert relerr(r1, r0) > 1e-3


def test_config4_workload_against_oracle(oracle):
    """The 64-problem ensemble workload itself (2D isentropic vortices, K
    log-uniform in [128, 8192], N in {2, 3, 4}, strength U[3, 6], centre
    U[-2, 2]^2, seed 0) with the autotuned launch shape and graph-replayed
    steps: every problem bit for bit equal to its own single-problem stages, and
    the smallest and the largest problem of each degree equal to the oracle run
    through the same steps (each problem is the reference's rk_run,
    euler.c:3672-3699, on its own mesh)."""
    from scratch_wadge_amd.ensemble import Ensemble, PerProblem, vortex_ensemble
    probs = vortex_ensemble(64, seed=0)
    steps = 4
    print(f"case config-4 ensemble: 64 problems, K_total {sum(p.K for p in probs)}, "
          f"{steps} steps", flush=True)
    ens, ref = Ensemble(probs), PerProblem(probs)
    ens.tune(reps=2)
    print("tuned launch shape", ens.config(), flush=True)
    ens.run(steps, use_graph=True)
    ref.run(steps)
    for q in range(len(probs)):
        a, b = ens.state(q), ref.state(q)
        for f in FIELDS:
            assert np.isfinite(a[f]).all(), (q, f)
            assert np.array_equal(a[f], b[f]), (q, f)
    picks = set()
    for N in (2, 3, 4):
        qs = [q for q, p in enumerate(probs) if p.N == N]
        assert qs, N
        picks.add(min(qs, key=lambda q: probs[q].K))
        picks.add(max(qs, key=lambda q: probs[q].K))
    for q in sorted(picks):
        p = probs[q]
        st = p.copy_state()
        for _ in range(steps):
            for s in range(5):
                oracle.stage(p, st, P.RK4A[s], P.RK4B[s], p.dt)
        a = ens.state(q)
        errs = {f: relerr(a[f], st[f]) for f in ("Q", "rhsQ", "Qf")}
        print(f"problem {q}: N={p.N} K={p.K} vs oracle " +
              " ".join(f"{f} {e:.2e}" for f, e in errs.items()), flush=True)
        assert max(errs.values()) <= 1e-12, (q, errs)


def test_graph_tables_and_eager_tables_interleaved():
    """A graph captured with one set of operators (B) and calls outside capture
    with another (A) on the same plan, interleaved: every call reads the tables
    of its own operators, equal bit for bit to fresh plans with A and with B."""
    import torch
    from scratch_wadge_amd.device import DeviceSolver
    prob = P.tgv3d_problem(n=3, N=3, curved=True)
    print("case graph / eager operator tables", flush=True)
    names = ("Drq", "Dsq", "Dtq", "VqLq", "VfPq")
    a = DeviceSolver(prob)
    A = {k: a.c[k] for k in names}
    B = {k: (1.5 * v).contiguous() for k, v in A.items()}      # other values, other pointers
    ra, rfa = DeviceSolver(prob).vol_only(prob.rhsQ, prob.Qf)
    b = DeviceSolver(prob)
    b.c.update(B)
    rb, rfb = b.vol_only(prob.rhsQ, prob.Qf)
    assert relerr(ra, rb) > 1e-3
    rq = torch.as_tensor(prob.rhsQ, device=a.device)
    qf = torch.as_tensor(prob.Qf, device=a.device)
    rg, rfg = rq.clone(), torch.zeros_like(qf)

    def eager_a():
        r, rf = a.vol_only(prob.rhsQ, prob.Qf)
        assert np.array_equal(r, ra) and np.array_equal(rf, rfa)

    def replay_b():
        rg.copy_(rq)                                 # vol overwrites its input
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(rg.cpu().numpy(), rb) and np.array_equal(rfg.cpu().numpy(), rfb)

    eager_a()                                        # tables of A prepared and cached
    g = torch.cuda.CUDAGraph()
    a.c.update(B)
    with torch.cuda.graph(g):
        a.vol(rg, qf, rfg)
    a.c.update(A)
    replay_b()
    eager_a()
    replay_b()
    replay_b()
    eager_a()
    eager_a()
