"""End-to-end runs of the small workloads on the GPU against the oracle run
through the same steps.

* 2D isentropic vortex, 16 x 16 x 2 = 512 perturbed triangles, N = 3, CFL 0.5,
  200 LSRK45 steps: the final state equals the oracle's to accumulated round-off,
  and both are close to the exact translated vortex with the same error.
* 3D Taylor-Green on 16^3 x 6 = 24,576 affine tets, N = 3: 20 steps against the
  oracle; the total entropy sum wq J (-rho s) does not increase (the penalty is
  dissipative; the kinetic energy alone may oscillate with the pressure waves)."""

import numpy as np
import pytest

from helpers import relerr
from scratch_wadge_amd import problem as P

pytestmark = pytest.mark.gpu


def _oracle_run(oracle, prob, nsteps):
    st = prob.copy_state()
    for _ in range(nsteps):
        for s in range(5):
            oracle.stage(prob, st, P.RK4A[s], P.RK4B[s], prob.dt)
    return st


def test_vortex_2d_200_steps(oracle):
    from scratch_wadge_amd.device import DeviceSolver
    prob = P.vortex2d_problem(n=16, N=3)
    steps = 200
    print(f"case vortex K={prob.K} dt={prob.dt:.4e} steps={steps}", flush=True)
    dev = DeviceSolver(prob)
    dev.run(steps, fused=True, graph=True)
    g = dev.get_state()
    ref = _oracle_run(oracle, prob, steps)
    e = relerr(g["Q"], ref["Q"])
    # exact solution: the vortex translated by (1, 1) t, periodic on [-5, 5]^2
    t = steps * prob.dt
    x, y = prob.xq
    xs = np.mod(x - t + 5.0, 10.0) - 5.0
    ys = np.mod(y - t + 5.0, 10.0) - 5.0
    Uex = P.vortex_2d(xs, ys)                                     # (4, Nq, K)
    Qex = P._to_layout(Uex)
    w = (prob.wq[:, None] * prob.Jq.reshape(prob.K, prob.Nq).T)   # (Nq, K)

    def l2(Q):
        d = (Q - Qex).reshape(prob.K, prob.nfields, prob.Nq).transpose(1, 2, 0)
        return float(np.sqrt(np.sum(w[None] * d * d)))

    eg, eo = l2(g["Q"]), l2(ref["Q"])
    diag = dev.diagnostics(Qex)
    print(f"GPU vs oracle after {steps} steps: {e:.2e}; L2 error vs exact: GPU {eg:.6e} "
          f"oracle {eo:.6e} device reduction {diag['l2_error']:.6e}", flush=True)
    assert e < 1e-10
    assert abs(eg - eo) <= 1e-8 * eo
    assert abs(diag["l2_error"] - eg) <= 1e-10 * eg
    assert eg < 0.05


def test_taylor_green_affine_20_steps(oracle):
    from scratch_wadge_amd.device import DeviceSolver
    prob = P.tgv3d_problem(n=16, N=3, curved=False)
    steps = 20
    print(f"case TGV affine K={prob.K} steps={steps}", flush=True)
    dev = DeviceSolver(prob)
    ent = [dev.diagnostics()["entropy"]]
    for _ in range(4):
        dev.run(steps // 4, fused=True)
        ent.append(dev.diagnostics()["entropy"])
    g = dev.get_state()
    ref = _oracle_run(oracle, prob, steps)
    e = relerr(g["Q"], ref["Q"])
    print(f"GPU vs oracle after {steps} steps: {e:.2e}; entropy {ent}", flush=True)
    assert e < 1e-11
    assert all(b <= a + 1e-13 * abs(a) for a, b in zip(ent, ent[1:]))
