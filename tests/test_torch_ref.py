"""The torch restatement used as the reference gradient agrees with the C++
oracle on the forward pass (volume, surface and update terms), so that gradients
taken through it are gradients of the reference's algorithm."""

import numpy as np
import pytest
import torch

import torch_ref as TR
from helpers import has_rule, relerr, rough_state, small_problem
from scratch_wadge_amd.problem import RK4A, RK4B


@pytest.mark.parametrize("dim,N", [(2, 2), (2, 3), (3, 3)])
def test_forward_matches_oracle(oracle, dim, N):
    if not has_rule(dim, N):
        pytest.skip("no quadrature rule")
    prob = small_problem(dim, N)
    Q, uq, uf = rough_state(prob, seed=2, amp=0.2 if dim == 2 else 0.05)
    ops = TR.Ops(prob, "cpu")
    T = lambda a: torch.as_tensor(a, dtype=torch.float64)
    r_t = TR.rhs(ops, T(uq), T(uf)).numpy()
    r_o, _ = oracle.rhs(prob, uq.copy(), uf.copy())
    assert relerr(r_t, r_o) < 1e-13
    rng = np.random.default_rng(0)
    res = 1e-3 * rng.standard_normal(Q.shape)
    fa, fb, dt = RK4A[2], RK4B[2], 1e-3
    outs = TR.update(ops, fa, fb, dt, T(r_o), T(res), T(Q))
    st = {"rhsQ": r_o.copy(), "resQ": res.copy(), "Q": Q.copy(), "Qf": uf.copy()}
    oracle.update(prob, fa, fb, dt, st["rhsQ"], st["resQ"], st["Q"], st["Qf"])
    for name, t in zip(("Q", "resQ", "rhsQ", "Qf"), outs):
        assert relerr(t.numpy(), st[name]) < 1e-13, name


def test_projection_matches_problem_setup():
    prob = small_problem(2, 3)
    ops = TR.Ops(prob, "cpu")
    uq, uf = TR.project(ops, torch.as_tensor(prob.Q))
    assert relerr(uq.numpy(), prob.rhsQ) < 1e-13
    assert relerr(uf.numpy(), prob.Qf) < 1e-13
