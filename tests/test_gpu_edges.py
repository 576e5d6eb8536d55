"""Edge cases of the C ABI on the GPU: empty launches (K = 0) and ragged ones
(K not a multiple of the elements a workgroup holds, within one workgroup and
across several) of the volume, surface, update and fused stage calls, which must
compute the first K elements as the oracle does and leave every other element
untouched; an ensemble that contains an empty problem."""

import dataclasses

import numpy as np
import pytest
import torch

from helpers import relerr, rough_state, small_problem
from scratch_wadge_amd.problem import RK4A, RK4B

pytestmark = pytest.mark.gpu

CASES = [(2, 1), (2, 2), (2, 3), (3, 3), (3, 4)]


def _geo(dev):
    from scratch_wadge_amd.device import _ptr
    c = dev.c
    return [_ptr(c[n]) for n in ("vgeo", "vfgeo", "fgeo", "mapPq", "nrJ", "nsJ", "ntJ", "Drq", "Dsq",
                                 "Dtq", "VqLq", "VqPq", "VfPq", "Jq")]


@pytest.mark.parametrize("dim,N", CASES)
@pytest.mark.parametrize("Ksub", [0, 1, 3, 7, 17, 29])
def test_ragged_and_empty_launches(oracle, dim, N, Ksub):
    from scratch_wadge_amd.device import DeviceSolver, _ptr
    prob = small_problem(dim, N)
    Q, rq, qf = rough_state(prob, seed=3, amp=0.2 if dim == 2 else 0.05)
    print(f"case ragged dim={dim} N={N} K={Ksub} of {prob.K}", flush=True)
    nq, nfc = prob.Nq * prob.nfields, prob.Nf * prob.nfields
    sub = dataclasses.replace(prob, K=Ksub)
    dev = DeviceSolver(prob)
    # volume + surface on the first Ksub elements
    r = torch.as_tensor(rq.copy(), device="cuda:0")
    qfd = torch.as_tensor(qf.copy(), device="cuda:0")
    rf = torch.full_like(qfd, 7.0)
    g = _geo(dev)
    dev.plan.vol(Ksub, g[0], g[1], g[4], g[5], g[6], g[7], g[8], g[9], None, g[10], g[12],
                 _ptr(dev.s["Q"]), _ptr(qfd), _ptr(r), _ptr(rf), dev.stream())
    dev.plan.surf(Ksub, g[2], g[3], g[10], _ptr(qfd), _ptr(rf), _ptr(r), dev.stream())
    torch.cuda.synchronize()
    r_o = rq.copy()
    rf_o = np.full_like(qf, 7.0)
    oracle.vol(sub, r_o, qf, rf_o)
    oracle.surf(sub, qf, rf_o, r_o)
    rg = r.cpu().numpy()
    assert np.array_equal(rg[Ksub * nq:], rq[Ksub * nq:]), "elements beyond K were written"
    assert np.array_equal(rf.cpu().numpy()[Ksub * nfc:], rf_o[Ksub * nfc:])
    if Ksub:
        assert relerr(rg[:Ksub * nq], r_o[:Ksub * nq]) < 1e-12
    # update on the first Ksub elements
    rng = np.random.default_rng(5)
    st = {"Q": Q.copy(), "resQ": 0.01 * rng.standard_normal(Q.shape),
          "rhsQ": 0.1 * rng.standard_normal(rq.shape), "Qf": qf.copy()}
    dev.set_state(**st)
    s = dev.s
    dev.plan.update(Ksub, g[13], g[11], g[12], -0.41, 0.37, 1e-3, _ptr(s["rhsQ"]), _ptr(s["resQ"]),
                    _ptr(s["Q"]), _ptr(s["Qf"]), dev.stream())
    got = dev.get_state()
    ref = {k: v.copy() for k, v in st.items()}
    oracle.update(sub, -0.41, 0.37, 1e-3, ref["rhsQ"], ref["resQ"], ref["Q"], ref["Qf"])
    for k, n in (("Q", nq), ("resQ", nq), ("rhsQ", nq), ("Qf", nfc)):
        assert np.array_equal(got[k][Ksub * n:], st[k][Ksub * n:]), k
        if Ksub:
            assert relerr(got[k][:Ksub * n], ref[k][:Ksub * n]) < 1e-12, k
    # fused stage on the first Ksub elements
    st = {"Q": Q.copy(), "resQ": 1e-3 * np.ones_like(Q), "rhsQ": rq.copy(), "Qf": qf.copy()}
    dev.set_state(**st)
    out = torch.full_like(dev.s["Qf"], 7.0)
    s = dev.s
    dev.plan.stage(Ksub, *g, RK4A[2], RK4B[2], prob.dt, _ptr(s["Qf"]), _ptr(out), _ptr(s["rhsQ"]),
                   _ptr(s["resQ"]), _ptr(s["Q"]), dev.stream())
    got = dev.get_state()
    got["Qf"] = out.cpu().numpy()
    ref = {k: v.copy() for k, v in st.items()}
    oracle.stage(sub, ref, RK4A[2], RK4B[2], prob.dt)
    for k in ("Q", "resQ", "rhsQ"):
        assert np.array_equal(got[k][Ksub * nq:], st[k][Ksub * nq:]), k
        if Ksub:
            assert relerr(got[k][:Ksub * nq], ref[k][:Ksub * nq]) < 1e-12, k
    assert np.all(got["Qf"][Ksub * nfc:] == 7.0)
    if Ksub:
        assert relerr(got["Qf"][:Ksub * nfc], ref["Qf"][:Ksub * nfc]) < 1e-12


def test_ensemble_with_empty_problem():
    from scratch_wadge_amd import _lib
    from scratch_wadge_amd import problem as P
    from scratch_wadge_amd.device import DeviceSolver
    from scratch_wadge_amd.ensemble import PerProblem, group_of
    print("case ensemble with an empty problem", flush=True)
    probs = [P.vortex2d_problem(n=4, N=3, beta=1.0), P.vortex2d_problem(n=3, N=2, beta=1.2)]
    solvers = [DeviceSolver(p) for p in probs]
    empty = _lib.Group()
    empty.N, empty.Nq, empty.Nfq, empty.K, empty.dt = 4, 16, 5, 0, 1e-3
    groups = [group_of(solvers[0]), empty, group_of(solvers[1])]
    b = _lib.Batch(2, groups, probs[0].gamma, probs[0].tau)
    b.run(3, use_graph=True)
    ref = PerProblem(probs)
    ref.run(3)
    torch.cuda.synchronize()
    cur = b.current()
    for sv, q in zip(solvers, range(2)):
        r = ref.state(q)
        qf = (sv.s["Qf"], sv.Qf_alt)[cur].cpu().numpy()
        assert np.array_equal(sv.s["Q"].cpu().numpy(), r["Q"])
        assert np.array_equal(qf, r["Qf"])
    b.close()
