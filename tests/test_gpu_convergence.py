"""End-to-end accuracy against the reference's published results: the
isentropic-vortex convergence study of its paper (docs/main.tex:1622-1675,
:1766-1845), whose L2 errors are tabulated in docs/figs/errors_3d.m, on the
reference's own gmsh meshes (meshes/periodicSquare1-4.msh, [0,20] x [-5,5],
h = 2, 1, 1/2, 1/4; meshes/periodicCube1-3.msh, [0,10] x [0,20] x [0,10],
1354 / 9543 / 72923 tets, h = 2, 1, 1/2; copied under tests/golden).

Each case runs the whole path -- mesh ingest and periodic matching, the
operator and geometry builder (the reference's quadrature tables and
interpolation nodes, Kopriva geofacs in 3D), the projected initial state, and
the fused stage on the GPU -- to T = 5 with the reference's rk_run step
(ceil(T/dt) steps, euler.c:3678-3683; dt = CFL hmin / CN, euler.c:4180-4192)
and measures the reference's error, sqrt(sum wq J |Q - Qexact|^2) over all
fields (euler.c:4227-4274), with the device reduction.  The published numbers
come from the reference's runs at a CFL the paper does not state for every
case; the step changes the errors in about the sixth digit (checked against the
oracle: 2D N = 3 on periodicSquare1, 0.5078212 at CFL 0.5 and 0.5078202 at
0.25, against 0.5078199 published).  The bar is a relative difference of
2e-4 from every published error, far below the factor 8-30 between
refinement levels, so a wrong rule, node set, geofac, flux or update fails it.
"""

import os

import numpy as np
import pytest

from scratch_wadge_amd import problem as P

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# docs/figs/errors_3d.m (2D: :3-10; 3D: :23-34).  3D curved: the runs with the
# Kopriva geofacs (errck, :25 and :29) where tabulated, else errc (:33).
PUBLISHED = {
    (2, False, 2): [1.067171, 0.1496387, 0.02569296, 0.003428271],
    (2, True, 2): [1.377095, 0.2195012, 0.03858958, 0.003990601],
    (2, False, 3): [0.5078199, 0.05160533, 0.002494255, 0.0001106185],
    (2, True, 3): [0.5350362, 0.06574176, 0.005015359, 0.0002430048],
    (2, False, 4): [0.2353737, 0.01257141, 0.0003215585, 1.06097e-05],
    (2, True, 4): [0.297758, 0.02076038, 0.000816006, 2.361021e-05],
    (3, False, 2): [1.055192, 0.1435153, 0.02126818],
    (3, True, 2): [1.209334, 0.2069238, 0.02845067],
    (3, False, 3): [0.3393179, 0.03143425, 0.001976993],
    (3, True, 3): [0.4646378, 0.05133794, 0.003346182],
    (3, False, 4): [0.1222291, 0.004884339, 0.0001924525],
    (3, True, 4): [0.1845467, 0.01043607, 0.0003960681],
}
MESHES = {2: [f"periodicSquare{i}.msh" for i in (1, 2, 3, 4)],
          3: [f"periodicCube{i}.msh" for i in (1, 2, 3)]}
BAR = 2e-4


def vortex_error(path, N, curved, cfl):
    from scratch_wadge_amd.device import DeviceSolver
    prob = P.vortex_reference_problem(path, N, curved=curved, cfl=cfl)
    nsteps, dt = P.rk_steps(prob.dt, 5.0)
    dev = DeviceSolver(prob)
    dev.run(nsteps, dt=dt, fused=True, graph=True)
    err = dev.diagnostics(P.vortex_reference_exact(prob, 5.0))["l2_error"]
    ok = bool(np.isfinite(dev.s["Q"].cpu().numpy()).all())
    del dev
    return prob.K, nsteps, err, ok


@pytest.mark.parametrize("dim,curved,N", sorted(PUBLISHED), ids=lambda v: str(v))
def test_vortex_errors_match_published(dim, curved, N):
    import torch
    # the tabulated 2D curved N = 4 run used CFL 1/4 (errors_3d.m:10); the others 1/2
    cfl = 0.25 if (dim, curved, N) == (2, True, 4) else 0.5
    pub = PUBLISHED[(dim, curved, N)]
    errs = []
    for mesh, ref in zip(MESHES[dim], pub):
        print(f"case vortex dim={dim} curved={curved} N={N} {mesh} CFL {cfl}", flush=True)
        K, nsteps, err, ok = vortex_error(os.path.join(GOLD, mesh), N, curved, cfl)
        torch.cuda.empty_cache()
        assert ok
        errs.append(err)
        print(f"  K={K} steps={nsteps}: L2 error {err:.7g}, published {ref:.7g}, "
              f"relative difference {abs(err - ref) / ref:.2e}", flush=True)
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    prates = np.log2(np.array(pub[:-1]) / np.array(pub[1:]))
    print(f"  observed rates {np.round(rates, 3)}, published {np.round(prates, 3)}", flush=True)
    for err, ref in zip(errs, pub):
        assert abs(err - ref) <= BAR * ref


def test_published_errors_pin_the_penalty_constant():
    """The published errors are sensitive to the Lax-Friedrichs penalty
    (okl/Euler3D.okl:3316-3330, tau/4 max(lamM, lamP) sJ): with the constant
    doubled (tau = 2) or the penalty left out (tau = 0), the error of 2D N = 3 on
    periodicSquare2.msh misses the published one by far more than the bar the
    reference's discretization meets, so a wrong penalty constant would fail
    test_vortex_errors_match_published."""
    import torch
    path = os.path.join(GOLD, "periodicSquare2.msh")
    pub = PUBLISHED[(2, False, 3)][1]

    def err(tau):
        from scratch_wadge_amd.device import DeviceSolver
        prob = P.vortex_reference_problem(path, 3, curved=False, cfl=0.5, tau=tau)
        nsteps, dt = P.rk_steps(prob.dt, 5.0)
        dev = DeviceSolver(prob)
        dev.run(nsteps, dt=dt, fused=True, graph=True)
        e = dev.diagnostics(P.vortex_reference_exact(prob, 5.0))["l2_error"]
        del dev
        torch.cuda.empty_cache()
        return e

    rel = lambda e: abs(e - pub) / pub
    e1 = err(1.0)
    print(f"case penalty pin: tau=1 L2 error {e1:.7g} (published {pub:.7g})", flush=True)
    assert rel(e1) <= BAR
    for tau in (0.0, 2.0):
        e = err(tau)
        print(f"  tau={tau}: L2 error {e:.7g}, relative difference {rel(e):.2e}", flush=True)
        assert rel(e) > 20 * BAR
