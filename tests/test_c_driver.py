"""The C ABI used from plain C (tests/c_driver/rk_driver.c, the reference's
rk_run written against include/wadges.h): it links, and on the GPU its LSRK45
steps, as three kernels per stage or one fused launch, equal the Python
driver's bit for bit."""

import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "c_driver", "rk_driver")


def test_driver_built_and_checks_arguments():
    assert os.path.exists(DRIVER), "run __graft_entry__.build() first"
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr


def _write(prob, d):
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{prob.dim} {prob.N} {prob.Nq} {prob.Nfq} {prob.K} {prob.dt!r} {prob.gamma!r} {prob.tau!r}\n")
    names = ["vgeo", "vfgeo", "fgeo", "Jq", "Drq", "Dsq", "VqLq", "VqPq", "VfPq", "nrJ", "nsJ",
             "Q", "rhsQ", "resQ", "Qf"] + (["Dtq", "ntJ"] if prob.dim == 3 else [])
    for n in names:
        np.ascontiguousarray(getattr(prob, n), dtype=np.float64).tofile(os.path.join(d, n + ".bin"))
    np.ascontiguousarray(prob.mapPq, dtype=np.uint32).tofile(os.path.join(d, "mapPq.bin"))


@pytest.mark.gpu
@pytest.mark.parametrize("dim,N,mode", [(2, 3, 0), (2, 3, 1), (3, 3, 0), (3, 3, 1)])
def test_c_driver_matches_python_driver(tmp_path, dim, N, mode):
    from helpers import small_problem
    from scratch_wadge_amd.device import DeviceSolver
    prob = small_problem(dim, N)
    print(f"case c driver dim={dim} N={N} mode={mode}", flush=True)
    _write(prob, str(tmp_path))
    r = subprocess.run([DRIVER, str(tmp_path), "3", str(mode)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    dev = DeviceSolver(prob)
    dev.run(3, fused=bool(mode))
    g = dev.get_state()
    for k in ("Q", "rhsQ", "resQ", "Qf"):
        c = np.fromfile(os.path.join(str(tmp_path), f"out_{k}.bin"), dtype=np.float64)
        assert np.array_equal(c, g[k]), k
