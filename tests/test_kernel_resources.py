"""Register budget of the forward kernels, from the compiler's resource remarks
(-Rpass-analysis=kernel-resource-usage) on a device-only compile with the
library's own flags.  The volume and fused stage kernels of every shape must not
spill to scratch (a spill there costs more than any register-hungry schedule
gains); the one-tile ensemble kernels may spill at most 32 bytes per lane (a
fourth wave per SIMD was once worth a 20-byte spill in the 2D one); the 3D N = 3
fused stage must keep the two waves per SIMD that its LDS allows."""

import os
import re
import shutil
import subprocess

import pytest

from scratch_wadge_amd import build as B

NO_SCRATCH = ("vol_kernel", "stage_kernel")
SMALL_SCRATCH = ("batch_stage_flat",)


@pytest.fixture(scope="module")
def remarks(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    flags = [f for f in B.HIPCC_FLAGS if f not in ("-shared",)]
    out = tmp_path_factory.mktemp("rr") / "w.o"
    res = subprocess.run([hipcc, *flags, "--cuda-device-only", "-c", os.path.join(B.CSRC, "wadges.hip"),
                          "-o", str(out), "-Rpass-analysis=kernel-resource-usage"],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    kern, info = None, {}
    for line in res.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            kern = m.group(1)
            info[kern] = {}
            continue
        m = re.search(r"remark:\s+(VGPRs|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]): (\d+)", line)
        if m and kern:
            info[kern][m.group(1).split()[0]] = int(m.group(2))
    assert info, "no resource remarks parsed"
    return info


def _shape(name):
    m = re.search(r"ShapeILi(\d)ELi(\d+)ELi(\d+)", name)
    return tuple(int(x) for x in m.groups()) if m else None


def test_forward_kernels_do_not_spill(remarks):
    seen = flat = 0
    for name, r in remarks.items():
        if any(k in name for k in NO_SCRATCH):
            seen += 1
            assert r.get("ScratchSize", 0) == 0, (name[:90], r)
        if any(k in name for k in SMALL_SCRATCH):
            flat += 1
            assert r.get("ScratchSize", 0) <= 32, (name[:90], r)
    assert seen >= 2 * 9 and flat == 6   # vol + stage for 9 shapes; flat ensemble: 2 dims x 3 widths


def test_3d_n3_stage_keeps_two_waves(remarks):
    st = [r for n, r in remarks.items() if "stage_kernel" in n and _shape(n) == (3, 23, 12)]
    assert len(st) == 1
    assert st[0]["Occupancy"] >= 2, st[0]
