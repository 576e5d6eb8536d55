"""Generated code of the volume kernels where a lane group's LDS additions must
stay separate instructions: the 3D N = 4 face-point walk (lanes L and L + 44 add
into the same row at every step) and the 2D N = 2 face walk (both lane groups on
the same face point).  One merged instruction would leave the order of the two
additions into an address to the hardware; the kernels give the second group's
atomic another scope so that the compiler keeps the groups apart.  Checked on the
gfx950 assembly of a device-only compile with the library's own flags: a masked
block of ds_add_f64 under the complementary lane mask (s_andn2_saveexec) must
follow the first group's."""

import os
import re
import shutil
import subprocess

import pytest

from scratch_wadge_amd import build as B


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    flags = [f for f in B.HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
    out = tmp_path_factory.mktemp("asm") / "w.s"
    res = subprocess.run([hipcc, *flags, "--cuda-device-only", "-S",
                          os.path.join(B.CSRC, "wadges.hip"), "-o", str(out)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    return out.read_text()


def _function(asm, mangled_prefix):
    start = asm.index("\n" + mangled_prefix) + 1
    end = asm.index("s_endpgm", start)
    return asm[start:end]


SPLIT = re.compile(r"s_andn2_saveexec_b64[^\n]*\n(?:\s*s_cbranch_execz[^\n]*\n)?(?:;[^\n]*\n)?\s*ds_add_f64")


@pytest.mark.parametrize("kernel,shape", [("vol_kernel", (3, 44, 16)), ("stage_kernel", (3, 44, 16)),
                                          ("vol_kernel", (2, 6, 3)), ("stage_kernel", (2, 6, 3))])
def test_second_lane_group_adds_separately(asm, kernel, shape):
    d, nq, nfq = shape
    name = f"_ZN6wadges{len(kernel)}{kernel}INS_5ShapeILi{d}ELi{nq}ELi{nfq}EEE"
    body = _function(asm, name)
    assert "ds_add_f64" in body
    assert SPLIT.search(body), f"{kernel}{shape}: no masked second group of LDS additions"
