"""The C-ABI library loads and exports what include/wadges.h declares (no GPU calls)."""

import ctypes
import os
import re

import pytest

from scratch_wadge_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "wadges.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wadges_[a-z_]+)\s*\(", src)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(_lib.EXPORTED)


def test_library_exports_every_symbol():
    L = _lib.load()
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_instances_and_error_strings():
    inst = _lib.instances()
    assert (3, 3, 23, 12) in inst and (3, 4, 44, 16) in inst and (2, 3, 12, 4) in inst
    L = _lib.load()
    for code in (0, -1, -2, -3, -4, -5):
        assert L.wadges_error_string(code) != b"unknown error"
    # argument checks run before any HIP call
    rc = L.wadges_vol(None, 10, *([None] * 15), None)
    assert rc == -1
    assert L.wadges_diagnostics_noalloc(None, 10, *([None] * 5), None) == -1
    h = ctypes.c_void_p()
    assert L.wadges_plan_create(ctypes.byref(h), 3, 3, 24, 12, 1.4, 1.0) == -2  # no such instance
    assert L.wadges_plan_create(ctypes.byref(h), 3, 3, 23, 12, 0.9, 1.0) == -1  # gamma <= 1


def test_workmodel_matches_oracle_flop_counts(oracle):
    from scratch_wadge_amd import workmodel as W
    for dim, Nq, Nfq in ((3, 23, 12), (2, 12, 4)):
        c = oracle.flop_counts(dim, Nq, Nfq)
        for k in ("c_F", "c_node", "c_vv", "c_vf"):
            assert W.PIECES[dim][k] == c[k], (dim, k)
        # the counted pair form computes what the two per-block contractions do
        assert c["vf_pair_deviation"] <= 1e-14
    # one volume-face pair: average 2 NG, normal DIM (2 DIM - 1), F.n NFLD (2 DIM - 1),
    # and 2 flops per field for each of the two rows
    for dim in (2, 3):
        ng, nf = dim * dim, dim + 2
        assert W.PIECES[dim]["c_vf"] == 2 * ng + (dim + nf) * (2 * dim - 1) + 4 * nf
    for (dim, Nq, Nf), (sf, uf) in W.SURF_UPDATE.items():
        c = oracle.flop_counts(dim, Nq, Nf // (dim + 1))
        assert (sf, uf) == (c["surf_elem"], c["update_elem"]), (dim, Nq)
    assert W.vol_flops(3, 23, 48) == 224004
    assert W.bytes_vol(3, 23, 48) == 10792
    assert W.bytes_surf(3, 23, 48) == 9328
    assert W.bytes_update(3, 23, 48) == 7624


def test_batch_argument_checks():
    """wadges_batch_create rejects bad descriptors before any HIP call."""
    L = _lib.load()
    h = ctypes.c_void_p()
    fake = 0x1000      # never dereferenced: validation fails first

    def group(Nq=12, Nfq=4, K=8, same_qf=False, dt=1e-3):
        g = _lib.Group()
        g.N, g.Nq, g.Nfq, g.K, g.dt = 3, Nq, Nfq, K, dt
        for name in ("vgeo", "vfgeo", "fgeo", "mapPq", "nrJ", "nsJ", "Drq", "Dsq", "VqLq", "VqPq",
                     "VfPq", "Jq", "Q", "rhsQ", "resQ"):
            setattr(g, name, fake)
        g.Qf[0] = fake
        g.Qf[1] = fake if same_qf else fake + 8
        return g

    def create(dim, groups, gamma=1.4, tau=1.0):
        arr = (_lib.Group * max(1, len(groups)))(*groups)
        return L.wadges_batch_create(ctypes.byref(h), dim, len(groups), arr, gamma, tau, None)

    assert create(4, [group()]) == -1                 # no such dimension
    assert create(2, [group(Nq=13)]) == -2            # no kernel instance
    assert create(2, [group(K=-1)]) == -1
    assert create(2, [group(same_qf=True)]) == -1     # the stage reads one buffer, writes the other
    assert create(2, [group(dt=float("nan"))]) == -1
    assert create(3, [group(Nq=23, Nfq=12)]) == -1    # 3D needs ntJ and Dtq
    assert create(2, [group()], gamma=1.0) == -1
    assert L.wadges_batch_create(ctypes.byref(h), 2, -1, None, 1.4, 1.0, None) == -1
    assert L.wadges_batch_current(None) == -1
    assert L.wadges_batch_stage(None, 0.0, 0.0, None) == -1
    assert L.wadges_batch_run(None, 1, 1, None) == -1
    assert L.wadges_batch_tune(None, 1, None, None) == -1
    assert L.wadges_plan_invalidate(None) == -1
    assert L.wadges_batch_destroy(None) == 0
