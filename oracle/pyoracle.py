"""ctypes loader for the CPU oracle (test infrastructure only).

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker (or the reported CPU baseline) -- never by the product.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwadges_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_up = ctypes.POINTER(ctypes.c_uint32)


class WoParams(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("Nq", ctypes.c_int), ("Nfq", ctypes.c_int),
                ("gamma", ctypes.c_double), ("tau", ctypes.c_double)]


def build(force=False):
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "all"], cwd=HERE)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(WoParams)
        L.wo_logmean.restype = ctypes.c_double
        L.wo_logmean.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.wo_ec_flux.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        L.wo_VU.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.wo_UV.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.wo_vol.argtypes = [P, ctypes.c_int] + [_dp] * 13
        L.wo_surf.argtypes = [P, ctypes.c_int, _dp, _up, _dp, _dp, _dp, _dp]
        L.wo_update.argtypes = [P, ctypes.c_int, _dp, _dp, _dp, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp]
        L.wo_flop_counts.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.wo_set_threads.argtypes = [ctypes.c_int]
        L.wo_set_logmean_variant.argtypes = [ctypes.c_int]
        L.wo_set_update_variant.argtypes = [ctypes.c_int]
        L.wo_get_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def params(prob):
    return WoParams(prob.dim, prob.Nq, prob.Nfq, prob.gamma, prob.tau)


def set_threads(n):
    lib().wo_set_threads(int(n))


def set_logmean_variant(v):
    """0 = intended log mean; 1 = the reference's as-shipped 3D log branch (tests only)."""
    lib().wo_set_logmean_variant(int(v))


def set_update_variant(v):
    """0 = the weight-adjusted update as written; 1 = its first product without the
    projection, a slip invisible on affine meshes (tests only)."""
    lib().wo_set_update_variant(int(v))


def logmean(dim, a, b):
    return lib().wo_logmean(dim, float(a), float(b))


def ec_flux(dim, gamma, UL, UR):
    UL = np.ascontiguousarray(UL, dtype=np.float64)
    UR = np.ascontiguousarray(UR, dtype=np.float64)
    F = np.zeros((dim, dim + 2))
    lib().wo_ec_flux(dim, gamma, _p(UL), _p(UR), _p(F))
    return F


def VU(dim, gamma, U):
    U = np.ascontiguousarray(U, dtype=np.float64)
    V = np.zeros(dim + 2)
    lib().wo_VU(dim, gamma, _p(U), _p(V))
    return V


def UV(dim, gamma, V):
    V = np.ascontiguousarray(V, dtype=np.float64)
    U = np.zeros(dim + 2)
    lib().wo_UV(dim, gamma, _p(V), _p(U))
    return U


def vol(prob, rhsQ, Qf, rhsQf):
    """In place: rhsQ (projected state in, volume RHS out), rhsQf out."""
    p = params(prob)
    lib().wo_vol(ctypes.byref(p), prob.K, _p(prob.vgeo), _p(prob.vfgeo), _p(prob.nrJ),
                 _p(prob.nsJ), _p(prob.ntJ), _p(prob.Drq), _p(prob.Dsq), _p(prob.Dtq),
                 _p(prob.VqLq), _p(prob.VfPq), _p(Qf), _p(rhsQ), _p(rhsQf))


def surf(prob, Qf, rhsQf, rhsQ):
    p = params(prob)
    m = np.ascontiguousarray(prob.mapPq, dtype=np.uint32)
    lib().wo_surf(ctypes.byref(p), prob.K, _p(prob.fgeo), m.ctypes.data_as(_up),
                  _p(prob.VqLq), _p(Qf), _p(rhsQf), _p(rhsQ))


def update(prob, fa, fb, dt, rhsQ, resQ, Q, Qf):
    p = params(prob)
    lib().wo_update(ctypes.byref(p), prob.K, _p(prob.Jq), _p(prob.VqPq), _p(prob.VfPq),
                    fa, fb, dt, _p(rhsQ), _p(resQ), _p(Q), _p(Qf))


def rhs(prob, rhsQ_in, Qf):
    """Volume + surface RHS of a projected state; returns (rhsQ, rhsQf)."""
    r = rhsQ_in.copy()
    rf = np.zeros_like(Qf)
    vol(prob, r, Qf, rf)
    surf(prob, Qf, rf, r)
    return r, rf


def stage(prob, st, fa, fb, dt):
    """One LSRK stage (vol, surf, update) on a state dict with Q, rhsQ, Qf, resQ."""
    rf = np.zeros_like(st["Qf"])
    vol(prob, st["rhsQ"], st["Qf"], rf)
    surf(prob, st["Qf"], rf, st["rhsQ"])
    update(prob, fa, fb, dt, st["rhsQ"], st["resQ"], st["Q"], st["Qf"])


def flop_counts(dim, Nq, Nfq):
    out = np.zeros(13)
    lib().wo_flop_counts(dim, Nq, Nfq, _p(out))
    keys = ["c_F", "c_node", "c_vv", "c_vf", "vol_elem", "surf_elem", "update_elem",
            "F_log", "F_exp", "F_pow", "F_sqrt", "F_div", "vf_pair_deviation"]
    return dict(zip(keys, out.tolist()))
