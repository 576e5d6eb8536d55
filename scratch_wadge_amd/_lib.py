"""ctypes binding of the C ABI declared in include/wadges.h.

The product path: every kernel call goes through libwadges.so (HIP, gfx950).
If the library is missing, loading fails with an error -- there is no CPU
fallback.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwadges.so")

# Every symbol the header declares (checked by the CPU tests).
EXPORTED = (
    "wadges_plan_create", "wadges_plan_destroy", "wadges_plan_invalidate", "wadges_vol", "wadges_surf", "wadges_update",
    "wadges_stage", "wadges_stage_record", "wadges_stage_vjp", "wadges_diagnostics",
    "wadges_diagnostics_noalloc",
    "wadges_rhs_vjp", "wadges_update_vjp",
    "wadges_project_vjp", "wadges_batch_create", "wadges_batch_destroy", "wadges_batch_stage",
    "wadges_batch_run", "wadges_batch_current", "wadges_batch_get_config", "wadges_batch_set_config",
    "wadges_batch_tune",
    "wadges_last_hip_error", "wadges_error_string", "wadges_instance_count", "wadges_instance",
)

_vp = ctypes.c_void_p
_lib = None


class WadgesError(RuntimeError):
    pass


class Group(ctypes.Structure):
    """struct wadges_group: one problem of an ensemble (device pointers)."""
    _fields_ = [("N", ctypes.c_int), ("Nq", ctypes.c_int), ("Nfq", ctypes.c_int), ("K", ctypes.c_int),
                ("dt", ctypes.c_double)] + [
        (name, _vp) for name in ("vgeo", "vfgeo", "fgeo", "mapPq", "nrJ", "nsJ", "ntJ", "Drq", "Dsq",
                                 "Dtq", "VqLq", "VqPq", "VfPq", "Jq", "Q", "rhsQ", "resQ")] + [
        ("Qf", _vp * 2)]


class BatchConfig(ctypes.Structure):
    """struct wadges_batch_config: launch shape of the ensemble kernel."""
    _fields_ = [("tiles_per_item", ctypes.c_int), ("grid", ctypes.c_int), ("dynamic", ctypes.c_int),
                ("order", ctypes.c_int), ("waves_per_group", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise WadgesError(f"{LIB_PATH} is missing: build it with `python -m scratch_wadge_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    L.wadges_plan_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.wadges_plan_destroy.argtypes = [_vp]
    L.wadges_plan_invalidate.argtypes = [_vp]
    L.wadges_vol.argtypes = [_vp, ctypes.c_int] + [_vp] * 15 + [_vp]
    L.wadges_surf.argtypes = [_vp, ctypes.c_int] + [_vp] * 6 + [_vp]
    L.wadges_update.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, _vp, _vp, _vp, _vp, _vp]
    L.wadges_stage.argtypes = [_vp, ctypes.c_int] + [_vp] * 14 + [ctypes.c_double] * 3 + [_vp] * 5 + [_vp]
    L.wadges_stage_record.argtypes = ([_vp, ctypes.c_int] + [_vp] * 14 + [ctypes.c_double] * 3 +
                                      [_vp] * 8 + [_vp])
    L.wadges_stage_vjp.argtypes = ([_vp, ctypes.c_int] + [_vp] * 14 + [ctypes.c_double] * 3 +
                                   [_vp] * 11 + [_vp])
    L.wadges_diagnostics.argtypes = [_vp, ctypes.c_int] + [_vp] * 6 + [_vp]
    L.wadges_diagnostics_noalloc.argtypes = [_vp, ctypes.c_int] + [_vp] * 5 + [_vp]
    L.wadges_rhs_vjp.argtypes = [_vp, ctypes.c_int] + [_vp] * 18 + [_vp]
    L.wadges_update_vjp.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double] + [_vp] * 6 + [_vp]
    L.wadges_project_vjp.argtypes = [_vp, ctypes.c_int] + [_vp] * 6 + [_vp]
    L.wadges_batch_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(Group), ctypes.c_double, ctypes.c_double, _vp]
    L.wadges_batch_destroy.argtypes = [_vp]
    L.wadges_batch_stage.argtypes = [_vp, ctypes.c_double, ctypes.c_double, _vp]
    L.wadges_batch_run.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp]
    L.wadges_batch_current.argtypes = [_vp]
    L.wadges_batch_get_config.argtypes = [_vp, ctypes.POINTER(BatchConfig)]
    L.wadges_batch_set_config.argtypes = [_vp, ctypes.POINTER(BatchConfig)]
    L.wadges_batch_tune.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), _vp]
    L.wadges_error_string.restype = ctypes.c_char_p
    L.wadges_error_string.argtypes = [ctypes.c_int]
    L.wadges_instance.argtypes = [ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 4
    _lib = L
    return L


def check(rc, what):
    if rc != 0:
        L = load()
        msg = L.wadges_error_string(rc).decode()
        raise WadgesError(f"{what} failed: {msg} (code {rc}, hip error {L.wadges_last_hip_error()})")


def instances():
    L = load()
    out = []
    for n in range(L.wadges_instance_count()):
        v = [ctypes.c_int() for _ in range(4)]
        check(L.wadges_instance(n, *[ctypes.byref(x) for x in v]), "wadges_instance")
        out.append(tuple(x.value for x in v))
    return out


class Plan:
    """Owns a wadges_plan (the reference's kernel defines as a runtime object)."""

    def __init__(self, dim, N, Nq, Nfq, gamma, tau):
        L = load()
        h = _vp()
        check(L.wadges_plan_create(ctypes.byref(h), dim, N, Nq, Nfq, gamma, tau), "wadges_plan_create")
        self.h = h
        self.dim, self.N, self.Nq, self.Nfq = dim, N, Nq, Nfq

    def close(self):
        if self.h:
            load().wadges_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def invalidate(self):
        check(load().wadges_plan_invalidate(self.h), "wadges_plan_invalidate")

    def vol(self, K, vgeo, vfgeo, nrJ, nsJ, ntJ, Drq, Dsq, Dtq, Drstq, VqLq, VfPq, Q, Qf, rhsQ,
            rhsQf, stream=None):
        check(load().wadges_vol(self.h, K, vgeo, vfgeo, nrJ, nsJ, ntJ, Drq, Dsq, Dtq, Drstq, VqLq,
                                VfPq, Q, Qf, rhsQ, rhsQf, stream), "wadges_vol")

    def surf(self, K, fgeo, mapPq, VqLq, Qf, rhsQf, rhsQ, stream=None):
        check(load().wadges_surf(self.h, K, fgeo, mapPq, VqLq, Qf, rhsQf, rhsQ, stream), "wadges_surf")

    def stage(self, K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, Drq, Dsq, Dtq, VqLq, VqPq, VfPq, Jq,
              fa, fb, fdt, Qf_in, Qf_out, rhsQ, resQ, Q, stream=None):
        check(load().wadges_stage(self.h, K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, Drq, Dsq, Dtq,
                                  VqLq, VqPq, VfPq, Jq, fa, fb, fdt, Qf_in, Qf_out, rhsQ, resQ, Q,
                                  stream), "wadges_stage")

    def stage_record(self, K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, Drq, Dsq, Dtq, VqLq, VqPq,
                     VfPq, Jq, fa, fb, fdt, Qf_in, Qf_out, rhsQ_in, rhsQ_out, resQ_in, resQ_out,
                     Q_in, Q_out, stream=None):
        check(load().wadges_stage_record(self.h, K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, Drq,
                                         Dsq, Dtq, VqLq, VqPq, VfPq, Jq, fa, fb, fdt, Qf_in, Qf_out,
                                         rhsQ_in, rhsQ_out, resQ_in, resQ_out, Q_in, Q_out, stream),
              "wadges_stage_record")

    def stage_vjp(self, K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, Drq, Dsq, Dtq, VqLq, VqPq, VfPq,
                  Jq, fa, fb, fdt, rhsQ_in, Qf_in, Q_out, Q_bar, resQ_bar, rhsQ_bar_out, Qf_bar_out,
                  R_bar, sbar, rhsQ_bar_in, Qf_bar_in, stream=None):
        check(load().wadges_stage_vjp(self.h, K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, Drq, Dsq,
                                      Dtq, VqLq, VqPq, VfPq, Jq, fa, fb, fdt, rhsQ_in, Qf_in, Q_out,
                                      Q_bar, resQ_bar, rhsQ_bar_out, Qf_bar_out, R_bar, sbar,
                                      rhsQ_bar_in, Qf_bar_in, stream), "wadges_stage_vjp")

    def diagnostics(self, K, wq, Jq, Q, Qexact, aux, totals, stream=None):
        check(load().wadges_diagnostics(self.h, K, wq, Jq, Q, Qexact, aux, totals, stream),
              "wadges_diagnostics")

    def rhs_vjp(self, K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, Drq, Dsq, Dtq, VqLq, VfPq, rhsQ, Qf,
                rhsQ_bar, sbar, uq_bar, uf_bar, stream=None):
        check(load().wadges_rhs_vjp(self.h, K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, Drq, Dsq, Dtq,
                                    VqLq, VfPq, rhsQ, Qf, rhsQ_bar, sbar, uq_bar, uf_bar, stream),
              "wadges_rhs_vjp")

    def update_vjp(self, K, Jq, VqPq, VfPq, fa, fb, fdt, Q, Q_bar, res_bar, uq_bar_out, uf_bar_out,
                   rhsQ_bar, stream=None):
        check(load().wadges_update_vjp(self.h, K, Jq, VqPq, VfPq, fa, fb, fdt, Q, Q_bar, res_bar,
                                       uq_bar_out, uf_bar_out, rhsQ_bar, stream), "wadges_update_vjp")

    def project_vjp(self, K, VqPq, VfPq, Q, uq_bar, uf_bar, Q_bar, stream=None):
        check(load().wadges_project_vjp(self.h, K, VqPq, VfPq, Q, uq_bar, uf_bar, Q_bar, stream),
              "wadges_project_vjp")

    def update(self, K, Jq, VqPq, VfPq, fa, fb, fdt, rhsQ, resQ, Q, Qf, stream=None):
        check(load().wadges_update(self.h, K, Jq, VqPq, VfPq, fa, fb, fdt, rhsQ, resQ, Q, Qf, stream),
              "wadges_update")


class Batch:
    """Owns a wadges_batch: an ensemble of problems advanced by one launch per
    stage.  `groups` is a sequence of Group structures; the caller keeps the
    device arrays they point to alive for the life of the batch."""

    def __init__(self, dim, groups, gamma, tau, stream=None):
        L = load()
        arr = (Group * len(groups))(*groups)
        h = _vp()
        check(L.wadges_batch_create(ctypes.byref(h), dim, len(groups), arr, gamma, tau, stream),
              "wadges_batch_create")
        self.h = h

    def close(self):
        if self.h:
            load().wadges_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stage(self, fa, fb, stream=None):
        check(load().wadges_batch_stage(self.h, fa, fb, stream), "wadges_batch_stage")

    def run(self, nsteps, use_graph=True, stream=None):
        check(load().wadges_batch_run(self.h, nsteps, 1 if use_graph else 0, stream), "wadges_batch_run")

    def current(self):
        return load().wadges_batch_current(self.h)

    def config(self):
        c = BatchConfig()
        check(load().wadges_batch_get_config(self.h, ctypes.byref(c)), "wadges_batch_get_config")
        return c.as_dict()

    def set_config(self, tiles_per_item=1, grid=0, dynamic=0, order=0, waves_per_group=1):
        c = BatchConfig(tiles_per_item, grid, dynamic, order, waves_per_group)
        check(load().wadges_batch_set_config(self.h, ctypes.byref(c)), "wadges_batch_set_config")

    def tune(self, reps=5, stream=None):
        ms = ctypes.c_double()
        check(load().wadges_batch_tune(self.h, reps, ctypes.byref(ms), stream), "wadges_batch_tune")
        return ms.value
