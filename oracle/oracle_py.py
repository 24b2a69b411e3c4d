"""ctypes loader of the FP64 CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the checker;
the engine never imports it.  Mesh and parameter structs are the engine's public C types
(include/rkdg.h), so both sides are fed byte-identical inputs.
"""
import ctypes as C
from pathlib import Path

import numpy as np

from dgswemv2_amd import _native as N

LIB = Path(__file__).resolve().parent / "liboracle.so"
_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise OSError(f"{LIB} missing: run make -C oracle")
        lib = C.CDLL(str(LIB))
        dp, ip, up = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_uint8)
        sig = {
            "orc_create": (C.c_void_p, [C.POINTER(N.MeshDesc), C.POINTER(N.Params)]),
            "orc_error": (C.c_char_p, []),
            "orc_destroy": (None, [C.c_void_p]),
            "orc_set_threads": (None, [C.c_void_p, C.c_int]),
            "orc_keep_prelimit": (None, [C.c_void_p, C.c_int]),
            "orc_round_state": (None, [C.c_void_p, C.c_int]),
            "orc_get_prelimit": (C.c_int, [C.c_void_p, dp]),
            "orc_limiter_geometry": (None, [C.c_void_p, ip, dp]),
            "orc_max_threads": (C.c_int, []),
            "orc_stage": (C.c_int, [C.c_void_p]),
            "orc_step": (C.c_int, [C.c_void_p, C.c_int]),
            "orc_nan_info": (None, [C.c_void_p, ip, ip, ip]),
            "orc_get_state": (None, [C.c_void_p, dp, up, up, up]),
            "orc_set_state": (None, [C.c_void_p, dp, up]),
            "orc_get_time": (None, [C.c_void_p, dp, ip, ip]),
            "orc_l2_residual": (C.c_double, [C.c_void_p]),
            "orc_num_interfaces": (C.c_int, [C.c_void_p]),
            "orc_num_boundaries": (C.c_int, [C.c_void_p]),
            "orc_dubiner": (C.c_int, [C.c_int, C.c_int, dp, dp, dp, dp]),
            "orc_dunavant": (C.c_int, [C.c_int, dp, dp]),
            "orc_gauss_legendre": (C.c_int, [C.c_int, dp, dp]),
            "orc_llf_flux": (None, [C.c_double, dp, dp, dp, dp, dp]),
            "orc_rk_coeffs": (C.c_int, [C.c_int, C.c_int, dp, dp, dp]),
            "orc_stage_times": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_int, dp]),
            "orc_register_scheme": (C.c_int, [C.c_int, C.c_int, dp, dp]),
            "orc_element_integrals": (C.c_int, [C.c_int, dp, dp, dp, dp, dp, dp, dp]),
            "orc_element_gp_global": (C.c_int, [C.c_int, dp, dp, dp]),
            "orc_element_edge_integrals": (C.c_int, [C.c_int, dp, dp, dp, dp, dp, dp]),
            "orc_wet_dry_single": (C.c_int, [dp, dp, dp, C.c_double, C.c_double, C.c_int, dp, ip, ip, dp]),
            "orc_set_plant": (None, [C.c_int]),
            "orc_R": (None, [C.c_double] * 6 + [dp]),
            "orc_invR": (None, [C.c_double] * 6 + [dp]),
            "orc_cs_limit_element": (None, [C.c_double] * 3 + [dp, dp, C.c_double, dp, dp, dp, dp, dp, ip]),
            "orc_apply_limiter": (C.c_int, [C.c_void_p]),
            "orc_limiter_element_geometry": (None, [C.c_void_p, C.c_int, dp, dp, dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
        _lib = lib
    return _lib


class OracleSim:
    """The oracle's serial (or OpenMP) RKDG stage on one mesh."""

    def __init__(self, mesh, params, threads=1):
        self.lib = load()
        self.mesh, self.params = mesh, params
        self.ne = mesh.n_elements
        self.ndof = (params.p + 1) * (params.p + 2) // 2
        self.h = self.lib.orc_create(C.byref(mesh.desc), C.byref(params))
        if not self.h:
            raise RuntimeError("oracle create failed: " + self.lib.orc_error().decode())
        self.lib.orc_set_threads(self.h, threads)

    @staticmethod
    def register_scheme(nstages, order, ark, brk):
        """Shu-Osher tables ([nstages, nstages], lower triangle) of a scheme the oracle does not
        restate, for oracles created afterwards."""
        a = np.ascontiguousarray(ark, np.float64)
        b = np.ascontiguousarray(brk, np.float64)
        return load().orc_register_scheme(nstages, order, N.dptr(a), N.dptr(b))

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        return self.lib.orc_step(self.h, n)

    def stage(self):
        return self.lib.orc_stage(self.h)

    def state(self):
        q = np.empty(3 * self.ndof * self.ne)
        wet = np.empty(self.ne, np.uint8)
        lim = np.empty(3 * self.ne, np.uint8)
        tr = np.empty(self.ne, np.uint8)
        self.lib.orc_get_state(self.h, N.dptr(q), N.uptr(wet), N.uptr(lim), N.uptr(tr))
        return {"q": q.reshape(3, self.ndof, self.ne), "wet": wet, "limited": lim.reshape(3, self.ne),
                "troubled": tr}

    def set_state(self, q, wet=None):
        q = np.ascontiguousarray(q, np.float64).reshape(-1)
        w = None if wet is None else np.ascontiguousarray(wet, np.uint8)
        self.lib.orc_set_state(self.h, N.dptr(q), N.uptr(w))

    def round_state(self, on=True):
        """Round the state to FP32 at the end of every stage (the FP32-storage error floor)."""
        self.lib.orc_round_state(self.h, 1 if on else 0)

    def keep_prelimit(self, on=True):
        """Keep each stage's state as it was after wetting-drying, before the limiter."""
        self.lib.orc_keep_prelimit(self.h, 1 if on else 0)

    def prelimit(self):
        q = np.empty(3 * self.ndof * self.ne)
        if not self.lib.orc_get_prelimit(self.h, N.dptr(q)):
            raise RuntimeError("no pre-limiter state kept")
        return q.reshape(3, self.ndof, self.ne)

    def limiter_geometry(self):
        """(neighbor [3][ne] by local face, -1 on a boundary; bathymetry at each barycentre [ne])."""
        nb = np.empty(3 * self.ne, np.int32)
        bb = np.empty(self.ne)
        self.lib.orc_limiter_geometry(self.h, N.iptr(nb), N.dptr(bb))
        return nb.reshape(3, self.ne), bb

    def time(self):
        t, s, k = C.c_double(), C.c_int32(), C.c_int32()
        self.lib.orc_get_time(self.h, C.byref(t), C.byref(s), C.byref(k))
        return t.value, s.value, k.value

    def nan_info(self):
        e, s, k = C.c_int32(), C.c_int32(), C.c_int32()
        self.lib.orc_nan_info(self.h, C.byref(e), C.byref(s), C.byref(k))
        return e.value, s.value, k.value

    def l2_residual(self):
        return self.lib.orc_l2_residual(self.h)
