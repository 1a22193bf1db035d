"""ctypes binding of the CPU oracle (oracle/liboforacle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, to check results or as the timed CPU baseline.  The
product path (oneflow_amd) never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboforacle.so"

OFX_F = dict(Q=0, TURBQ=1, RES=2, TURBRES=3, DQDX=4, DTDX=5, LIMITER=6, TIMESTEP=7, INVSR=8, VISSR=9,
             DQ=10, TURBDQ=11, TEMPR=12, VISL=13, VIST=14, DIST=15, VOL=16, XCC=17, FACE_NORMAL=18,
             FACE_CENTER=19, TURBDQDX=20, TURBIMPSR=21, IMPSR=22)


class OfxNorms(ctypes.Structure):
    _fields_ = [("ns", ctypes.c_double * 5), ("sa", ctypes.c_double)]


def build() -> None:
    """Compile the oracle with its Makefile (g++, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        P = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.ofo_last_error.restype = ctypes.c_char_p
        L.ofo_create.restype = P
        L.ofo_create.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ip, ip, ip, ip, ip]
        L.ofo_destroy.argtypes = [P]
        for name in ("ofo_init_freestream", "ofo_ns_boundary", "ofo_sa_boundary", "ofo_residual_ns",
                     "ofo_residual_sa", "ofo_lusgs_ns", "ofo_lusgs_sa", "ofo_update_ns", "ofo_update_sa"):
            getattr(L, name).argtypes = [P]
        L.ofo_mesh_info.argtypes = [P, ip]
        L.ofo_config_set.argtypes = [P, ctypes.c_void_p]
        L.ofo_get_field.argtypes = [P, ctypes.c_int, dp]
        L.ofo_set_field.argtypes = [P, ctypes.c_int, dp]
        L.ofo_timestep.argtypes = [P, ctypes.c_int]
        L.ofo_lhs.argtypes = [P, ctypes.c_int, ctypes.c_double]
        L.ofo_norms_get.argtypes = [P, ctypes.POINTER(OfxNorms)]
        L.ofo_iterate.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OfxNorms)]
        L.ofo_aero.argtypes = [P, dp]
        L.ofo_aero_row.argtypes = [P, dp]
        L.ofo_wall_count.argtypes = [P]
        L.ofo_wall_data.argtypes = [P, dp]
        L.ofo_face_flux.argtypes = [P, dp, dp]
        L.ofo_pow.argtypes = [dp, dp, ctypes.c_int, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class OracleCase:
    """One zone + case on the CPU oracle; methods mirror oneflow_amd.Solver."""

    def __init__(self, grid, cfg):
        L = lib()
        g = grid
        self._keep = [np.ascontiguousarray(a) for a in (g.x, g.y, g.z)]
        self._ikeep = [np.ascontiguousarray(a, dtype=np.int32) for a in
                       (g.face_node_count, g.face_nodes, g.lc, g.rc, g.bc_type)]
        x, y, z = self._keep
        fnc, fnn, lc, rc, bct = self._ikeep
        self.h = L.ofo_create(g.dim, g.n_nodes, _dp(x), _dp(y), _dp(z), g.n_faces, g.n_cells, g.n_bfaces,
                              _ip(fnc), _ip(fnn), _ip(lc), _ip(rc), _ip(bct))
        if not self.h:
            raise RuntimeError(L.ofo_last_error().decode())
        info = np.zeros(5, np.int32)
        L.ofo_mesh_info(self.h, _ip(info))
        self.n_cells, self.n_faces, self.n_bfaces, self.n_tcell = (int(v) for v in info[:4])
        self.cfg = cfg
        self._c = cfg.to_c()
        self._check(L.ofo_config_set(self.h, ctypes.byref(self._c)))

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}: {lib().ofo_last_error().decode()}")

    def __del__(self):
        if getattr(self, "h", None):
            lib().ofo_destroy(self.h)
            self.h = None

    def field_shape(self, name):
        T, F = self.n_tcell, self.n_faces
        return {"Q": (5, T), "RES": (5, T), "LIMITER": (5, T), "DQ": (5, T), "DQDX": (3, 5, T),
                "DTDX": (3, T), "TURBDQDX": (3, T), "XCC": (3, T), "FACE_NORMAL": (4, F),
                "FACE_CENTER": (3, F), "DIST": (self.n_cells,)}.get(name, (T,))

    def get(self, name):
        a = np.zeros(self.field_shape(name), np.float64)
        self._check(lib().ofo_get_field(self.h, OFX_F[name], _dp(a)))
        return a

    def set(self, name, value):
        a = np.ascontiguousarray(value, dtype=np.float64).reshape(self.field_shape(name))
        self._check(lib().ofo_set_field(self.h, OFX_F[name], _dp(a)))

    def init_freestream(self):
        self._check(lib().ofo_init_freestream(self.h))

    def ns_boundary(self):
        self._check(lib().ofo_ns_boundary(self.h))

    def sa_boundary(self):
        self._check(lib().ofo_sa_boundary(self.h))

    def timestep(self, it):
        self._check(lib().ofo_timestep(self.h, it))

    def residual_ns(self):
        self._check(lib().ofo_residual_ns(self.h))

    def residual_sa(self):
        self._check(lib().ofo_residual_sa(self.h))

    def lusgs_ns(self):
        self._check(lib().ofo_lusgs_ns(self.h))

    def lusgs_sa(self):
        self._check(lib().ofo_lusgs_sa(self.h))

    def update_ns(self):
        self._check(lib().ofo_update_ns(self.h))

    def update_sa(self):
        self._check(lib().ofo_update_sa(self.h))

    def lhs(self, solver, lhscoef):
        self._check(lib().ofo_lhs(self.h, solver, lhscoef))

    def norms(self):
        n = OfxNorms()
        self._check(lib().ofo_norms_get(self.h, ctypes.byref(n)))
        return np.array(list(n.ns) + [n.sa])

    def iterate(self, first_iter, n):
        arr = (OfxNorms * n)()
        self._check(lib().ofo_iterate(self.h, first_iter, n, arr))
        return np.array([list(a.ns) + [a.sa] for a in arr])

    def aero(self):
        a = np.zeros(4)
        self._check(lib().ofo_aero(self.h, _dp(a)))
        return a

    def aero_row(self):
        a = np.zeros(12)
        self._check(lib().ofo_aero_row(self.h, _dp(a)))
        return a

    def wall_data(self):
        """[n_wall][5]: x, y, z of each wall face centre, -cp, cf (face order)."""
        n = lib().ofo_wall_count(self.h)
        a = np.zeros((n, 5))
        self._check(lib().ofo_wall_data(self.h, _dp(a)))
        return a

    def face_flux(self):
        inv = np.zeros((5, self.n_faces))
        vis = np.zeros((5, self.n_faces))
        self._check(lib().ofo_face_flux(self.h, _dp(inv), _dp(vis)))
        return inv, vis
