"""Python view of libofx.so: one zone of the unstructured NS + SA solver on one GPU.

`Solver` mirrors the reference's task interface for the path: each method is
one registered task of the reference solver (UPDATE_RESIDUALS, CALC_TIME_STEP,
LUSGS_LOWER/UPPER_SWEEP, UPDATE_FLOWFIELD, CALC_BOUNDARY, ...) and `iterate`
is the whole pseudo-time step.  All compute runs in the HIP library; this
module only moves arrays across the C ABI.  It raises if the library is not
built or no GPU is visible -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from .config import CaseConfig
from .mesh import Grid

LIB_PATH = Path(__file__).resolve().parent / "libofx.so"

FIELD = dict(Q=0, TURBQ=1, RES=2, TURBRES=3, DQDX=4, DTDX=5, LIMITER=6, TIMESTEP=7, INVSR=8, VISSR=9,
             DQ=10, TURBDQ=11, TEMPR=12, VISL=13, VIST=14, DIST=15, VOL=16, XCC=17, FACE_NORMAL=18,
             FACE_CENTER=19, TURBDQDX=20, TURBIMPSR=21, IMPSR=22)


class OfxNorms(ctypes.Structure):
    _fields_ = [("ns", ctypes.c_double * 5), ("sa", ctypes.c_double)]


class OfxError(RuntimeError):
    """A libofx call failed; `code` is the OFX_E* return value."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


OFX_ENAN = -5


_lib = None


def load_library():
    """Load libofx.so (built by oneflow_amd.build); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise OfxError(f"{LIB_PATH} is not built; run __graft_entry__.build() or python oneflow_amd/build.py")
    L = ctypes.CDLL(str(LIB_PATH))
    P = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.ofx_last_error.restype = ctypes.c_char_p
    L.ofx_ctx_create.argtypes = [ctypes.POINTER(P), ctypes.c_int]
    L.ofx_ctx_destroy.argtypes = [P]
    L.ofx_mesh_upload.argtypes = [P, ctypes.c_int, ctypes.c_int, dp, dp, dp, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ip, ip, ip, ip, ip]
    L.ofx_mesh_info.argtypes = [P, ip]
    L.ofx_set_storage_order.argtypes = [P, ctypes.c_int]
    L.ofx_config_set.argtypes = [P, ctypes.c_void_p]
    for name in ("ofx_init_freestream", "ofx_ns_boundary", "ofx_sa_boundary", "ofx_residual_ns",
                 "ofx_residual_sa", "ofx_lusgs_ns", "ofx_lusgs_sa", "ofx_update_ns", "ofx_update_sa", "ofx_sync"):
        getattr(L, name).argtypes = [P]
    L.ofx_get_field.argtypes = [P, ctypes.c_int, dp]
    L.ofx_set_field.argtypes = [P, ctypes.c_int, dp]
    L.ofx_timestep.argtypes = [P, ctypes.c_int]
    L.ofx_lhs.argtypes = [P, ctypes.c_int, ctypes.c_double]
    L.ofx_norms_get.argtypes = [P, ctypes.POINTER(OfxNorms)]
    L.ofx_iterate.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OfxNorms)]
    L.ofx_aero.argtypes = [P, dp]
    L.ofx_aero_row.argtypes = [P, dp]
    L.ofx_time_kernel.argtypes = [P, ctypes.c_int, ctypes.c_int, dp]
    L.ofx_copy_bench.argtypes = [P, ctypes.c_size_t, ctypes.c_int, dp]
    L.ofx_zones_set.argtypes = [P, ctypes.c_int, ip, ip, ctypes.c_void_p]
    L.ofx_zones_norms.argtypes = [P, ctypes.POINTER(OfxNorms)]
    L.ofx_zones_aero.argtypes = [P, dp]
    L.ofx_set_face_tile.argtypes = [P, ctypes.c_int, ctypes.c_int]
    L.ofx_wall_count.argtypes = [P, ip]
    L.ofx_wall_data.argtypes = [P, dp]
    L.ofx_set_kernel_variant.argtypes = [P, ctypes.c_int, ctypes.c_int]
    L.ofx_get_kernel_variant.argtypes = [P, ctypes.c_int, ip]
    L.ofx_autotune.argtypes = [P, ctypes.c_int, ip]
    L.ofx_batch_create.argtypes = [ctypes.POINTER(P), ctypes.POINTER(P), ctypes.c_int, ctypes.c_int]
    L.ofx_batch_destroy.argtypes = [P]
    L.ofx_batch_iterate.argtypes = [P, ctypes.c_int, ctypes.c_int, ip]
    L.ofx_batch_norms.argtypes = [P, ctypes.POINTER(OfxNorms)]
    L.ofx_batch_autotune.argtypes = [P, ctypes.c_int, ip]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Solver:
    """One zone + case resident on one GPU (device buffers owned by the context)."""

    def __init__(self, grid: Grid, cfg: CaseConfig, device: int = 0, storage_order=None):
        L = load_library()
        self._L = L
        h = ctypes.c_void_p()
        self._check(L.ofx_ctx_create(ctypes.byref(h), device))
        self.h = h
        if storage_order is not None:  # 0 the grid's order (the library's default), 1 spatial
            self._check(L.ofx_set_storage_order(h, storage_order))
        g = grid
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (g.x, g.y, g.z))
        fnc, fnn, lc, rc, bct = (np.ascontiguousarray(a, dtype=np.int32) for a in
                                 (g.face_node_count, g.face_nodes, g.lc, g.rc, g.bc_type))
        self._check(L.ofx_mesh_upload(h, g.dim, g.n_nodes, _dp(x), _dp(y), _dp(z), g.n_faces, g.n_cells,
                                      g.n_bfaces, _ip(fnc), _ip(fnn), _ip(lc), _ip(rc), _ip(bct)))
        info = np.zeros(5, np.int32)
        self._check(L.ofx_mesh_info(h, _ip(info)))
        self.n_cells, self.n_faces, self.n_bfaces, self.n_tcell, self.n_levels = (int(v) for v in info)
        self.cfg = cfg
        self._c = cfg.to_c()
        self._check(L.ofx_config_set(h, ctypes.byref(self._c)))

    def _check(self, rc):
        if rc != 0:
            raise OfxError(f"libofx error {rc}: {self._L.ofx_last_error().decode()}", rc)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self._L.ofx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field_shape(self, name):
        T, F = self.n_tcell, self.n_faces
        return {"Q": (5, T), "RES": (5, T), "LIMITER": (5, T), "DQ": (5, T), "DQDX": (3, 5, T),
                "DTDX": (3, T), "TURBDQDX": (3, T), "XCC": (3, T), "FACE_NORMAL": (4, F),
                "FACE_CENTER": (3, F), "DIST": (self.n_cells,)}.get(name, (T,))

    def get(self, name):
        a = np.zeros(self.field_shape(name), np.float64)
        self._check(self._L.ofx_get_field(self.h, FIELD[name], _dp(a)))
        return a

    def set(self, name, value):
        a = np.ascontiguousarray(value, dtype=np.float64).reshape(self.field_shape(name))
        self._check(self._L.ofx_set_field(self.h, FIELD[name], _dp(a)))

    # ---- tasks (reference task names in brackets)
    def init_freestream(self):        # INIT_FLOWFIELD + INIT_FINAL
        self._check(self._L.ofx_init_freestream(self.h))

    def ns_boundary(self):            # CALC_BOUNDARY (NS)
        self._check(self._L.ofx_ns_boundary(self.h))

    def sa_boundary(self):            # CALC_BOUNDARY (SA)
        self._check(self._L.ofx_sa_boundary(self.h))

    def timestep(self, it):           # CALC_TIME_STEP
        self._check(self._L.ofx_timestep(self.h, it))

    def residual_ns(self):            # UPDATE_RESIDUALS (NS)
        self._check(self._L.ofx_residual_ns(self.h))

    def residual_sa(self):            # UPDATE_RESIDUALS (SA)
        self._check(self._L.ofx_residual_sa(self.h))

    def lusgs_ns(self):               # LUSGS_LOWER_SWEEP + LUSGS_UPPER_SWEEP (NS)
        self._check(self._L.ofx_lusgs_ns(self.h))

    def lusgs_sa(self):               # LUSGS_LOWER_SWEEP + LUSGS_UPPER_SWEEP (SA)
        self._check(self._L.ofx_lusgs_sa(self.h))

    def update_ns(self):              # UPDATE_FLOWFIELD_LUSGS (NS)
        self._check(self._L.ofx_update_ns(self.h))

    def update_sa(self):              # UPDATE_FLOWFIELD_LUSGS (SA)
        self._check(self._L.ofx_update_sa(self.h))

    def lhs(self, solver, lhscoef):  # CALC_LHS (multi-stage): solver 0 NS, 1 SA
        self._check(self._L.ofx_lhs(self.h, solver, lhscoef))

    def sync(self):
        self._check(self._L.ofx_sync(self.h))

    def norms(self, check=True):      # DUMP_RESIDUAL
        """Residual norms [ns0..ns4, sa]; a NaN raises (OFX_ENAN) unless check is False."""
        n = OfxNorms()
        rc = self._L.ofx_norms_get(self.h, ctypes.byref(n))
        if check or rc != OFX_ENAN:
            self._check(rc)
        return np.array(list(n.ns) + [n.sa])

    def iterate(self, first_iter, n, with_norms=True):
        if with_norms:
            arr = (OfxNorms * n)()
            self._check(self._L.ofx_iterate(self.h, first_iter, n, arr))
            return np.array([list(a.ns) + [a.sa] for a in arr])
        self._check(self._L.ofx_iterate(self.h, first_iter, n, None))
        return None

    def aero(self):                   # DUMP_AERODYNAMIC: CL, CD, CD_pres, CD_visc
        a = np.zeros(4)
        self._check(self._L.ofx_aero(self.h, _dp(a)))
        return a

    def aero_row(self):               # the 12 value columns of an aero.dat row
        a = np.zeros(12)
        self._check(self._L.ofx_aero_row(self.h, _dp(a)))
        return a

    KERNELS = {"grad": 0, "limiter": 1, "face": 2, "gather": 3, "ns_residual": 4, "sa_residual": 5,
               "sa_grad": 6, "sa_face": 7, "sa_cell": 8, "residual": 9, "pseudo_step": 10,
               "timestep": 11, "lusgs_ns": 12}

    def wall_data(self):               # DUMP_PRESSURE_COEFF
        """[n_wall][5]: x, y, z of each wall face centre, -cp, cf (face order)."""
        n = np.zeros(1, np.int32)
        self._check(self._L.ofx_wall_count(self.h, _ip(n)))
        a = np.zeros((int(n[0]), 5))
        if n[0]:
            self._check(self._L.ofx_wall_data(self.h, _dp(a)))
        return a

    def time_kernel(self, which, iters=100):
        """Average ms of one residual part over `iters` back-to-back launches."""
        a = np.zeros(1)
        self._check(self._L.ofx_time_kernel(self.h, self.KERNELS[which], iters, _dp(a)))
        return float(a[0])

    def set_face_tile(self, ns_tile, sa_tile):
        """Threads per block of the NS / SA face-flux kernels (64..512)."""
        self._check(self._L.ofx_set_face_tile(self.h, ns_tile, sa_tile))

    def kernel_variant(self, kernel):
        v = np.zeros(1, np.int32)
        self._check(self._L.ofx_get_kernel_variant(self.h, kernel, _ip(v)))
        return int(v[0])

    def set_kernel_variant(self, kernel, variant):
        """Launch variant 0 or 1 of kernel 0 or 1 (the NS gradient with limiter), 2 (gather),
        3 (NS face split), 4 (SA face geometry), 5 (SA gradient) or 6 (NS time step and
        LU-SGS sweeps); results are identical."""
        self._check(self._L.ofx_set_kernel_variant(self.h, kernel, variant))

    def autotune(self, iters=20):
        """Pick the fastest face tiles and gradient variant on the current state; returns (ns_tile, sa_tile)."""
        a = np.zeros(2, np.int32)
        self._check(self._L.ofx_autotune(self.h, iters, _ip(a)))
        return int(a[0]), int(a[1])

    def copy_bench(self, nbytes, iters=20):
        """Average ms of a plain device copy of nbytes (bandwidth reference)."""
        a = np.zeros(1)
        self._check(self._L.ofx_copy_bench(self.h, ctypes.c_size_t(nbytes), iters, _dp(a)))
        return float(a[0])


class Batch:
    """Several solvers advanced together: steps past the CFL ramp are recorded
    once into one HIP graph (one branch per case) and replayed.  Per case the
    results equal `Solver.iterate` bit for bit."""

    def __init__(self, solvers, steps_per_graph=2):
        self._L = load_library()
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s.h for s in self.solvers])
        self.h = ctypes.c_void_p()
        rc = self._L.ofx_batch_create(ctypes.byref(self.h), arr, len(self.solvers), steps_per_graph)
        if rc != 0:
            raise OfxError(f"ofx_batch_create: {self._L.ofx_last_error().decode()}")

    def _check(self, rc):
        if rc != 0:
            raise OfxError(f"ofx error {rc}: {self._L.ofx_last_error().decode()}", rc)

    def iterate(self, first_iter, n):
        """Advance every case n steps; returns the number of graph replays used."""
        k = np.zeros(1, np.int32)
        self._check(self._L.ofx_batch_iterate(self.h, first_iter, n, _ip(k)))
        return int(k[0])

    def norms(self):
        arr = (OfxNorms * len(self.solvers))()
        self._check(self._L.ofx_batch_norms(self.h, arr))
        return np.array([list(a.ns) + [a.sa] for a in arr])

    def autotune(self, iters=20):
        a = np.zeros(2 * len(self.solvers), np.int32)
        self._check(self._L.ofx_batch_autotune(self.h, iters, _ip(a)))
        return a.reshape(-1, 2)

    def close(self):
        if self.h:
            self._L.ofx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ZoneSolver(Solver):
    """Several independent cases in one context: their grids concatenated
    (meshgen.concat_grids) and every kernel run over all of them at once, with
    per-case freestream constants and reductions.  Each case's results equal
    a Solver on that case alone, bit for bit."""

    def __init__(self, grids, cfgs, device: int = 0, storage_order=None):
        from .meshgen import concat_grids
        g, cs, bs = concat_grids(grids)
        super().__init__(g, cfgs[0], device, storage_order)
        self.n_zones = len(grids)
        self.cell_start, self.bface_start = cs, bs
        self.zone_grids = list(grids)
        self.zone_cfgs = list(cfgs)
        arr = (type(self._c) * len(cfgs))(*[c.to_c() for c in cfgs])
        self._zc = arr
        self._check(self._L.ofx_zones_set(self.h, len(grids), _ip(cs), _ip(bs), ctypes.cast(arr, ctypes.c_void_p)))

    def zone_norms(self):
        arr = (OfxNorms * self.n_zones)()
        self._check(self._L.ofx_zones_norms(self.h, arr))
        return np.array([list(a.ns) + [a.sa] for a in arr])

    def zone_aero(self):
        a = np.zeros(4 * self.n_zones)
        self._check(self._L.ofx_zones_aero(self.h, _dp(a)))
        return a.reshape(-1, 4)

    def zone_slice(self, z):
        """(cell slice, ghost slice) of zone z in cell-indexed fields."""
        C = self.n_cells
        return (slice(int(self.cell_start[z]), int(self.cell_start[z + 1])),
                slice(C + int(self.bface_start[z]), C + int(self.bface_start[z + 1])))

    def get_zone(self, name, z):
        """A cell field of zone z in the zone's own layout (cells, then ghosts)."""
        a = self.get(name)
        cs, gs = self.zone_slice(z)
        return np.concatenate([a[..., cs], a[..., gs]], axis=-1)

    def set_zone(self, name, z, value):
        a = self.get(name)
        cs, gs = self.zone_slice(z)
        nc = cs.stop - cs.start
        a[..., cs] = value[..., :nc]
        a[..., gs] = value[..., nc:]
        self.set(name, a)
