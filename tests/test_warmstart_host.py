"""Host logic of the warm-start solve loops (no GPU): per-case status, the
NaN stop and the removal of finished cases, on a scripted stand-in for the
zone solver whose norms follow a fixed schedule per case."""
import numpy as np

from oneflow_amd import warmstart as ws
from oneflow_amd.solver import OFX_ENAN, OfxError


class FakeZones:
    """Zones whose density norm at step k is schedule[z](k); iterate raises
    OFX_ENAN after the call when any zone's norm is NaN, as ofx_iterate does."""

    class _Cfg:
        ncfl = 0

    def __init__(self, schedules, ids=None):
        self.sched = schedules
        self.ids = list(range(len(schedules))) if ids is None else ids
        self.n_zones = len(self.ids)
        self.cfg = self._Cfg()
        self.step = 0
        self.cell_start = np.arange(self.n_zones + 1)
        self.bface_start = np.zeros(self.n_zones + 1, int)
        self.closed = False

    def sync(self):
        pass

    def iterate(self, it, n, with_norms=False):
        self.step = it + n - 1
        if np.any(np.isnan(self.zone_norms()[:, 0])):
            raise OfxError("libofx error -5: NotANumber in the residual", OFX_ENAN)

    def zone_norms(self):
        v = np.array([self.sched[c](self.step) for c in self.ids])
        return np.stack([v] * 6, axis=1)

    def zone_aero(self):
        return np.array([[float(c), 1.0, 0.0, 0.0] for c in self.ids])

    def zone_slice(self, z):
        return slice(z, z + 1), slice(self.n_zones + z, self.n_zones + z)

    def get(self, name):
        return np.zeros((5, self.n_zones)) if name == "Q" else np.zeros(self.n_zones)

    def set(self, name, v):
        pass

    def ns_boundary(self):
        pass

    def sa_boundary(self):
        pass

    def close(self):
        self.closed = True


def schedules():
    conv = lambda k: 10.0 ** (-k / 100.0)          # 1e-8 of 1 at step 800
    slow = lambda k: 1.0                           # never converges
    nan_at = lambda k: np.nan if k >= 300 else 1.0  # diverges at step 300
    return [conv, slow, nan_at]


def test_status_per_case_without_rebuild():
    fz = FakeZones(schedules())
    t, steps, aero, status = ws.solve_zones_to_tolerance(fz, [1.0, 1.0, 1.0], tol=1e-8, chunk=100,
                                                         max_steps=1000)
    assert status == [ws.CONVERGED, ws.MAX_STEPS, ws.DIVERGED]
    assert steps == [800, 1000, 300]
    assert aero[2] is None and aero[0][0] == 0.0 and aero[1][0] == 1.0


def test_diverged_case_leaves_the_context():
    made = []

    def rebuild(ids):
        z = FakeZones(schedules(), ids)
        made.append(ids)
        return z
    fz = FakeZones(schedules())
    t, steps, aero, status = ws.solve_zones_to_tolerance(fz, [1.0] * 3, tol=1e-8, chunk=100, max_steps=1000,
                                                         rebuild=rebuild, compact=0.0)
    # the NaN case forces a rebuild with the two running cases at once
    assert made[0] == [0, 1]
    assert status == [ws.CONVERGED, ws.MAX_STEPS, ws.DIVERGED]


def test_case_status_rules():
    assert ws._case_status(np.nan, 1.0, 10, 100) == ws.DIVERGED
    assert ws._case_status(np.inf, 1.0, 10, 100) == ws.DIVERGED
    assert ws._case_status(0.5, 1.0, 10, 100) == ws.CONVERGED
    assert ws._case_status(2.0, 1.0, 101, 100) == ws.MAX_STEPS
    assert ws._case_status(2.0, 1.0, 10, 100) is None


def test_every_case_leaves_once_with_its_context():
    """on_leave sees each case exactly once, in the context that holds its
    last state (the first context for the cases finished before a rebuild)."""
    seen = []

    def rebuild(ids):
        return FakeZones(schedules(), ids)
    fz = FakeZones(schedules())
    ws.solve_zones_to_tolerance(fz, [1.0] * 3, tol=1e-8, chunk=100, max_steps=1000, rebuild=rebuild, compact=0.0,
                                on_leave=lambda ctx, pairs: seen.extend((ctx.ids[z], cid) for z, cid in pairs))
    assert sorted(cid for _, cid in seen) == [0, 1, 2]
    assert all(a == b for a, b in seen)


def test_field_cache_reads_each_field_once():
    calls = []

    class Z(FakeZones):
        zone_cfgs = [None] * 3

        def get(self, name):
            calls.append(name)
            return np.arange(5 * self.n_zones, dtype=float).reshape(5, -1) if name == "Q" else np.zeros(self.n_zones)
    fc = ws.FieldCache(Z(schedules()))
    a = [ws.ZoneView(fc, z).get("Q") for z in range(3)]
    assert calls == ["Q"] and a[2][0, 0] == 2.0
