"""What the trained network gains: the learned warm start against the plain
solver from freestream, end to end on the GPU.

Nine training airfoils (one NACA 4-digit section, Mach 0.30/0.40/0.50 x AoA
0/2/4 deg, Re 5e6, Roe + SA, first order) are solved as zones of one
context; the BF16 MLP is trained on their fields by the library's own
kernels; two unseen cases inside that range (Mach 0.35 / 0.45, AoA 1 / 3)
then start from the network's prediction and from freestream, and both
iterate with the same solver to the same absolute residual level (a 1e-8 drop
from the plain run's peak density residual).  The test asserts that the
learned start needs fewer steps on every unseen case, and that both reach the
same lift and drag to 1e-6 relative -- the solver acting as fallback and
judge.
"""
import numpy as np
import pytest

from oneflow_amd.cases import AirfoilCase

pytestmark = pytest.mark.gpu

TOL = 1e-8
SECTION = dict(m=0.02, p=0.4, t=0.12, reynolds=5.0e6, target_cells=6000)


def cfg_of(case):
    return case.config(ilim=-1, cfled=10.0)


def test_learned_start_saves_steps_at_equal_forces():
    from oneflow_amd import warmstart as ws
    from oneflow_amd.mlp import Mlp
    from oneflow_amd.solver import ZoneSolver
    train = [AirfoilCase(mach=m, aoa=a, **SECTION) for m in (0.30, 0.40, 0.50) for a in (0.0, 2.0, 4.0)]
    test = [AirfoilCase(mach=0.35, aoa=1.0, **SECTION), AirfoilCase(mach=0.45, aoa=3.0, **SECTION)]
    g = train[0].grid("multicolor")
    print(f"[warmstart] grid cells={g.n_cells}", flush=True)

    # 1. training fields: the nine cases converged together as zones, each to
    #    a 1e-8 drop from its peak density residual over the first 1000 steps
    zt = ZoneSolver([g] * len(train), [cfg_of(c) for c in train])
    zt.init_freestream()
    peak = np.zeros(len(train))
    for it in range(1, 1001, 100):
        zt.iterate(it, 100, with_norms=False)
        peak = np.maximum(peak, zt.zone_norms()[:, 0])
    _, steps, _, status = ws.solve_zones_to_tolerance(zt, list(peak), tol=TOL, chunk=2000, max_steps=250_000,
                                                      first_iter=1001)
    drop = zt.zone_norms()[:, 0] / peak
    print(f"[warmstart] training fields: steps {steps}, status {status}, drop {drop}", flush=True)
    assert status == ["converged"] * len(train)
    views = [ws.ZoneView(zt, z) for z in range(len(train))]
    X = np.concatenate([ws.cell_features(v, c) for v, c in zip(views, train)])
    Y = np.concatenate([ws.cell_targets(v, c, v.cfg) for v, c in zip(views, train)])
    zt.close()

    # 2. the network, trained on the GPU
    net = Mlp(seed=5)
    losses = net.train(X, Y, batch=4096, steps=3000, lr=1e-3)
    print(f"[warmstart] {X.shape[0]} samples, loss {losses[:20].mean():.3e} -> {losses[-50:].mean():.3e}",
          flush=True)
    assert losses[-50:].mean() < 0.05 * losses[:20].mean()

    # 3. unseen cases: the reference level is the plain run's peak density norm
    refs = []
    for c in test:
        z = ZoneSolver([g], [cfg_of(c)])
        z.init_freestream()
        refs.append(float(np.max(z.iterate(1, 1000)[:, 0])))
        z.close()

    def run(learned):
        z = ZoneSolver([g] * len(test), [cfg_of(c) for c in test])
        z.init_freestream()
        if learned:
            ws.apply_predictions_zones(z, test, net)
        out = ws.solve_zones_to_tolerance(z, refs, tol=TOL, chunk=200, max_steps=250_000)
        z.close()
        return out
    _, st_plain, aero_plain, stat_plain = run(False)
    _, st_learn, aero_learn, stat_learn = run(True)
    print(f"[warmstart] steps plain {st_plain}, learned {st_learn}", flush=True)
    assert stat_plain == ["converged"] * len(test) and stat_learn == ["converged"] * len(test)
    for i in range(len(test)):
        assert st_learn[i] < st_plain[i], (i, st_learn[i], st_plain[i])
        a, b = np.array(aero_learn[i][:2]), np.array(aero_plain[i][:2])
        rel = np.abs(a - b) / np.abs(b)
        print(f"[warmstart] case {i}: CL {a[0]:.9f} vs {b[0]:.9f}, CD {a[1]:.9f} vs {b[1]:.9f}, rel {rel}",
              flush=True)
        assert np.all(rel <= 1e-6), (i, rel)
    net.close()
