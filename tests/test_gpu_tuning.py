"""Launch variants (amrlbm_set_tuning) and the autotuner (amrlbm_autotune).
The leaves' launch order, the number of hg fill operands in flight per
thread and the execution (step graph or persistent segment kernel) change no
arithmetic, so every variant must leave the state bit for bit the same. The
autotuner times all 24 on the live mesh, keeps the fastest, and leaves the
state as it found it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm as mod
    mod.lib()
    return mod


def _state(m):
    rows, free = m.topology()
    arrays = {n: m.get_block(n) for n in range(1, rows.shape[0] + 1) if rows[n - 1, 0] >= 0 and rows[n - 1, 5]}
    return rows, free, arrays


def _same(a, b):
    ra, fa, xa = a
    rb, fb, xb = b
    assert np.array_equal(ra, rb) and fa == fb and xa.keys() == xb.keys()
    for n in xa:
        for f in ("C", "P", "Ux", "Uy"):
            assert np.array_equal(xa[n][f], xb[n][f]), (n, f)
        for f in ("h", "g"):   # interiors (the fused stream leaves the previous ghost ring)
            assert np.array_equal(xa[n][f][:, 1:-1, 1:-1], xb[n][f][:, 1:-1, 1:-1]), (n, f)


@pytest.mark.parametrize("precision", [64, 32])
def test_variants_are_bitwise_equal(amrlbm, workloads, precision):
    w = workloads["cfg2"]
    variants = [(o, d, e) for e in range(2) for o in range(4) for d in range(3)] if precision == 64 else \
        [(0, 0, 0), (1, 1, 0), (2, 2, 0), (3, 1, 0), (0, 1, 1), (3, 0, 1), (2, 2, 1)]
    ref = None
    for order, depth, ex in variants:
        print(f"[tuning] FP{precision} cfg2 variant order {order} depth {depth} execution {ex}: 2000 steps", flush=True)
        with amrlbm.Mesh(w, precision=precision) as m:
            m.setup_rising_bubble()
            m.set_tuning(order, depth, ex)
            assert m.get_tuning() == (order, depth, ex)
            m.step(2000, 10)
            st = _state(m)
            assert m.regrid_stats()["changes"] >= 2   # across mesh changes
        if ref is None:
            ref = st
        else:
            _same(st, ref)


def test_autotune_keeps_the_fastest_and_the_state(amrlbm, workloads):
    w = workloads["cfg2"]
    with amrlbm.Mesh(w) as a, amrlbm.Mesh(w) as b:
        a.setup_rising_bubble()
        b.setup_rising_bubble()
        a.step(100, 10)
        b.step(100, 10)
        before = _state(a)
        table = a.autotune(reps=10)
        for order, depth, ex, us in table:
            print(f"[tuning] cfg2 order {order} depth {depth} execution {ex}: {us:.2f} us per step")
        assert sorted((o, d, e) for o, d, e, _ in table) == [(o, d, e) for o in range(4) for d in range(3)
                                                              for e in range(2)]
        assert all(us > 0 for _, _, e, us in table if e == 0)
        best = min((r for r in table if r[3] > 0), key=lambda r: r[3])
        assert a.get_tuning() == best[:3]
        # the fields as they were (the diagnostics sum leaves in launch order, so they may differ in the last bit)
        _same(_state(a), before)
        a.step(200, 10)
        b.step(200, 10)
        _same(_state(a), _state(b))


def test_regrid_interval_sweep(amrlbm, workloads):
    """The regrid-interval knob (amrlbm.tuning) on the faster-rising 128x256
    bubble over 2,200 steps: the every-step run is the baseline, the
    interval-10 run states its distance from it, and the choice is one of
    the swept intervals within the bounds."""
    from amrlbm import tuning
    rows = tuning.sweep(workloads["cfg1_rise"], intervals=(1, 10), steps=2200)
    for r in rows:
        print(f"[interval] every {r['regrid_every']}: {r['us_per_step']} us/step, {r['changes']} changes, "
              f"rise velocity {r['rise_velocity_rel']:.2e}, centroid {r['centroid_abs']:.3f}, mass {r['mass_rel']:.2e}",
              flush=True)
    assert rows[0]["rise_velocity_rel"] == 0.0 and rows[0]["changes"] >= 3
    assert rows[1]["changes"] >= 1 and rows[1]["diag"][8] == 0
    best = tuning.choose(rows)
    assert best["regrid_every"] in (1, 10)
