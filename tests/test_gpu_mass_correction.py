"""The phase-field mass correction (amrlbm_set_mass_correction): each
instance's FP64 phase-field mass S = sum C dx^2 is held at its budget, the
value when the correction was switched on plus the change each later mesh
change's restrict and prolong make, by a correction at every segment end.

Output's mass (AMR_Rising_2D_OMP.f90:478) is sum (rhol + C (rhoh - rhol))
dx^2 = rhol A + (rhoh - rhol) S over a domain of A finest cells, so the
checks read it from amrlbm_output: after every stepping call it must equal
the budget, and the budget must move at a mesh change by exactly what that
change moved Output's mass by (read before and after the regrid). The
stepped runs cross mesh changes (the rising bubble at 1024x2048 changes its
mesh at steps 709 and 1079 with a regrid every 10 steps). The uncorrected
path is the reference's: off, nothing changes (bit for bit)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm as mod
    mod.lib()
    return mod


def _budget_mass(w, target):
    """Output's mass of a state whose phase-field mass is target."""
    return w.rhol * w.fine_cells + (w.rhoh - w.rhol) * target


@pytest.mark.parametrize("precision", [32, 64])
def test_mass_held_at_budget_across_mesh_changes(amrlbm, workloads, precision):
    w = workloads["cfg2"]
    cycles, k = 120, 10
    print(f"[mass] cfg2 FP{precision}: {cycles} x ({k} steps, regrid), correction on", flush=True)
    with amrlbm.Mesh(w, precision=precision) as m:
        m.setup_rising_bubble()
        m.set_mass_correction(True)
        budget = m.output()[3]
        t0 = m.mass_budget()[0, 0]
        np.testing.assert_allclose(_budget_mass(w, t0), budget, rtol=1e-14)
        changes, worst, lam = 0, 0.0, 0.0
        for c in range(cycles):
            m.step(k, 1 << 30)   # one segment, no regrid in it: corrected at its end
            a = m.output()[3]
            worst = max(worst, abs(a - budget) / budget)
            lam = max(lam, abs(m.mass_budget()[0, 3]))
            n0 = m.counts()["nleaf"]
            m.refine_derefine_mesh()
            b = m.output()[3]
            if m.counts()["nleaf"] != n0:
                changes += 1
                budget += b - a   # the budget takes the change's own mass change
            else:
                assert b == a
            tgt = m.mass_budget()[0, 0]
            np.testing.assert_allclose(_budget_mass(w, tgt), budget, rtol=1e-12)
        print(f"[mass] FP{precision}: {changes} mesh changes, largest |mass - budget| / mass {worst:.2e}, "
              f"largest |lambda| {lam:.2e}, budget moved by {(budget - m.output()[3]) / budget:.2e} since")
        assert changes >= 2
        assert worst <= 1e-11
        assert 0.0 < lam < 1e-4


def test_corrected_run_stays_close_to_the_reference_path(amrlbm, workloads):
    """The correction moves C by lambda 4 C (1 - C) and keeps each cell's
    velocity and pressure: over 4,000 FP32 steps with mesh changes the
    bubble's centroid and rise velocity stay within 0.05 cells (the FP32
    path's own bound against FP64) and 0.1 % of the uncorrected run's, while the uncorrected run's own mass moves (the
    reference path's interpolation). (Near step 1,200 the rise velocity
    passes through a minimum of 5e-3, 7 % of its value at step 400, so
    relative differences there say little.)"""
    w = workloads["cfg2"]
    steps, k = 4000, 10
    out = {}
    for on in (False, True):
        with amrlbm.Mesh(w, precision=32) as m:
            m.setup_rising_bubble()
            m0 = m.output()[3]
            if on:
                m.set_mass_correction(True)
            m.step(steps, k)
            out[on] = (m0, m.output(), m.counts()["nleaf"])
    (m0, d0, n0), (_, d1, n1) = out[False], out[True]
    cen = max(abs(d1[5] - d0[5]), abs(d1[6] - d0[6]))
    rise = abs(d1[7] - d0[7]) / abs(d0[7])
    print(f"[mass] FP32 {steps} steps: uncorrected mass moved {(d0[3] - m0) / m0:.2e}, corrected "
          f"{(d1[3] - m0) / m0:.2e}; centroid {cen:.2e} cells, rise velocity {rise:.2e} apart; leaves {n0}, {n1}")
    assert cen <= 0.05 and rise <= 1e-3
    assert abs(d1[3] - m0) < abs(d0[3] - m0)


def test_per_instance_budgets_in_an_ensemble(amrlbm):
    """Five instances of uneven size, level range and fluid in one FP32 mesh:
    each instance's Output mass stays at its own budget."""
    from amrlbm.ensemble import Ensemble, member
    members = (member("a", 64, 2, 20.0, 10.0, 1000.0, 100.0), member("b", 128, 3, 32.0, 40.0, 100.0, 10.0),
               member("c", 256, 2, 48.0, 5.0, 10.0, 1.0), member("d", 128, 2, 24.0, 80.0, 500.0, 50.0),
               member("e", 256, 4, 40.0, 20.0, 1000.0, 10.0))
    ens = Ensemble("t5", members, block_cap=4096)
    with amrlbm.Mesh(ens, precision=32) as e:
        e.setup_rising_bubble()
        e.set_mass_correction(True)
        e.step(300, 10)
        d = e.output_instances()
        tg = e.mass_budget()
        for q, w in enumerate(members):
            want = _budget_mass(w, tg[q, 0])
            print(f"[mass] instance {w.name}: mass {d[q, 3]:.12e}, budget {want:.12e}, lambda {tg[q, 3]:.2e}")
            # FP32 storage: a small instance's corrected C is rounded to its ulp
            assert abs(d[q, 3] - want) <= 1e-9 * want


def test_off_leaves_the_reference_path(amrlbm, workloads):
    """Switched on and off again before stepping, the FP64 path is bit for
    bit the one that never had it."""
    w = workloads["cfg1"]
    arrays = []
    for toggle in (False, True):
        with amrlbm.Mesh(w) as m:
            m.setup_rising_bubble()
            if toggle:
                m.set_mass_correction(True)
                m.set_mass_correction(False)
                with pytest.raises(amrlbm.AmrLbmError):
                    m.mass_budget()
            m.step(40, 10)
            arrays.append({n: m.get_block(n) for n in m.leaves()})
    assert arrays[0].keys() == arrays[1].keys()
    for n, a in arrays[0].items():
        for f in ("C", "P", "Ux", "Uy"):
            assert np.array_equal(a[f], arrays[1][n][f]), (n, f)
