"""CPU-side checks of the C-ABI library: it loads, exports every function
include/amrlbm.h declares, and fails cleanly (error code + message, no
crash) where no GPU is present."""
import ctypes
import os

import pytest


def test_library_exports_header_symbols():
    import amrlbm
    names = amrlbm.exported_functions()
    assert len(names) >= 25
    L = amrlbm.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    """Every C-ABI function has an argtypes signature in the binding."""
    import amrlbm
    L = amrlbm.lib()
    for n in amrlbm.exported_functions():
        f = getattr(L, n)
        assert f.argtypes is not None or n == "amrlbm_last_error", n


def test_create_rejects_bad_parameters():
    import amrlbm
    w = amrlbm.WORKLOADS["tiny"]
    p = amrlbm.params_from_workload(w)
    p.nx = p.ny = 8
    h = ctypes.c_void_p()
    rc = amrlbm.lib().amrlbm_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == 2
    assert b"16" in amrlbm.lib().amrlbm_last_error()
    p = amrlbm.params_from_workload(w)
    p.L0 = 100
    assert amrlbm.lib().amrlbm_create(ctypes.byref(p), ctypes.byref(h)) == 2


def test_workload_physics_matches_reference_derivation(workloads):
    """Set_Rising_Bubble_Parameters (AMR_Rising_2D_OMP.f90:241-259) at D = 20:
    the values the reference build prints for the tiny workload."""
    w = workloads["tiny"]
    assert w.sigma == 4.9950000000000008e-05
    assert w.beta == 1.1988000000000002e-04
    assert w.kappa == 3.7462500000000007e-04
    assert w.taul == 1.4992498124061915e-01
    assert w.tauh == 1.4992498124061913e-02
    assert w.max_levels == 2 and w.max_blocks == 96
    c2 = workloads["cfg2"]
    assert c2.fine_shape == (1024, 2048) and c2.max_levels == 6 and c2.min_levels == 3


def test_regrid_interval_choice():
    """amrlbm.tuning.choose keeps the fastest interval whose diagnostics stay
    within every bound against the baseline row, and the baseline itself when
    no other row qualifies."""
    from amrlbm import tuning

    def row(k, us, rise=0.0, cen=0.0, mass=0.0):
        return {"regrid_every": k, "us_per_step": us, "rise_velocity_rel": rise, "centroid_abs": cen,
                "mass_rel": mass}
    rows = [row(1, 40.0), row(10, 26.0, 1e-3, 0.02), row(20, 25.0, 4e-3, 0.05), row(50, 24.0, 1e-2, 0.3)]
    assert tuning.choose(rows)["regrid_every"] == 20
    assert tuning.choose(rows, rise_rel=2e-3)["regrid_every"] == 10
    rows[2]["mass_rel"] = 2e-6
    assert tuning.choose(rows)["regrid_every"] == 10
    assert tuning.choose([row(1, 40.0), row(10, 26.0, 1.0)])["regrid_every"] == 1
