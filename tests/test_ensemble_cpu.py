"""The ensemble generator and the ensemble's host description (no GPU):
SplitMix64 against its published outputs, the draw's ranges and relaxation
times, and the C structure the library reads."""
import ctypes
import math

import pytest


def test_splitmix64_known_outputs():
    """The first outputs of SplitMix64 from seed 0 (java.util.SplittableRandom(0).nextLong()
    and the reference C implementation of the generator)."""
    from amrlbm.ensemble import splitmix64
    assert [splitmix64(0, k) for k in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_draw_is_reproducible_and_in_range():
    from amrlbm.ensemble import ensemble, TAU_MIN
    a, b = ensemble(64, 2024), ensemble(64, 2024)
    assert a.members == b.members
    assert ensemble(64, 2025).members != a.members
    for w in a.members:
        levels = w.max_levels - w.min_levels + 1
        assert w.L0 in (256, 512, 1024) and 2 <= levels <= 4
        assert w.max_levels == int(math.log2(w.L0 // 16))
        assert 16.0 <= w.D <= 64.0 and (w.x0, w.y0) == (w.L0 / 2, w.L0 / 2)
        assert 1.0 <= w.Bo <= 100.0
        assert 10.0 <= 1.0 / w.rhol <= 1000.0 and 1.0 <= w.visc_ratio <= 100.0
        # tau_l / tau_h = density ratio / viscosity ratio; the lighter one is TAU_MIN
        assert math.isclose(w.taul / w.tauh, (1.0 / w.rhol) / w.visc_ratio, rel_tol=1e-12)
        assert math.isclose(min(w.tauh, w.taul), TAU_MIN, rel_tol=1e-9)
        # Eo = g (rho_h - rho_l) D^2 / sigma
        assert math.isclose(w.gy * (w.rhoh - w.rhol) * w.D ** 2 / w.sigma, w.Bo, rel_tol=1e-12)
    e = ensemble(64, 2024)
    assert e.nroots == 128 and e.L0 == 1024 and e.max_levels == 6
    assert e.fine_cells == sum(2 * w.L0 * w.L0 for w in e.members)


def test_instance_struct_layout():
    """amrlbm_instance: three ints, then 13 doubles at 8-byte alignment."""
    from amrlbm import Instance, instance_from_workload
    from amrlbm.ensemble import member
    assert ctypes.sizeof(Instance) == 16 + 13 * 8
    assert Instance.M0.offset == 16 and Instance.y0.offset == 16 + 12 * 8
    w = member("x", 128, 2, 24.0, 10.0, 100.0, 10.0)
    d = instance_from_workload(w)
    assert (d.L0, d.min_levels, d.max_levels) == (128, 2, 3)
    assert (d.R, d.x0, d.y0, d.W) == (12.0, 64.0, 64.0, 5.0)
    assert d.tauh == w.tauh and d.sigma == w.sigma
