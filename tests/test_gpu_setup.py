"""Device-resident set-up (amrlbm_setup_bubbles): the phase field profile of
Initialize_AMR_Rising_Bubble (AMR_Rising_2D_OMP.f90:174-185) evaluated by the
init kernel on the GPU, for one or several bubbles.

* One bubble: the same mesh as the host set-up (amrlbm_setup_rising_bubble,
  which matches the oracle bit for bit), with fields equal to the host's up to
  the last bits of tanh (the device's tanh against the host libm's).
* Several bubbles: each interface is resolved at the finest level, the bubble
  area and centroid are those of the union of discs, and the run steps on.
* Bad arguments are rejected before anything is set up.
"""
import math

import numpy as np
import pytest

from conftest import mesh_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm as mod
    mod.lib()
    return mod


@pytest.mark.parametrize("name", ["tiny", "cfg1", "cfg2"])
def test_device_profile_matches_host_setup(amrlbm, workloads, name):
    w = workloads[name]
    print(f"[setup] {name}: device set-up of one bubble against the host set-up", flush=True)
    with amrlbm.Mesh(w) as a, amrlbm.Mesh(w) as b:
        a.setup_rising_bubble()
        b.setup_bubbles([w.radius], [w.x0], [w.y0])
        ra, fa, ka, xa = mesh_state(a)
        rb, fb, kb, xb = mesh_state(b)
        assert np.array_equal(ra, rb) and list(fa) == list(fb) and ka == kb
        worst, nbit, ncell = 0.0, 0, 0
        for n in xa:
            for f in ("h", "g", "C", "P", "Ux", "Uy"):
                x, y = xb[n][f], xa[n][f]
                if f in ("h", "g"):
                    x, y = x[:, 1:-1, 1:-1], y[:, 1:-1, 1:-1]
                d = float(np.max(np.abs(x - y) / np.maximum(np.abs(y), 1e-3)))
                worst = max(worst, d)
                nbit += int(np.count_nonzero(x != y))
                ncell += x.size
        print(f"[setup] {name}: {len(xa)} leaves, {nbit} of {ncell} values differ in the last bits, "
              f"max rel diff {worst:.2e}")
        assert worst <= 1e-12
        da, db = a.output(), b.output()
        assert np.allclose(da, db, rtol=1e-12, atol=1e-15)


def _leaf_at(mesh, x, y, L0):
    """Level of the leaf covering finest-grid point (x, y)."""
    rows, _ = mesh.topology()
    for r in rows:
        if r[0] >= 0 and r[5]:
            size = 16 * int(r[1])
            if r[2] * size <= x < (r[2] + 1) * size and r[3] * size <= y < (r[3] + 1) * size:
                return int(r[0])
    raise AssertionError(f"no leaf covers ({x}, {y})")


@pytest.mark.parametrize("precision", [64, 32])
def test_two_bubbles(amrlbm, workloads, precision):
    w = workloads["cfg2"]
    R = [48.0, 56.0]
    X = [352.0, 672.0]
    Y = [400.0, 560.0]
    with amrlbm.Mesh(w, precision=precision) as m:
        m.setup_bubbles(R, X, Y)
        c = m.counts()
        print(f"[setup] two bubbles, FP{precision}: {c}", flush=True)
        assert c["levelmax"] == w.max_levels
        for r, x, y in zip(R, X, Y):
            for a in np.linspace(0, 2 * math.pi, 16, endpoint=False):
                px, py = x + r * math.cos(a), y + r * math.sin(a)
                assert _leaf_at(m, px, py, w.L0) == w.max_levels, (r, x, y, a)
        d = m.output()
        area = math.pi * (R[0] ** 2 + R[1] ** 2)
        cx = (R[0] ** 2 * X[0] + R[1] ** 2 * X[1]) / (R[0] ** 2 + R[1] ** 2)
        cy = (R[0] ** 2 * Y[0] + R[1] ** 2 * Y[1]) / (R[0] ** 2 + R[1] ** 2)
        print(f"[setup] area {d[4]:.2f} (discs {area:.2f}), centroid ({d[5]:.3f}, {d[6]:.3f}) "
              f"(discs {cx:.3f}, {cy:.3f})")
        assert abs(d[4] - area) <= 5e-3 * area
        assert abs(d[5] - cx) <= 0.5 and abs(d[6] - cy) <= 0.5
        mass0 = d[3]
        m.step(100, 10)
        d = m.output()
        print(f"[setup] after 100 steps: {m.counts()}, rise velocity {d[7]:.3e}, mass drift {d[3] / mass0 - 1:.2e}")
        assert d[8] == 0
        assert abs(d[3] / mass0 - 1) <= 1e-6
        assert d[7] > 0   # the light bubbles rise


def test_bad_bubble_arguments(amrlbm, workloads):
    w = workloads["tiny"]
    with amrlbm.Mesh(w) as m:
        for args in (([], [], []), ([-1.0], [8.0], [8.0]), ([0.0], [8.0], [8.0])):
            with pytest.raises(amrlbm.AmrLbmError):
                m.setup_bubbles(*args)
