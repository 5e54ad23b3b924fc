"""The 32-bubble case (amrlbm.multibubble): 2048 x 4096 finest-equivalent,
levels 3..7, 32 bubbles set up on the device. At set-up every bubble is a
gas region at its centre with its tanh profile's area (2%); after the same
steps the FP32 path keeps the FP64 path's total mass (1e-6) and per-bubble
centroids."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 300


def _profile_area(r, w):
    """The weighted area sum((1 - C) dx^2) over C < 1/2 of one bubble whose
    set-up profile is C = 1/2 + 1/2 tanh(2 (d - R) / W) at distance d from
    its centre: the integral of (1 - C) 2 pi d over 0 <= d <= R, which falls
    short of pi R^2 by about 2 pi R W ln2 / 4 (W ln2 / 4 = 0.87 cells at W = 5)."""
    d = np.linspace(0.0, r, 20001)
    f = 0.5 * (1.0 + np.tanh(2.0 * (r - d) / w)) * 2.0 * math.pi * d
    return float(np.sum(0.5 * (f[1:] + f[:-1]) * np.diff(d)))


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm as mod
    mod.lib()
    return mod


def test_multibubble_setup_and_fp32_vs_fp64(amrlbm):
    from amrlbm import multibubble as mb
    bub = mb.bubbles()
    assert len(bub) == 32
    res = {}
    for prec in (64, 32):
        print(f"[multi] FP{prec}: set-up of {len(bub)} bubbles", flush=True)
        with amrlbm.Mesh(mb.WORKLOAD, precision=prec) as m:
            mb.setup(m, bub)
            c0 = m.counts()
            g0 = mb.gas_regions(m)
            print(f"[multi] FP{prec}: {c0['nleaf']} leaves, levels {c0['levelmin']}..{c0['levelmax']}, "
                  f"{len(g0)} gas regions", flush=True)
            assert c0["levelmax"] == mb.WORKLOAD.max_levels
            assert len(g0) == len(bub)
            for r, x, y in bub:   # each bubble is one region at its centre with its profile's area
                a, xc, yc = min(g0, key=lambda g: math.hypot(g[1] - x, g[2] - y))
                assert math.hypot(xc - x, yc - y) < 0.5, (r, x, y, xc, yc)
                a0 = _profile_area(r, mb.WORKLOAD.W)
                assert abs(a - a0) < 0.02 * a0, (r, a, a0)
            m.step(STEPS, 10)
            res[prec] = (m.output(), mb.gas_regions(m), m.counts())
    (d64, g64, c64), (d32, g32, c32) = res[64], res[32]
    mass = abs(d32[3] - d64[3]) / abs(d64[3])
    dist = max(mb.match(g64, g32), mb.match(g32, g64))
    print(f"[multi] after {STEPS} steps: leaves {c64['nleaf']} / {c32['nleaf']}, regions {len(g64)} / {len(g32)}, "
          f"mass rel {mass:.2e}, centroid max diff {dist:.2e} cells")
    assert d64[8] == 0 and d32[8] == 0
    assert mass <= 1e-6
    assert len(g64) == len(g32)
    assert dist <= 0.05


def test_multibubble_learned_cadence_fp32(amrlbm):
    """The learned regrid's cadence (full regrid every learned.FULL_EVERY
    steps, the plain criterion judging every step on the device, no network) on the 32 bubbles
    in FP32 over 3000 steps: the state stays finite and the mesh the same
    size as under the plain regrid every 10 steps. Regrids issued back to
    back (no step between them) refined every just-derefined block again:
    3731 leaves against 2804 in FP64, and the FP32 state turned non-finite
    near step 3000."""
    from amrlbm import learned
    from amrlbm import multibubble as mb
    bub = mb.bubbles()
    leaves = {}
    for name in ("learned", "plain"):
        print(f"[multi] FP32 {name} cadence: 3000 steps", flush=True)
        with amrlbm.Mesh(mb.WORKLOAD, precision=32) as m:
            mb.setup(m, bub)
            if name == "learned":
                st = learned.run(m, 3000, None)
                assert st["checks"] == 3000 and st["predicted_regrids"] == 3000 // learned.FULL_EVERY
            else:
                m.step(3000, 10)
            d = m.output(allow_nonfinite=True)
            leaves[name] = m.counts()["nleaf"]
            print(f"[multi] FP32 {name}: {leaves[name]} leaves, mass {d[3]:.9g}, non-finite {d[8]}")
            assert d[8] == 0
    assert abs(leaves["learned"] - leaves["plain"]) <= 0.02 * leaves["plain"]
