"""The FP32 production form (amrlbm_create_fp32) against the FP64 path, which
matches the oracle bit for bit (test_gpu_parity.py). Bounds: bubble mass
within 1e-6 relative and rise velocity within 0.5 % of FP64 after the run;
populations, C and U, P within the absolute bounds below wherever both runs
have the same mesh. Every error is printed."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MASS_REL = 1e-6          # total mass sum(rho dx^2), relative
RISE_REL = 5e-3          # bubble rise velocity, relative
CENTROID_ABS = 0.05      # bubble centroid, finest cells
# accumulated FP32 rounding after the runs below, with margin (measured at
# cfg2, 400 steps: h 1.4e-4, g 2.6e-7, C 3.2e-4, P 4.2e-7, Ux 4.2e-4, Uy 5.1e-4
# against a rise velocity of 7.4e-2); reported, not a target
FIELD_ABS = {"h": 1e-3, "g": 5e-6, "C": 2e-3, "P": 5e-6, "Ux": 3e-3, "Uy": 3e-3}


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm as mod
    mod.lib()
    return mod


def _run(amrlbm, w, precision, steps, k):
    m = amrlbm.Mesh(w, precision=precision)
    m.setup_rising_bubble()
    m.step(steps, k)
    return m


def test_fp32_setup_matches_fp64(amrlbm, workloads):
    """The set-up is computed in FP64 and stored rounded: same mesh, values
    within FP32 rounding of the FP64 ones (population interiors; the ghost
    ring holds nothing until the first fill)."""
    w = workloads["cfg1"]
    with amrlbm.Mesh(w) as a, amrlbm.Mesh(w, precision=32) as b:
        a.setup_rising_bubble()
        b.setup_rising_bubble()
        ra, _ = a.topology()
        rb, _ = b.topology()
        assert np.array_equal(ra, rb)
        worst = 0.0
        for n in a.leaves():
            x, y = a.get_block(n), b.get_block(n)
            for f in ("h", "g", "C", "P", "Ux", "Uy"):
                u, v = x[f], y[f]
                if f in ("h", "g"):   # the ghost ring is written by the first fill
                    u, v = u[:, 1:-1, 1:-1], v[:, 1:-1, 1:-1]
                d = float(np.max(np.abs(u - v) / np.maximum(np.abs(u), 1.0)))
                worst = max(worst, d)
        print(f"[fp32] set-up: max relative rounding {worst:.2e}")
        assert worst <= 1.2e-7


@pytest.mark.parametrize("name,steps,k", [("cfg1", 1000, 5), ("cfg2", 400, 10)])
def test_fp32_run_against_fp64(amrlbm, workloads, name, steps, k):
    w = workloads[name]
    print(f"[fp32] {name}: {steps} steps, regrid every {k}", flush=True)
    a = _run(amrlbm, w, 64, steps, k)
    b = _run(amrlbm, w, 32, steps, k)
    try:
        da, db = a.output(), b.output()
        mass = abs(db[3] - da[3]) / abs(da[3])
        rise = abs(db[7] - da[7]) / abs(da[7])
        cen = max(abs(db[5] - da[5]), abs(db[6] - da[6]))
        print(f"[fp32] {name}: mass rel {mass:.2e}, rise velocity rel {rise:.2e} "
              f"({da[7]:.6e} vs {db[7]:.6e}), centroid {cen:.2e} cells", flush=True)
        ra, _ = a.topology()
        rb, _ = b.topology()
        same_mesh = ra.shape == rb.shape and np.array_equal(ra[:, :7], rb[:, :7])
        print(f"[fp32] {name}: same mesh as FP64: {same_mesh} "
              f"(leaves {a.counts()['nleaf']} vs {b.counts()['nleaf']})")
        if same_mesh:
            err = {f: 0.0 for f in FIELD_ABS}
            for n in a.leaves():
                x, y = a.get_block(n), b.get_block(n)
                for f in FIELD_ABS:
                    u, v = x[f], y[f]
                    if f in ("h", "g"):
                        u, v = u[:, 1:-1, 1:-1], v[:, 1:-1, 1:-1]
                    if f == "C":
                        u, v = u[2:-2, 2:-2], v[2:-2, 2:-2]
                    err[f] = max(err[f], float(np.max(np.abs(u - v))))
            print("[fp32] " + name + ": max abs error " + ", ".join(f"{f} {e:.2e}" for f, e in err.items()))
            for f, e in err.items():
                assert e <= FIELD_ABS[f], (f, e)
        assert mass <= MASS_REL
        assert rise <= RISE_REL
        assert cen <= CENTROID_ABS
        assert db[8] == 0
    finally:
        a.close()
        b.close()


def test_fp32_replicas_are_identical(amrlbm, workloads):
    """Two copies of cfg1 in one FP32 mesh evolve bit for bit alike."""
    w = workloads["cfg1x2"]
    with amrlbm.Mesh(w, precision=32) as m:
        m.setup_rising_bubble()
        m.step(200, 5)
        rows, _ = m.topology()
        pos = {}
        for n in range(1, rows.shape[0] + 1):
            r = rows[n - 1]
            if r[0] >= 0 and r[5]:
                pos[(int(r[0]), int(r[2]), int(r[3]))] = n
        roots = w.nroots // w.instances
        count = 0
        for (lev, I, J), n in pos.items():
            if J >= (roots << lev):
                continue
            x, y = m.get_block(n), m.get_block(pos[(lev, I, J + (roots << lev))])
            for f in ("C", "P", "Ux", "Uy"):
                assert np.array_equal(x[f], y[f]), (lev, I, J, f)
            for f in ("h", "g"):
                assert np.array_equal(x[f][:, 1:-1, 1:-1], y[f][:, 1:-1, 1:-1]), (lev, I, J, f)
            count += 1
        print(f"[fp32] cfg1x2: {count} leaf pairs bitwise equal")
        assert count == m.counts()["nleaf"] // 2


def test_fp32_long_cfg2_run_within_targets(amrlbm, workloads):
    """6,000 steps of the 1024x2048 bubble (regrid every 10, mesh changes on
    the way): mass within 1e-6 and rise velocity within 0.5 % of FP64, every
    2,000 steps. Beside them, FP64 from a set-up whose bubble radius is 1e-7
    larger gives the flow's own sensitivity, printed for comparison."""
    w = workloads["cfg2"]
    with amrlbm.Mesh(w) as a, amrlbm.Mesh(w, precision=32) as b, amrlbm.Mesh(w) as c:
        a.setup_rising_bubble()
        b.setup_rising_bubble()
        c.setup_rising_bubble(radius=w.radius * (1.0 + 1e-7))
        for chk in range(1, 4):
            for m in (a, b, c):
                m.step(2000, 10)
            da, db, dc = a.output(), b.output(), c.output()
            mass = abs(db[3] - da[3]) / abs(da[3])
            rise = abs(db[7] - da[7]) / abs(da[7])
            print(f"[fp32] cfg2 step {2000 * chk}: FP32 mass rel {mass:.2e}, rise velocity rel {rise:.2e}; "
                  f"FP64 perturbed: mass rel {abs(dc[3] - da[3]) / abs(da[3]):.2e}, rise velocity rel "
                  f"{abs(dc[7] - da[7]) / abs(da[7]):.2e}", flush=True)
            assert mass <= MASS_REL
            assert rise <= RISE_REL
            assert db[8] == 0
        assert a.regrid_stats()["changes"] >= 2
