"""Ensembles of independent instances of uneven size, level range and fluid
in one mesh (amrlbm_setup_ensemble): every step launch covers the leaves of
all instances, each kernel takes its block's instance's fluid constants and
each regrid its instance's level range. Walls separate the instances, so
each must evolve as it does alone. Every leaf of every instance is compared
with the leaf at the same (level, I, J) of a one-domain FP64 run of that
instance (which matches the oracle bit for bit, test_gpu_parity.py), across
regrids that change one instance's mesh and not another's: the FP64
ensemble bitwise, the FP32 ensemble within the FP32 path's bounds against
FP64 (those of test_gpu_fp32.py) and bitwise against the instance run alone
in FP32 (the FP32 cell functions fuse multiply-adds per source expression,
so the ensemble's kernels compute the same bits). The 64-instance ensemble
drawn with seed 2024 runs as drawn, and two of its instances whose fluids
are not the reference's run against the oracle with their own constants."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm as mod
    mod.lib()
    return mod


def _members():
    from amrlbm.ensemble import member
    return (   # name, L0, levels, D, Eo, density ratio, viscosity ratio
        member("a", 64, 2, 20.0, 10.0, 1000.0, 100.0),
        member("b", 128, 3, 32.0, 40.0, 100.0, 10.0),
        member("c", 256, 2, 48.0, 5.0, 10.0, 1.0),
        member("d", 128, 2, 24.0, 80.0, 500.0, 50.0),
        member("e", 256, 4, 40.0, 20.0, 1000.0, 10.0),
    )


def _leaves(mesh, ensemble):
    """{(instance, level, I, J in the instance's frame): id} of the leaves."""
    from amrlbm.ensemble import instance_of
    rows, _ = mesh.topology()
    out = {}
    for n in range(1, rows.shape[0] + 1):
        r = rows[n - 1]
        if r[0] >= 0 and r[5]:
            L, I, J = int(r[0]), int(r[2]), int(r[3])
            q, Jl = instance_of(L, J) if ensemble else (0, J)
            out[(q, L, I, Jl)] = n
    return out


# the FP32 path's bounds against FP64 (test_gpu_fp32.py)
MASS_REL, RISE_REL, CENTROID_ABS = 1e-6, 5e-3, 0.05
FIELD_ABS = {"h": 1e-3, "g": 5e-6, "C": 2e-3, "P": 5e-6, "Ux": 3e-3, "Uy": 3e-3}


@pytest.mark.parametrize("precision", [64, 32])
def test_instances_evolve_as_alone(amrlbm, precision):
    from amrlbm.ensemble import Ensemble
    members = _members()
    ens = Ensemble("t5", members, block_cap=4096)
    steps, k = 300, 10
    print(f"[ensemble] FP{precision}: {len(members)} instances, {steps} steps, regrid every {k}", flush=True)
    with amrlbm.Mesh(ens, precision=precision) as e:
        e.setup_rising_bubble()
        e.step(steps, k)
        le = _leaves(e, True)
        de = e.output_instances()
        nbits = 0
        for q, w in enumerate(members):
            with amrlbm.Mesh(w, precision=64) as a:
                a.setup_rising_bubble()
                a.step(steps, k)
                la = {key[1:]: n for key, n in _leaves(a, False).items()}   # (level, I, J): id
                da = a.output()
                mine = {key[1:] for key in le if key[0] == q}
                if precision == 64:
                    assert mine == set(la), f"instance {q}: the leaf sets differ"
                    for (L, I, J), n in la.items():
                        ref = a.get_block(n)
                        got = e.get_block(le[(q, L, I, J)])
                        for f in ("C", "P", "Ux", "Uy"):
                            assert np.array_equal(got[f], ref[f]), (q, L, I, J, f)
                        for f in ("h", "g"):   # interiors (the fused stream leaves the previous ghost ring)
                            assert np.array_equal(got[f][:, 1:-1, 1:-1], ref[f][:, 1:-1, 1:-1]), (q, L, I, J, f)
                        nbits += 1
                    # the same cells summed in another order: FP64 rounding only
                    np.testing.assert_allclose(de[q], da, rtol=1e-12, atol=1e-15)
                    print(f"[ensemble] instance {w.name}: L0 {w.L0}, levels {w.min_levels}..{w.max_levels}, "
                          f"{len(la)} leaves, tau_h {w.tauh:.4f}, tau_l {w.taul:.4f}, rise velocity {da[7]:.3e}")
                    continue
                common = [key for key in la if key in mine]
                assert len(common) >= 0.95 * len(la), f"instance {q}: the meshes differ"
                err = {f: 0.0 for f in FIELD_ABS}
                for key in common:
                    ref = a.get_block(la[key])
                    got = e.get_block(le[(q,) + key])
                    for f in err:
                        x, y = (got[f], ref[f]) if f not in ("h", "g") else (got[f][:, 1:-1, 1:-1], ref[f][:, 1:-1, 1:-1])
                        err[f] = max(err[f], float(np.max(np.abs(x - y))))
                mass = abs(de[q][3] - da[3]) / abs(da[3])
                rise = abs(de[q][7] - da[7]) / abs(da[7])
                cen = max(abs(de[q][5] - da[5]), abs(de[q][6] - da[6]))
                print(f"[ensemble] FP32 instance {w.name} vs FP64 alone ({len(common)}/{len(la)} leaves): "
                      + ", ".join(f"{f} {v:.1e}" for f, v in err.items())
                      + f"; mass {mass:.1e}, rise velocity {rise:.1e}, centroid {cen:.1e} cells")
                for f, v in err.items():
                    assert v <= FIELD_ABS[f], (q, f, v)
                assert mass <= MASS_REL and rise <= RISE_REL and cen <= CENTROID_ABS
                with amrlbm.Mesh(w, precision=32) as a32:   # the same instance alone, FP32: the same bits
                    a32.setup_rising_bubble()
                    a32.step(steps, k)
                    l32 = {key[1:]: n for key, n in _leaves(a32, False).items()}
                    assert mine == set(l32), f"instance {q}: the FP32 leaf sets differ"
                    for (L, I, J), n in l32.items():
                        ref = a32.get_block(n)
                        got = e.get_block(le[(q, L, I, J)])
                        for f in ("C", "P", "Ux", "Uy"):
                            assert np.array_equal(got[f], ref[f]), (q, L, I, J, f)
                        for f in ("h", "g"):
                            assert np.array_equal(got[f][:, 1:-1, 1:-1], ref[f][:, 1:-1, 1:-1]), (q, L, I, J, f)
                    print(f"[ensemble] FP32 instance {w.name}: {len(l32)} leaves bitwise equal to its FP32 run alone")
        if precision == 64:
            print(f"[ensemble] FP64: {nbits} leaves bitwise equal to the instances run alone")


def test_drawn_ensemble_sets_up_and_steps(amrlbm):
    """Eight instances drawn from a seed (the batched-ensemble generator):
    every instance has its level range at set-up and stays finite over 500
    FP32 steps, with its bubble rising (positive mean velocity of the gas)."""
    from amrlbm.ensemble import ensemble, instance_of
    ens = ensemble(8, seed=7, block_cap=8192)
    with amrlbm.Mesh(ens, precision=32) as e:
        e.setup_rising_bubble()
        rows, _ = e.topology()
        for q, w in enumerate(ens.members):
            levels = {int(r[0]) for r in rows if r[0] >= 0 and r[5] and instance_of(int(r[0]), int(r[3]))[0] == q}
            assert min(levels) == w.min_levels and max(levels) == w.max_levels, (q, levels)
        d0 = e.output_instances()
        e.step(500, 10)
        d1 = e.output_instances()
        for q, w in enumerate(ens.members):
            print(f"[ensemble] drawn {q}: L0 {w.L0}, levels {w.min_levels}..{w.max_levels}, D {w.D:.1f}, "
                  f"Eo {w.Bo:.1f}, rho ratio {1 / w.rhol:.1f}, mu ratio {w.visc_ratio:.2f}: "
                  f"centroid y {d0[q, 6]:.3f} -> {d1[q, 6]:.3f}, rise velocity {d1[q, 7]:.3e}")
            assert abs(d1[q, 3] - d0[q, 3]) <= 1e-4 * abs(d0[q, 3])
            assert d1[q, 7] > 0.0


# ---- the 64-instance ensemble drawn with seed 2024 ---------------------------
DRAWN_N, DRAWN_SEED, DRAWN_MIN_STEPS, DRAWN_MAX_STEPS, DRAWN_K = 64, 2024, 500, 5000, 10


def _largest(members, n=3):
    """The n instances with the most finest cells and levels (ties: larger bubble)."""
    order = sorted(range(len(members)),
                   key=lambda q: (members[q].L0, members[q].max_levels - members[q].min_levels, members[q].D))
    return order[-n:]


def test_drawn_64_instances_fp64_and_fp32(amrlbm):
    """The batched ensemble as drawn (64 instances, seed 2024: domains up to
    1024 x 2048, 2-4 levels, density ratios 10-1000, viscosity ratios 1-100),
    FP64 and FP32 with the regrid every 10 steps, for at least 500 steps and
    on, 250 at a time, until a mesh has changed: every instance stays finite
    with its bubble rising, the three largest instances equal their one-domain
    FP64 runs bit for bit, and every instance's FP32 mass and rise velocity
    stay within 1e-6 and 0.5 % of its FP64 run."""
    from amrlbm.ensemble import ensemble
    ens = ensemble(DRAWN_N, seed=DRAWN_SEED)
    members = ens.members
    out, leaves = {}, {}
    steps = None
    for precision in (64, 32):
        print(f"[drawn] FP{precision}: {DRAWN_N} instances (seed {DRAWN_SEED}), {ens.block_cap} block slots, "
              f"regrid every {DRAWN_K}", flush=True)
        with amrlbm.Mesh(ens, precision=precision) as e:
            e.setup_rising_bubble()
            d0 = e.output_instances()
            before = _leaves(e, True)
            if steps is None:   # FP64 sets the length
                e.step(DRAWN_MIN_STEPS, DRAWN_K)
                steps = DRAWN_MIN_STEPS
                while e.regrid_stats()["changes"] == 0 and steps < DRAWN_MAX_STEPS:
                    e.step(250, DRAWN_K)
                    steps += 250
            else:
                e.step(steps, DRAWN_K)
            print(f"[drawn] FP{precision}: {steps} steps, {e.regrid_stats()['changes']} mesh changes", flush=True)
            d1 = e.output_instances()
            after = _leaves(e, True)
            out[precision] = d1
            changed = [q for q in range(DRAWN_N)
                       if {k for k in before if k[0] == q} != {k for k in after if k[0] == q}]
            print(f"[drawn] FP{precision}: {len(after)} leaves; instances whose mesh changed: {changed}", flush=True)
            assert changed, "no instance's mesh changed"
            assert np.all(np.isfinite(d1)) and np.all(d1[:, 8] == 0), "non-finite instance"
            assert np.all(d1[:, 7] > 0.0), np.nonzero(d1[:, 7] <= 0.0)
            dy = d1[:, 6] - d0[:, 6]
            print(f"[drawn] FP{precision}: rise velocity {d1[:, 7].min():.2e}..{d1[:, 7].max():.2e}; centroid "
                  f"moved {dy.min():+.3f}..{dy.max():+.3f} cells (lowest: instance {int(np.argmin(dy))})", flush=True)
            if precision == 64:
                leaves = after
                for q in _largest(members):
                    w = members[q]
                    with amrlbm.Mesh(w, precision=64) as a:
                        a.setup_rising_bubble()
                        a.step(steps, DRAWN_K)
                        la = {key[1:]: n for key, n in _leaves(a, False).items()}
                        mine = {key[1:] for key in leaves if key[0] == q}
                        assert mine == set(la), f"instance {q}: the leaf sets differ"
                        for (L, I, J), n in la.items():
                            ref = a.get_block(n)
                            got = e.get_block(leaves[(q, L, I, J)])
                            for f in ("C", "P", "Ux", "Uy"):
                                assert np.array_equal(got[f], ref[f]), (q, L, I, J, f)
                            for f in ("h", "g"):
                                assert np.array_equal(got[f][:, 1:-1, 1:-1], ref[f][:, 1:-1, 1:-1]), (q, L, I, J, f)
                        print(f"[drawn] instance {q} (L0 {w.L0}, levels {w.min_levels}..{w.max_levels}, D {w.D:.1f}): "
                              f"{len(la)} leaves bitwise equal to its run alone", flush=True)
    a, b = out[64], out[32]
    mass = np.abs(b[:, 3] - a[:, 3]) / np.abs(a[:, 3])
    rise = np.abs(b[:, 7] - a[:, 7]) / np.abs(a[:, 7])
    worst = int(np.argmax(rise))
    print(f"[drawn] FP32 vs FP64 per instance: mass max {mass.max():.2e} (median {np.median(mass):.2e}), "
          f"rise velocity max {rise.max():.2e} (instance {worst}, median {np.median(rise):.2e})")
    assert mass.max() <= MASS_REL, int(np.argmax(mass))
    assert rise.max() <= RISE_REL, worst


@pytest.mark.parametrize("q", [1, 35])
def test_drawn_instance_fluid_matches_oracle(amrlbm, oracle_mod, q):
    """A drawn instance whose fluid is not the reference's (instance 1: density
    ratio 84, viscosity ratio 47, tau_l 0.431; instance 35: tau_h 0.527), run
    alone, against the oracle with that instance's constants: 20 steps with
    the regrid every 10, every value bit for bit (the ensemble's instances
    equal their one-domain runs, above)."""
    from amrlbm.ensemble import ensemble
    from conftest import compare_state
    w = ensemble(DRAWN_N, seed=DRAWN_SEED).members[q]
    print(f"[drawn-oracle] instance {q}: L0 {w.L0}, levels {w.min_levels}..{w.max_levels}, D {w.D:.1f}, "
          f"rho ratio {1 / w.rhol:.1f}, mu ratio {w.visc_ratio:.2f}, tau_h {w.tauh:.4f}, tau_l {w.taul:.4f}", flush=True)
    o = oracle_mod.Oracle(w)
    o.setup()
    o.run(20, 10)
    with amrlbm.Mesh(w) as m:
        m.setup_rising_bubble()
        m.step(20, 10)
        worst, nbit = compare_state(m, o.snapshot(), pop_ghosts=False)
        d = m.output()
    print(f"[drawn-oracle] instance {q}: max rel {worst:.2e}, bit-differing values {nbit}")
    np.testing.assert_allclose(d[:8], o.diagnostics()[:8], rtol=1e-10)
