"""Parity of the HIP library (through its C-ABI) with the oracle.

FP64 tolerance: every array value within 1e-12 relative (|a-b| <= 1e-12
max(|b|,1)); topology, free list and KK bit-exact. The tests also report how
many values differ in any bit (the build aims at zero).
"""
import numpy as np
import pytest

from conftest import compare_state, load_into_mesh

pytestmark = pytest.mark.gpu

PHASES = ["c2", "collide", "hg", "stream", "macro_h", "c", "macro_g", "regrid"]


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm
    amrlbm.lib()
    return amrlbm


def _oracle_after(oracle_mod, w, steps, regrid_every=1):
    o = oracle_mod.Oracle(w)
    o.setup()
    if steps:
        o.run(steps, regrid_every)
    return o


@pytest.mark.parametrize("name", ["tiny", "cfg1", "cfg2"])
def test_setup_matches_oracle(amrlbm, oracle_mod, workloads, name):
    w = workloads[name]
    print(f"[setup] {name}", flush=True)
    o = _oracle_after(oracle_mod, w, 0)
    with amrlbm.Mesh(w) as m:
        m.setup_rising_bubble()
        worst, nbit = compare_state(m, o.snapshot())
    print(f"[setup] {name}: max rel {worst:.2e}, bit-differing values {nbit}")


@pytest.mark.parametrize("name,steps", [("tiny", 3), ("cfg1", 5)])
def test_each_phase_matches_oracle(amrlbm, oracle_mod, workloads, name, steps):
    """Every phase of Main_Loop from the same state: full arrays, ghosts included."""
    w = workloads[name]
    o = _oracle_after(oracle_mod, w, steps)
    with amrlbm.Mesh(w) as m:
        load_into_mesh(m, o.snapshot())
        o.phase("levels")
        m.determine_min_max_levels()
        for ph in PHASES:
            print(f"[phase] {name} step {steps} {ph}", flush=True)
            if ph == "regrid":
                o.regrid()
            else:
                o.phase(ph)
            m.phase(ph)
            worst, nbit = compare_state(m, o.snapshot())
            print(f"[phase] {name} {ph}: max rel {worst:.2e}, bit-differing values {nbit}")


@pytest.mark.parametrize("name,steps,k", [("tiny", 20, 1), ("cfg1", 60, 10), ("cfg2", 12, 4)])
def test_fused_steps_match_oracle(amrlbm, oracle_mod, workloads, name, steps, k):
    """amrlbm_step (fused per-level kernels in a HIP graph) vs the oracle's
    Main_Loop. The fused stream leaves the population ghost ring of the
    previous buffer in place, so populations are compared on interiors."""
    w = workloads[name]
    print(f"[steps] {name} {steps} steps regrid every {k}", flush=True)
    o = _oracle_after(oracle_mod, w, steps, k)
    with amrlbm.Mesh(w) as m:
        m.setup_rising_bubble()
        m.step(steps, k)
        worst, nbit = compare_state(m, o.snapshot(), pop_ghosts=False)
        print(f"[steps] {name}: max rel {worst:.2e}, bit-differing values {nbit}")
        d_gpu = m.output()
        d_orc = o.diagnostics()
        rel = np.abs(d_gpu - d_orc) / np.maximum(np.abs(d_orc), 1e-300)
        print("[steps] diagnostics rel err", rel)
        assert np.all(rel[:8] <= 1e-10), rel
        assert d_gpu[8] == 0


def test_regrid_walk_matches_oracle(amrlbm, oracle_mod, workloads):
    """Move the analytic phase field between regrids: refine, derefine and
    free-list reuse on every call, with prolong/restrict data."""
    w = workloads["cfg1"]
    o = _oracle_after(oracle_mod, w, 0)
    with amrlbm.Mesh(w) as m:
        m.setup_rising_bubble()
        for k in range(1, 15):
            print(f"[walk] move {k}", flush=True)
            o.set_bubble_c(w.L0 // 2 + 3.0 * k, w.L0 // 2 + 5.0 * k)
            snap = o.snapshot()
            for n in snap.leaves:
                m.put_block(n, C=snap.blocks[n].C)
            o.phase("levels")
            m.determine_min_max_levels()
            o.regrid()
            m.refine_derefine_mesh()
            worst, nbit = compare_state(m, o.snapshot())
            c = m.counts()
            print(f"[walk] {k}: nblks {c['nblks']} nleaf {c['nleaf']} kk {c['kk']} "
                  f"max rel {worst:.2e} bitdiff {nbit}")


def test_per_block_kernels(amrlbm, oracle_mod, workloads):
    """Per-block entry points against the oracle's per-block routines."""
    import ctypes
    w = workloads["cfg1"]
    o = _oracle_after(oracle_mod, w, 3)
    o.phase("levels")
    o.phase("c2")
    snap = o.snapshot()
    L = oracle_mod.lib()
    P = oracle_mod.make_params(w)
    for n in snap.leaves[:6]:
        b = snap.blocks[n]
        print(f"[block] {n} dx {b.dx}", flush=True)
        h1, g1 = np.asfortranarray(b.h.copy()), np.asfortranarray(b.g.copy())
        h2, g2 = np.asfortranarray(b.h.copy()), np.asfortranarray(b.g.copy())
        C, Pp, Ux, Uy = (np.asfortranarray(a.copy()) for a in (b.C, b.P, b.Ux, b.Uy))
        amrlbm.collision_block(w, b.dx, C, Pp, Ux, Uy, h1, g1)
        L.orc_collision_block(ctypes.byref(P), b.dx, oracle_mod.dptr(C), oracle_mod.dptr(Pp),
                              oracle_mod.dptr(Ux), oracle_mod.dptr(Uy), oracle_mod.dptr(h2), oracle_mod.dptr(g2))
        np.testing.assert_allclose(h1, h2, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(g1, g2, rtol=1e-12, atol=1e-15)
        outs = [np.zeros((16, 16), order="F") for _ in range(6)]
        amrlbm.macro_g_block(w, b.dx, g1, C, *outs[:3])
        L.orc_macro_g_block(ctypes.byref(P), b.dx, oracle_mod.dptr(g1), oracle_mod.dptr(C),
                            *[oracle_mod.dptr(a) for a in outs[3:]])
        for a, r in zip(outs[:3], outs[3:]):
            np.testing.assert_allclose(a, r, rtol=1e-12, atol=1e-15)
        for dx in (1, 2, 4, 8):
            f1 = np.asfortranarray(b.h.copy())
            f2 = np.asfortranarray(b.h.copy())
            amrlbm.stream_block(w, dx, f1)
            if dx == 1:
                L.orc_perfect_shift_block(ctypes.byref(P), oracle_mod.dptr(f2))
            else:
                L.orc_lax_wendroff_block(ctypes.byref(P), 1.0 / dx, oracle_mod.dptr(f2))
            np.testing.assert_allclose(f1, f2, rtol=1e-12, atol=1e-15)


def test_deterministic(amrlbm, workloads):
    """Two runs of the same steps give bitwise equal states."""
    w = workloads["cfg1"]
    states = []
    for _ in range(2):
        with amrlbm.Mesh(w) as m:
            m.setup_rising_bubble()
            m.step(30, 10)
            states.append([m.get_block(n) for n in m.leaves()] + [m.output()])
    for a, b in zip(states[0][:-1], states[1][:-1]):
        for f in a:
            assert np.array_equal(a[f], b[f])
    assert np.array_equal(states[0][-1], states[1][-1])


@pytest.mark.parametrize("name,steps,k,min_changes", [("tiny", 300, 1, 0), ("cfg1", 400, 5, 0),
                                                     ("cfg2", 1000, 10, 1), ("cfg2", 2400, 7, 1)])
def test_device_regrid_decision_matches_host_regrid(amrlbm, workloads, name, steps, k, min_changes):
    """step() decides each regrid on the device and launches the next segment
    before the host knows the answer; profile_steps() runs the host's
    Refine_Derefine_Mesh after every segment with the device idle. Both must
    give the same topology, flags, free list and bitwise equal fields, and the
    runs must include regrids that change the mesh (taken-back segments)."""
    from conftest import mesh_state
    w = workloads[name]
    print(f"[spec] {name} {steps} steps regrid every {k}", flush=True)
    with amrlbm.Mesh(w) as a, amrlbm.Mesh(w) as b:
        a.setup_rising_bubble()
        b.setup_rising_bubble()
        a.step(steps, k)
        b.profile_steps(steps, k)
        st = a.regrid_stats()
        print(f"[spec] {name}: {st}", flush=True)
        assert st["checks"] == steps // k
        assert st["changes"] >= min_changes
        ra, fa, ka, xa = mesh_state(a)
        rb, fb, kb, xb = mesh_state(b)
        assert np.array_equal(ra, rb)
        assert list(fa) == list(fb) and ka == kb
        assert xa.keys() == xb.keys()
        for n in xa:
            for f in xa[n]:
                assert np.array_equal(xa[n][f], xb[n][f]), (n, f)
        assert np.array_equal(a.output(), b.output())


def test_errors_reported(amrlbm, workloads):
    """Argument errors come back as codes with a message, not crashes."""
    w = workloads["tiny"]
    with amrlbm.Mesh(w) as m:
        with pytest.raises(amrlbm.AmrLbmError) as e:
            m.fill_ghost_cells_c2()      # no mesh set up yet
        assert e.value.code == 6
        with pytest.raises(amrlbm.AmrLbmError) as e:
            m.step(1, 0)
        assert e.value.code in (2, 6)
