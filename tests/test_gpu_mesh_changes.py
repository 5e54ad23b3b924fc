"""FP64 parity of the HIP library with the reference build across mesh changes
during time stepping.

The golden sequences (tests/golden/reference_hashes.json, "rise") are runs of
the reference build whose mesh changes while it steps: refinement,
derefinement and refinements that reuse freed ids, each with prolong or
restrict of stepped fields and Err_Interface from the step's C fill. The
library writes snapshots in the reference harness's layout
(amrlbm_save_snapshot), so equal SHA-256 means bit-identical topology, flags,
free list, KK and every array, ghost rings included. Population ghost rings
are written to memory by the phase-by-phase fill (amrlbm_fill_ghost_cells_hg),
not by the fused amrlbm_step, so the comparisons after fused steps start at the
hg fill of a step run phase by phase.
"""
import hashlib
import json
import os
import tempfile

import numpy as np
import pytest

from conftest import compare_state

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RISE = json.load(open(os.path.join(GOLDEN, "reference_hashes.json")))["rise"]
PHASES = ["c2", "collide", "hg", "stream", "macro_h", "c", "macro_g"]


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm
    amrlbm.lib()
    return amrlbm


def snap_hash(m):
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
        path = f.name
    try:
        m.save_snapshot(path)
        return hashlib.sha256(open(path, "rb").read()).hexdigest()
    finally:
        os.unlink(path)


def step_by_phases(m, t, k, g=None):
    """Main_Loop step t phase by phase on the library; with g, the snapshot
    hashes from the hg fill on (p3..p8) against the reference build's."""
    m.set_step(t)
    m.determine_min_max_levels()
    for p, ph in enumerate(PHASES, start=1):
        m.phase(ph)
        if g is not None and p >= 3:
            assert snap_hash(m) == g["phases"][f"s{t}_p{p}"], f"step {t} after {ph}"
    if (t + 1) % k == 0:
        m.refine_derefine_mesh()
    if g is not None:
        assert snap_hash(m) == g["phases"][f"s{t}_p8"], f"step {t} after the regrid"
    m.set_step(t + 1)


@pytest.mark.parametrize("execution", [0, 1])
@pytest.mark.parametrize("name", ["tiny_rise", "cfg1_rise", "cfg2_regrid"])
def test_fused_steps_across_mesh_changes_equal_reference(amrlbm, workloads, name, execution):
    """amrlbm_step (fused kernels, regrid decided and flagged on the device,
    segments taken back when the mesh changes; the step graph or the
    persistent segment kernel) over the whole run, then its last step phase by
    phase: the final state's SHA-256 equals the reference build's."""
    g = RISE[name]
    w = workloads[g["workload"]]
    n, k = g["steps"], g["regrid_every"]
    print(f"[rise] {name}: {n} steps, regrid every {k}, execution {execution}, "
          f"mesh changes at {g['change_steps']}", flush=True)
    with amrlbm.Mesh(w) as m:
        m.setup_rising_bubble()
        m.set_tuning(0, 1, execution)
        assert snap_hash(m) == g["init"]
        m.step(n - 1, k)
        st = m.regrid_stats()
        print(f"[rise] {name}: {st}", flush=True)
        assert st["changes"] >= len(g["change_steps"]), st
        assert st["device_flag_changes"] == st["changes"], st   # every change flagged on the device
        step_by_phases(m, n - 1, k)
        m.determine_min_max_levels()
        assert snap_hash(m) == g["final"], "final state differs from the reference build"


@pytest.mark.parametrize("name", ["tiny_rise", "cfg1_rise", "cfg2_regrid"])
def test_steps_changing_the_mesh_equal_reference_phase_by_phase(amrlbm, workloads, name):
    """Fused steps up to each step whose regrid changes the mesh, that step
    phase by phase with every snapshot from the hg fill on against the
    reference build's (the regrid itself included), then on to the end."""
    g = RISE[name]
    w = workloads[g["workload"]]
    n, k = g["steps"], g["regrid_every"]
    with amrlbm.Mesh(w) as m:
        m.setup_rising_bubble()
        for t in g["change_steps"]:
            print(f"[rise] {name}: fused steps to {t}, then step {t} by phases", flush=True)
            c = m.counts()
            m.step(t - c["T"], k)
            before = m.counts()
            step_by_phases(m, t, k, g)
            after = m.counts()
            assert (before["nleaf"], before["kk"]) != (after["nleaf"], after["kk"]), (before, after)
        m.step(n - 1 - m.counts()["T"], k)
        step_by_phases(m, n - 1, k)
        m.determine_min_max_levels()
        assert snap_hash(m) == g["final"]


def test_fused_steps_across_mesh_changes_equal_oracle(amrlbm, oracle_mod, workloads):
    """The same run against the oracle's arrays, value by value (interiors of
    the populations, every other array with its rings), bit for bit."""
    g = RISE["cfg1_rise"]
    w = workloads[g["workload"]]
    o = oracle_mod.Oracle(w)
    o.setup()
    o.run(g["steps"], g["regrid_every"])
    with amrlbm.Mesh(w) as m:
        m.setup_rising_bubble()
        m.step(g["steps"], g["regrid_every"])
        worst, nbit = compare_state(m, o.snapshot(), pop_ghosts=False)
        assert nbit == 0
        d_gpu, d_orc = m.output(), o.diagnostics()
        rel = np.abs(d_gpu - d_orc) / np.maximum(np.abs(d_orc), 1e-300)
        print("[rise] diagnostics rel err", rel)
        assert np.all(rel[:8] <= 1e-12), rel


@pytest.mark.parametrize("case", ["cfg2-64", "cfg2-32", "ens64-32"])
def test_device_flags_equal_host_flags(amrlbm, workloads, monkeypatch, case):
    """Every mesh change of amrlbm_step takes its Refine and Deref flags from
    the device decision (k_regrid_check: the criterion, the 2:1 closure, the
    derefinement flags and their level-by-level check), and uploads only the
    table rows the regrid marked, and rebuilds only the ghost programs that
    read a block whose row changed. With AMRLBM_VERIFY_FLAGS set the library
    also flags each such state on the host (Err_Interface downloaded afresh),
    with AMRLBM_VERIFY_TABLES it rebuilds every table row after each upload,
    with AMRLBM_VERIFY_PROGRAMS every program, and it fails on any
    difference: cfg2 over 4,000 steps in FP64 and FP32,
    and the 64-instance ensemble (per-instance level ranges, about 45,000
    block slots) over 1,000 steps."""
    from amrlbm.ensemble import ensemble
    name, prec = case.rsplit("-", 1)
    w = ensemble(64, seed=2024) if name == "ens64" else workloads[name]
    steps = 1000 if name == "ens64" else 4000
    monkeypatch.setenv("AMRLBM_VERIFY_FLAGS", "1")
    monkeypatch.setenv("AMRLBM_VERIFY_TABLES", "1")
    monkeypatch.setenv("AMRLBM_VERIFY_PROGRAMS", "1")
    print(f"[flags] {case}: {steps} steps, regrid every 10", flush=True)
    with amrlbm.Mesh(w, precision=int(prec)) as m:
        m.setup_rising_bubble()
        m.step(steps, 10)
        st = m.regrid_stats()
    print(f"[flags] {case}: {st}", flush=True)
    assert st["changes"] >= 2, st
    assert st["device_flag_changes"] == st["changes"], st


@pytest.mark.parametrize("name", ["tiny_rise", "cfg1_rise"])
def test_table_rows_marked_on_every_regrid(amrlbm, workloads, monkeypatch, name):
    """Regrid after every step (the reference's cadence) on the fast-rising
    bubbles, with AMRLBM_VERIFY_TABLES and AMRLBM_VERIFY_PROGRAMS set: after
    each upload every table row and every ghost program equals a fresh build,
    so every row a refinement, derefinement or freed-id reuse changed was
    marked and uploaded, and every program reading it rebuilt."""
    g = RISE[name]
    monkeypatch.setenv("AMRLBM_VERIFY_TABLES", "1")
    monkeypatch.setenv("AMRLBM_VERIFY_PROGRAMS", "1")
    with amrlbm.Mesh(workloads[g["workload"]]) as m:
        m.setup_rising_bubble()
        m.step(g["steps"], 1)
        st = m.regrid_stats()
    print(f"[tables] {name}: {st}", flush=True)
    assert st["changes"] >= len(g["change_steps"]), st
