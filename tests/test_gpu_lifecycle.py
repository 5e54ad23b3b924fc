"""Meshes created and destroyed one after another in one process, FP64 and FP32
interleaved, with mesh changes inside each run. The library keeps per-mesh
pinned staging, events and device arenas, and per-thread build memos; each new
mesh must start clean and evolve bit for bit as the first one did, and no
error may be left behind for the next HIP call to report (a pinned buffer
freed twice did exactly that once)."""
import numpy as np
import pytest

from conftest import mesh_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm as mod
    mod.lib()
    return mod


def _run(amrlbm, w, precision, steps):
    with amrlbm.Mesh(w, precision=precision) as m:
        m.setup_rising_bubble()
        m.step(steps, 10)
        changes = m.regrid_stats()["changes"]
        return mesh_state(m), changes


def _same(a, b):
    ra, fa, ka, xa = a
    rb, fb, kb, xb = b
    assert np.array_equal(ra, rb) and list(fa) == list(fb) and ka == kb and xa.keys() == xb.keys()
    for n in xa:
        for f in ("C", "P", "Ux", "Uy"):
            assert np.array_equal(xa[n][f], xb[n][f]), (n, f)
        for f in ("h", "g"):
            assert np.array_equal(xa[n][f][:, 1:-1, 1:-1], xb[n][f][:, 1:-1, 1:-1]), (n, f)


def test_meshes_in_sequence_start_clean(amrlbm, workloads):
    w = workloads["cfg2"]
    steps = 1000
    first = {}
    for rep in range(3):
        for precision in (64, 32):
            print(f"[lifecycle] mesh {rep} FP{precision}: cfg2, {steps} steps", flush=True)
            st, changes = _run(amrlbm, w, precision, steps)
            assert changes >= 1   # the run goes through the regrid and upload path
            if precision not in first:
                first[precision] = st
            else:
                _same(st, first[precision])


def test_step_error_leaves_no_skip_word(amrlbm, workloads):
    """A regrid that fails inside amrlbm_step (MaxBlks exceeded: the reference
    STOPs, AMR_Rising_2D_OMP.f90:527-531) returns the error with the device
    decision cleared: later phase calls still do their work."""
    import dataclasses
    w = dataclasses.replace(workloads["tiny_rise"], block_cap=26)   # the set-up mesh has 26 blocks
    with amrlbm.Mesh(w) as m:
        m.setup_rising_bubble()
        assert m.counts()["nblks"] == 26
        with pytest.raises(amrlbm.AmrLbmError) as e:
            m.step(1200, 1)   # the first refinement, at step 1077, needs blocks 27..34
        assert e.value.code == 3
        c = m.counts()
        assert c["T"] == 1078 and c["nblks"] == 26, c
        n = m.leaves()[0]
        before = m.get_block(n)
        m.fill_ghost_cells_c2()
        m.collide()
        after = m.get_block(n)
        assert not np.array_equal(before["g"][:, 1:-1, 1:-1], after["g"][:, 1:-1, 1:-1])
        m.fill_ghost_cells_hg()
        m.stream()
        m.macroscopic_values_h()
        assert not np.array_equal(before["C"], m.get_block(n)["C"])


def test_set_topology_rejects_bad_ids(amrlbm, workloads):
    """A block table whose ids would lead outside the table (a child group past
    Nblks, a parent that does not list the block, a free-list id in use, KK
    not 4 per freed group) is rejected with an error and leaves no mesh."""
    w = workloads["cfg1"]
    with amrlbm.Mesh(w) as m:
        m.setup_rising_bubble()
        rows, free = m.topology()
        kk = m.counts()["kk"]
        m.set_topology(rows, kk, free)   # the mesh's own table is accepted
        parent = next(n for n in range(rows.shape[0]) if rows[n, 0] >= 0 and not rows[n, 5])
        leaf = next(n for n in range(rows.shape[0]) if rows[n, 5] and rows[n, 4] > 0)
        cases = []
        r = rows.copy(); r[parent, 9:13] = rows.shape[0] - 1 + np.arange(4); cases.append((r, kk, free))
        r = rows.copy(); r[leaf, 4] = leaf + 1; cases.append((r, kk, free))
        r = rows.copy(); r[leaf, 6] = 1; cases.append((r, kk, free))
        cases.append((rows, kk + 4, free))
        cases.append((rows, kk + 4, list(free) + [1]))
        for r, k, f in cases:
            with pytest.raises(amrlbm.AmrLbmError) as e:
                m.set_topology(r, k, f)
            assert e.value.code == 4, e.value
            with pytest.raises(amrlbm.AmrLbmError) as e:
                m.fill_ghost_cells_c2()   # no mesh left behind
            assert e.value.code == 6
            m.set_topology(rows, kk, free)
