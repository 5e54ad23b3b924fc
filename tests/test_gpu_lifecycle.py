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
