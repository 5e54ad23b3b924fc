"""Restart files (amrlbm_save_snapshot / amrlbm_load_snapshot) in the layout of
the reference dump harness (oracle/ref_harness.f90).

* The set-up state of the tiny workload saved by the library is the same file,
  byte for byte, as the reference harness's init.bin (golden SHA-256).
* A reference snapshot loaded and saved again is unchanged.
* A run restarted from a file continues exactly as the run that wrote it.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import mesh_state

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm as mod
    mod.lib()
    return mod


def _sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def test_setup_snapshot_is_the_reference_file(amrlbm, workloads, tmp_path):
    golden = json.load(open(os.path.join(GOLDEN, "reference_hashes.json")))["tiny"]
    with amrlbm.Mesh(workloads["tiny"]) as m:
        m.setup_rising_bubble()
        path = str(tmp_path / "init.bin")
        m.save_snapshot(path)
    print(f"[snapshot] tiny set-up saved: {_sha(path)}")
    assert _sha(path) == golden["init"]


def test_reference_snapshot_round_trip(amrlbm, workloads, tmp_path):
    src = os.path.join(GOLDEN, "tiny_s1_p3.bin")
    with amrlbm.Mesh(workloads["tiny"]) as m:
        m.load_snapshot(src)
        out = str(tmp_path / "again.bin")
        m.save_snapshot(out)
    assert open(out, "rb").read() == open(src, "rb").read()


@pytest.mark.parametrize("precision", [64, 32])
def test_restart_continues_the_run(amrlbm, workloads, tmp_path, precision):
    w = workloads["cfg1"]
    path = str(tmp_path / "restart.bin")
    with amrlbm.Mesh(w, precision=precision) as a:
        a.setup_rising_bubble()
        a.step(60, 10)
        a.save_snapshot(path)
        a.step(40, 10)
        ra, fa, ka, xa = mesh_state(a)
        da = a.output()
    with amrlbm.Mesh(w, precision=precision) as b:
        b.load_snapshot(path)
        assert b.counts()["T"] == 60
        b.step(40, 10)
        rb, fb, kb, xb = mesh_state(b)
        db = b.output()
    assert np.array_equal(ra, rb) and list(fa) == list(fb) and ka == kb
    for n in xa:
        for f in ("C", "P", "Ux", "Uy"):
            assert np.array_equal(xa[n][f], xb[n][f]), (n, f)
        for f in ("h", "g"):
            assert np.array_equal(xa[n][f][:, 1:-1, 1:-1], xb[n][f][:, 1:-1, 1:-1]), (n, f)
    assert np.array_equal(da, db)
    print(f"[snapshot] FP{precision} restart at step 60: 40 more steps bitwise equal ({len(xa)} leaves)")
