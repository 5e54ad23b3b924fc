"""The Tecplot writer (amrlbm_tecplot_cell_center) against the reference's
own TecPlot_CellCenter (TecPlot_ASCII.f90), compiled from the reference
sources into the harness (oracle/build_ref.sh): after the same steps the
library holds the reference's state bit for bit, so the files must be the
same bytes (SHA-256 in tests/golden/reference_hashes.json, written by
tests/golden/make_golden.py)."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm as mod
    mod.lib()
    return mod


@pytest.mark.parametrize("name", ["tiny", "cfg1", "cfg2"])
def test_tecplot_file_is_the_reference_file(amrlbm, workloads, tmp_path, name):
    golden = json.load(open(os.path.join(GOLDEN, "reference_hashes.json")))[name]
    w = workloads[name]
    print(f"[tecplot] {name}: {golden['steps']} steps, regrid every {golden['regrid_every']}", flush=True)
    with amrlbm.Mesh(w) as m:
        m.setup_rising_bubble()
        m.step(golden["steps"], golden["regrid_every"])
        path = str(tmp_path / "Rising_Bubble_2D_00.dat")
        m.write_tecplot(path)
    data = open(path, "rb").read()
    print(f"[tecplot] {name}: {len(data)} bytes, header {data.splitlines()[2].decode()}")
    assert hashlib.sha256(data).hexdigest() == golden["tecplot"]


def _read(path):
    lines = open(path).read().splitlines()
    nn = int(lines[1].split("N =")[1].split(",")[0])
    ne = int(lines[1].split("E =")[1].split(",")[0])
    vals = " ".join(lines[3:]).split()
    return nn, ne, vals


def test_fp32_tecplot_close_to_fp64(amrlbm, workloads, tmp_path):
    """The FP32 mesh writes the same layout; its fields agree with FP64's to
    the FP32 bounds of the step (C to 1e-3)."""
    w = workloads["cfg1"]
    out = {}
    for prec in (64, 32):
        with amrlbm.Mesh(w, precision=prec) as m:
            m.setup_rising_bubble()
            m.step(100, 5)
            p = str(tmp_path / f"tec{prec}.dat")
            m.write_tecplot(p)
            out[prec] = _read(p)
    (n64, e64, v64), (n32, e32, v32) = out[64], out[32]
    assert (n64, e64) == (n32, e32) and len(v64) == len(v32)
    xy = 2 * n64
    c0 = xy + 2 * e64
    c64 = np.array([float(v) for v in v64[c0:c0 + n64]])   # C - 1/2 lies in E+00 .. E-99
    c32 = np.array([float(v) for v in v32[c0:c0 + n64]])
    assert v64[:xy] == v32[:xy]                     # node coordinates
    assert v64[c0 + n64 + e64:] == v32[c0 + n64 + e64:]   # connectivity
    print(f"[tecplot] FP32 vs FP64 node C: max diff {np.max(np.abs(c64 - c32)):.2e}")
    assert np.max(np.abs(c64 - c32)) <= 1e-3
