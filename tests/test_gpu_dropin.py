"""The reference's own program on the GPU library: its driver, set-up and
Tecplot writer (AMR_Rising_2D_OMP.f90, TecPlot_ASCII.f90) linked with the
Fortran shim and libamrlbm.so in place of the hot-path files
(oracle/build_dropin.sh). It must print the reference program's Output table
and write its Tecplot files byte for byte (16x16 blocks, L0=64, three roots,
12800 steps, regrid every step; mesh grows from 31 to 63 blocks), with each
of the reference's phase timers timing its own phase."""
import hashlib
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin", "dropin_program")


def output_table(stdout):
    rows = []
    for line in stdout.splitlines():
        parts = line.split()
        if len(parts) == 10:
            try:
                float(parts[0])
                [int(v) for v in parts[1:4]]
                rows.append(" ".join(parts))
            except ValueError:
                pass
    return rows


def test_reference_program_on_gpu_library(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("drop-in program not built (oracle/build_dropin.sh needs the reference sources)")
    golden = json.load(open(os.path.join(GOLDEN, "reference_hashes.json")))["dropin"]
    want = open(os.path.join(GOLDEN, "dropin_output.txt")).read().split("\n")[:-1]
    print("[dropin] running the reference driver on libamrlbm.so", flush=True)
    res = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    got = output_table(res.stdout)
    assert len(got) == len(want), (len(got), len(want), res.stdout[-1500:])
    for a, b in zip(got, want):
        assert a == b, f"Output row differs:\n gpu: {a}\n ref: {b}"
    tec = {f: hashlib.sha256(open(os.path.join(tmp_path, f), "rb").read()).hexdigest()
           for f in sorted(os.listdir(tmp_path)) if f.startswith("Rising_Bubble_2D_")}
    assert tec == golden["tecplot"]
    print(f"[dropin] {len(got)} Output rows and {len(tec)} Tecplot files identical to the reference program")
    # the reference's phase timers (AMR_Rising_2D_OMP.f90:366-375) each time their own phase on the GPU
    timers = {}
    for line in res.stdout.splitlines():
        parts = line.replace(":", " ").split()
        if len(parts) == 2 and parts[0].startswith("t_"):
            timers[parts[0]] = float(parts[1])
    print(f"[dropin] phase timers (s): {timers}")
    for t in ("t_collision", "t_advection", "t_Ghost_C2", "t_Ghost_hg", "t_Ghost_C", "t_Macroscopic_h",
              "t_Macroscopic_g", "t_refine"):
        assert timers.get(t, 0.0) > 0.0, (t, timers)
