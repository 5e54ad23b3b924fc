"""Writes the golden vectors of the reference build (oracle/_ref, compiled from
the reference sources by oracle/build_ref.sh) into tests/golden/.

    bash oracle/build_ref.sh && python tests/golden/make_golden.py

Outputs
  reference_hashes.json   SHA-256 of every binary snapshot the reference
                          harness writes in the sequences below (the file
                          layout of oracle/ref_harness.f90)
  tiny_init.bin           the tiny workload's state after set-up (full data)
  tiny_s1_p3.bin          the tiny workload after the hg ghost fill of step 1
  (tecplot keys)          SHA-256 of the reference's Tecplot file (TecPlot_CellCenter,
                          TecPlot_ASCII.f90) of the final state of the tiny, cfg1 and
                          cfg2 runs
  cfg1_diag.txt           reference FP64 diagnostics, cfg1, 300 steps
  cfg2_diag.txt           reference FP64 diagnostics, cfg2, 40 steps, regrid every 10
Sequences
  tiny   20 steps, regrid every step, snapshots after every phase of steps 0, 1, 10
  cfg1   300 steps, regrid every step (init + final)
  cfg2   40 steps, regrid every 10 (init + final)
  walk   cfg1 set-up, then 14 moves of the analytic bubble by (3, 5) cells with a
         Refine_Derefine_Mesh after each move
  rise   runs whose mesh changes while stepping (refinement, derefinement and
         reuse of freed ids), snapshots after every phase of each step whose
         regrid changes the mesh:
           tiny_rise  tiny set-up, gravity 1e-4 in the steps, 2400 steps, regrid every step
           cfg1_rise  cfg1 set-up, gravity 5e-5 in the steps, 2200 steps, regrid every step
           cfg2_regrid cfg2 as it is, 1100 steps, regrid every 10
  dropin the reference's complete program (oracle/build_dropin.sh: 16x16 blocks,
         L0=64, three roots, Tf=12800 steps, regrid every step): its Output table
         (dropin_output.txt) and the SHA-256 of its 21 Tecplot files
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def output_table(stdout):
    """The rows Output prints (AMR_Rising_2D_OMP.f90:486-487): T, Nblks, nLeaf,
    KK, rho_min, rho_max, U_max, mass, levels."""
    rows = []
    for line in stdout.splitlines():
        parts = line.split()
        if len(parts) == 10:
            try:
                float(parts[0]); [int(v) for v in parts[1:4]]
                rows.append(" ".join(parts))
            except ValueError:
                pass
    return rows


def run(cfg, d, args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    e["AMRLBM_TECPLOT"] = "1"   # the final state's Tecplot file, Rising_Bubble_2D_00.dat in d
    subprocess.run([os.path.join(REF, cfg, "ref_harness"), d] + [str(a) for a in args], check=True,
                   env=e, stdout=subprocess.DEVNULL, cwd=d)


# (sequence, workload of amrlbm.workloads, reference build, gravity override
# of the time steps, steps, regrid interval, steps whose regrid changes the
# mesh); the set-up's parameters are the build's own (the harness reads
# AMRLBM_GY_OVERRIDE after the set-up)
RISE = [
    ("tiny_rise", "tiny_rise", "tiny", "1e-4", 2400, 1, [1077, 1517, 2347]),
    ("cfg1_rise", "cfg1_rise", "cfg1", "5e-5", 2200, 1, [813, 1740, 1924, 2161]),
    ("cfg2_regrid", "cfg2", "cfg2", None, 1100, 10, [709, 1079]),
]


def rise_sequences(tmp):
    out = {}
    for name, workload, cfg, gy, steps, k, change_steps in RISE:
        d = os.path.join(tmp, name)
        os.makedirs(d)
        env = {"AMRLBM_GY_OVERRIDE": gy} if gy else {}
        run(cfg, d, [steps, k, 0, 8] + change_steps, env=env)
        out[name] = {"workload": workload, "steps": steps, "regrid_every": k, "gy_step": float(gy) if gy else None,
                     "change_steps": change_steps,
                     "init": sha(os.path.join(d, "init.bin")), "final": sha(os.path.join(d, "final.bin")),
                     "phases": {f"s{t}_p{p}": sha(os.path.join(d, f"s{t:06d}_p{p}.bin"))
                                for t in change_steps for p in range(9)}}
    return out


def main():
    for cfg in ("tiny", "cfg1", "cfg2"):
        if not os.path.exists(os.path.join(REF, cfg, "ref_harness")):
            sys.exit(f"reference build missing: run oracle/build_ref.sh ({cfg})")
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        d = os.path.join(tmp, "tiny")
        os.makedirs(d)
        run("tiny", d, [20, 1, 1, 1, 0, 1, 10])
        out["tiny"] = {"steps": 20, "regrid_every": 1, "phase_steps": [0, 1, 10],
                       "init": sha(os.path.join(d, "init.bin")), "final": sha(os.path.join(d, "final.bin")),
                       "phases": {f"s{t}_p{k}": sha(os.path.join(d, f"s{t:06d}_p{k}.bin"))
                                  for t in (0, 1, 10) for k in range(9)},
                       "tecplot": sha(os.path.join(d, "Rising_Bubble_2D_00.dat"))}
        shutil.copy(os.path.join(d, "init.bin"), os.path.join(HERE, "tiny_init.bin"))
        shutil.copy(os.path.join(d, "s000001_p3.bin"), os.path.join(HERE, "tiny_s1_p3.bin"))

        d = os.path.join(tmp, "cfg1")
        os.makedirs(d)
        run("cfg1", d, [300, 1, 50, 4, -1])
        out["cfg1"] = {"steps": 300, "regrid_every": 1, "init": sha(os.path.join(d, "init.bin")),
                       "final": sha(os.path.join(d, "final.bin")),
                       "tecplot": sha(os.path.join(d, "Rising_Bubble_2D_00.dat"))}
        shutil.copy(os.path.join(d, "diag.txt"), os.path.join(HERE, "cfg1_diag.txt"))

        d = os.path.join(tmp, "cfg2")
        os.makedirs(d)
        run("cfg2", d, [40, 10, 10, 8, -1])
        out["cfg2"] = {"steps": 40, "regrid_every": 10, "init": sha(os.path.join(d, "init.bin")),
                       "final": sha(os.path.join(d, "final.bin")),
                       "tecplot": sha(os.path.join(d, "Rising_Bubble_2D_00.dat"))}
        shutil.copy(os.path.join(d, "diag.txt"), os.path.join(HERE, "cfg2_diag.txt"))

        d = os.path.join(tmp, "walk")
        os.makedirs(d)
        run("cfg1", d, [0, 1, 1, 1, -1], env={"AMRLBM_WALK": "14 3.0 5.0"})
        out["walk"] = {"workload": "cfg1", "moves": 14, "step": [3.0, 5.0],
                       "snapshots": [sha(os.path.join(d, f"walk{k:03d}.bin")) for k in range(1, 15)]}
        out["rise"] = rise_sequences(tmp)
        exe = os.path.join(REF, "dropin", "ref_program")
        if os.path.exists(exe):
            d = os.path.join(tmp, "dropin")
            os.makedirs(d)
            res = subprocess.run([exe], cwd=d, check=True, capture_output=True, text=True)
            rows = output_table(res.stdout)
            with open(os.path.join(HERE, "dropin_output.txt"), "w") as f:
                f.write("\n".join(rows) + "\n")
            out["dropin"] = {"tecplot": {f: sha(os.path.join(d, f)) for f in sorted(os.listdir(d))
                                         if f.startswith("Rising_Bubble_2D_")},
                             "rows": len(rows)}
    with open(os.path.join(HERE, "reference_hashes.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "reference_hashes.json"))


if __name__ == "__main__":
    main()
