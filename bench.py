#!/usr/bin/env python3
"""Throughput of the two-phase AMR-LBM step on one MI355X.

Workload: the 2D rising bubble on a 1024 x 2048 finest-equivalent lattice,
16 x 16-cell blocks, levels 3..6 (dx = 8, 4, 2, 1), D = 128 at (512, 512),
x periodic, bounce-back walls in y, regrid every 10 steps. A step is one
Main_Loop iteration (C2 fill, collision, hg fill, stream, C = sum h, C fill,
U and P, plus the regrid every 10th step), all on the GPU through the C-ABI.

value = finest-equivalent MLUPS = 1024 * 2048 * steps / wall seconds / 1e6,
timed over exactly --steps steps with the device synchronised on both sides.
The state starts resident in HBM (set-up is untimed).

The roofline entry is for the dominant kernel (the one with the largest share
of a step), timed with HIP events on the mesh's stream around 200
back-to-back launches over every leaf at the state the timed run ends in;
its algorithmic bytes per leaf cell are counted from the values the kernel
must read and write once (FP64):
  C2 fill + collision : read h,g (18) + C,P,Ux,Uy (4), write h,g (18) = 320 B
  hg fill + stream    : read h,g (18), write h,g (18) + C (1)          = 296 B
  C fill + U,P        : read g (9) + C (1), write P,Ux,Uy (3)          = 104 B
Ghost-cell, neighbour-table and regrid traffic is overhead, not counted.

fp32 repeats the measurement with the FP32 production form (fields stored in
FP32 as deviations from the liquid at rest, FP32 collision; bytes per leaf
cell halved) and states its error against the FP64 run after the same steps,
beside the FP64 run from a set-up perturbed by 1e-7 (the flow's own
sensitivity). fp32_learned is the metric's pair: FP32 with the learned regrid
(the committed network, amrlbm/regrid_cnn_weights.npy; full regrid every 20
steps, the plain criterion judging every step on the device, the network's
flags joining every regrid that runs) on the same bubble, timed the same way,
against plain FP64 and against the same cadence without the network.
fp32.learned_regrid_cfg1 deploys the network on the 128x256 bubble (the
learned configuration). regrid_interval sweeps the regrid interval of the
FP64 run over the configuration's 20,000 steps against the reference's
every-step cadence (amrlbm.tuning).

large_working_set repeats the measurement on 16 independent copies of the
same domain in one mesh (about 4,600 leaf blocks, 0.5 GB of FP64 leaf state,
past the 256 MB Infinity Cache), FP64 and FP32; second measurements, not the
headline.

multi_bubble runs the 32-bubble case (2048 x 4096, levels 3..7) in FP64,
FP32 and FP32 with the learned regrid; ensemble runs 64 rising-bubble
instances of uneven size, level range and fluid drawn from a seed, in one
mesh (every launch over the leaves of all instances), FP32 and FP64, with
each instance's FP32 error against FP64. Both report finest-equivalent MLUPS
summed over the whole mesh.

--sensitivity is an accuracy run, not a throughput one: every 2,000 steps the
FP32 run's Output against the FP64 run's, beside FP64 runs whose h and g got
relative noise of 1e-7 (FP32's size) once at set-up or again before every
--noise-every steps.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "amr-lbm-openmp-2d_amd"))

METRIC = "finest-equiv MLUPS + % HBM peak: plain FP64 vs FP32+learned regrid, 1 GPU"
HBM_PEAK_GBS = 8000.0
ALG_BYTES = {"c2_collide": 320, "hg_stream": 296, "c_macro_g": 104}
KERNEL_NAMES = {"c2_collide": "k_prog_c<OP_MACRO_G_COLLIDE>", "hg_stream": "k_prog_hg",
                "c_macro_g": "k_prog_c<OP_MACRO_G>"}
LAUNCHES_PER_STEP = {"c2_collide": 1.0, "hg_stream": 1.0}   # c_macro_g: once per segment


def progress(what, t0=[time.perf_counter()]):
    """One line on stderr as each entry of the line finishes (the JSON line
    itself goes to stdout at the end)."""
    print(f"bench: {what} done, {time.perf_counter() - t0[0]:.1f} s", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload, steps, regrid_every):
    """The reference Fortran/OpenMP build (oracle/_ref) on the host cores for a
    bounded sample of the same workload; falls back to the oracle's C
    restatement (one thread) when the reference build is absent."""
    nthreads = min(os.cpu_count() or 1, 16)
    exe = os.path.join(ROOT, "oracle", "_ref", workload.name, "ref_harness")
    fine = workload.fine_cells
    if os.path.exists(exe):
        with tempfile.TemporaryDirectory() as d:
            subprocess.run([exe, d, str(steps), str(regrid_every), "0", str(nthreads), "-1"],
                           check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            t = {}
            for line in open(os.path.join(d, "timing.txt")):
                k, v = line.split()
                t[k] = float(v)
        step_s = t["step_total"]
        return {"value": fine * steps / step_s / 1e6, "unit": "MLUPS (finest-equivalent)",
                "cores": nthreads, "kind": "reference",
                "sample": f"{steps} steps of the same workload, reference Fortran/OpenMP build "
                          f"(amdflang -O3 -fopenmp), {nthreads} threads on {cpu_model()} "
                          f"(nproc {os.cpu_count()}), step time = sum of the reference's phase timers"}
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    o = oracle.Oracle(workload)
    o.setup()
    n = max(steps // 10, 10)
    t0 = time.perf_counter()
    o.run(n, regrid_every)
    dt = time.perf_counter() - t0
    return {"value": fine * n / dt / 1e6, "unit": "MLUPS (finest-equivalent)", "cores": 1,
            "kind": "port", "sample": f"{n} steps of the same workload, C restatement, 1 thread"}


def calibration_copy(nbytes):
    """One hipMemcpy device-to-device of nbytes: the plain copy kernel whose
    FETCH_SIZE / WRITE_SIZE calibrate the rocprofv3 traffic counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def pmc_traffic(kernel_key, workload, precision=64):
    """HBM bytes per launch from committed rocprofv3 --pmc summaries, if present."""
    tag = "_fp32" if precision == 32 else ""
    path = os.path.join(ROOT, "profiles", f"{kernel_key}_{workload}{tag}_pmc.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get("hbm_bytes_per_launch")
    except Exception:
        return None


def roofline(key, kt, nleaf, w, precision):
    """Roofline entry of step kernel `key` from its average launch time kt[key] (ms)."""
    scale = 0.5 if precision == 32 else 1.0   # 4-byte values
    b = ALG_BYTES[key] * scale * nleaf * 256
    ach = b / (kt[key] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(key, w.name, precision),
            "kernel": KERNEL_NAMES[key], "avg_launch_us": round(kt[key] * 1e3, 3),
            "algorithmic_bytes_per_leaf_cell": ALG_BYTES[key] * scale, "leaf_blocks": nleaf,
            "bytes_per_launch": round(b)}


def kernels_only(amrlbm, w, calibrate_mb, precisions, variant=None, regrid_checks=0):
    """The profiling run: the set-up mesh of workload w in launch variant
    (order, depth) if given (the autotuner's choice in the default run; the
    tuner itself launches step graphs, which this run avoids), then the three
    step kernels launched back to back (amrlbm_time_kernels, the launches the
    bench line's roofline times), with no whole steps, then regrid_checks
    regrid decisions on the same state (amrlbm_regrid_check: Err_Interface and
    the decision kernel); one JSON line per precision."""
    for prec in precisions:
        with amrlbm.Mesh(w, precision=prec) as m:
            m.setup_rising_bubble()
            if variant:
                m.set_tuning(*variant)
            kt = m.time_kernels(200)
            for _ in range(regrid_checks):
                m.regrid_check()
            nleaf = m.counts()["nleaf"]
        if calibrate_mb:
            calibration_copy(calibrate_mb << 20)
        print(json.dumps({"mode": "kernels-only", "workload": w.name, "dtype": f"f{prec}", "leaf_blocks": nleaf,
                          "variant": list(variant) if variant else [0, 1],
                          "roofline": {kk: roofline(kk, kt, nleaf, w, prec) for kk in kt}}), flush=True)


def profile_steps(amrlbm, w, steps, k, precisions, variant=None):
    """The whole-step profiling run: steps steps of workload w with the regrid
    every k (its mesh changes included), every launch of the step as in the
    step graph but issued one by one with HIP events between them
    (amrlbm_profile_steps; no graph); per kernel class the device time per
    launch and the achieved rate over the leaf cells it covered, one JSON
    line per precision."""
    for prec in precisions:
        with amrlbm.Mesh(w, precision=prec) as m:
            m.setup_rising_bubble()
            if variant:
                m.set_tuning(*variant)
            st = m.profile_steps(steps, k)
            nleaf = m.counts()["nleaf"]
        scale = 0.5 if prec == 32 else 1.0
        cls = {}
        for key, v in st.items():
            if key == "regrid":
                continue
            n = max(v["launches"], 1)
            gbs = ALG_BYTES[key] * scale * v["cells"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 else 0.0
            cls[key] = {"kernel": KERNEL_NAMES[key], "launches": v["launches"],
                        "avg_launch_us": round(v["ms"] * 1e3 / n, 3), "achieved": round(gbs, 2),
                        "frac": round(gbs / HBM_PEAK_GBS, 4)}
        print(json.dumps({"mode": "profile-steps", "workload": w.name, "dtype": f"f{prec}", "steps": steps,
                          "regrid_every": k, "leaf_blocks_end": nleaf, "regrid_calls": st["regrid"]["count"],
                          "regrid_ms_per_call": round(st["regrid"]["ms"] / max(st["regrid"]["count"], 1), 4),
                          "variant": list(variant) if variant else None, "kernels": cls}), flush=True)


TUNE = {"on": True}


def tuning_entry(table, chosen):
    """The autotuner's table: the chosen launch variant and its time per step
    (the steps between two regrids) against the default variant's (order 0,
    depth 1, step graph)."""
    t = {(o, d, e): us for o, d, e, us in table if us > 0}
    return {"chosen": {"order": chosen[0], "hg_depth": chosen[1],
                       "execution": ["step graph", "segment kernel"][chosen[2]]},
            "us_per_step": round(t[chosen], 3), "default_us_per_step": round(t[(0, 1, 0)], 3),
            "gain_vs_default": round(t[(0, 1, 0)] / t[chosen], 4),
            "variants_us": {f"{o}/{d}/{e}": round(us, 3) for (o, d, e), us in sorted(t.items())}}


def windows_for(steps):
    """Timed windows of `steps` steps each: enough that about 20,000 steps are
    timed in all (one window at the configuration's 20,000), so that a short
    window's launch and synchronisation latency is measured many times over
    and reported as a median with its spread."""
    return max(1, min(1000, 20000 // max(steps, 1)))


def measure(amrlbm, w, steps, warmup, k, calibrate_mb=0, precision=64, radius=None, windows=None,
            mass_correction=False):
    """Time `steps` steps of workload w after `warmup` (and, untimed, the
    autotuner's choice of launch variant): windows_for(steps) consecutive
    windows of exactly `steps` steps, each between two device
    synchronisations; then the step kernels back to back. Returns the numbers
    of one bench entry (wall = the median window). mass_correction: the
    phase-field mass correction on from the set-up."""
    m = amrlbm.Mesh(w, precision=precision)
    m.setup_rising_bubble(radius=radius)
    diag0 = m.output()
    if mass_correction:
        m.set_mass_correction(True)
    m.step(warmup, k)
    tuning = tuning_entry(m.autotune(20, k), m.get_tuning()) if TUNE["on"] else None
    m.synchronize()
    c0 = m.counts()
    walls = []
    for _ in range(windows or windows_for(steps)):
        t0 = time.perf_counter()
        m.step(steps, k)
        m.synchronize()
        walls.append(time.perf_counter() - t0)
    walls.sort()
    wall = walls[len(walls) // 2]
    c1 = m.counts()
    rstats = m.regrid_stats()
    kt = m.time_kernels(200)
    if calibrate_mb:
        calibration_copy(calibrate_mb << 20)
    diag = m.output()
    m.close()
    share = {kk: kt[kk] * LAUNCHES_PER_STEP.get(kk, 1.0 / k) for kk in kt}
    dom = max(share, key=share.get)

    def roof(key):
        return roofline(key, kt, c1["nleaf"], w, precision)
    pct = lambda q: walls[min(len(walls) - 1, int(q * len(walls)))]
    timing = {"windows": len(walls), "steps_per_window": steps, "median_ms": round(wall * 1e3, 4),
              "min_ms": round(walls[0] * 1e3, 4), "p10_ms": round(pct(0.1) * 1e3, 4),
              "p90_ms": round(pct(0.9) * 1e3, 4), "max_ms": round(walls[-1] * 1e3, 4),
              "steps_timed_in_all": steps * len(walls)}
    return {
        "value": w.fine_cells * steps / wall / 1e6, "wall": wall, "c0": c0, "c1": c1, "regrids": rstats,
        "timing": timing,
        "roofline": roof(dom), "roofline_all": {kk: roof(kk) for kk in kt},
        "kernel_us": {kk: round(v * 1e3, 3) for kk, v in kt.items()},
        "step_kernel_us": sum(kt[kk] * LAUNCHES_PER_STEP.get(kk, 1.0 / k) for kk in kt) * 1e3,
        "diag": diag, "diag0": diag0, "tuning": tuning,
    }


def learned_regrid(amrlbm, w, steps=10000, full_every=50):
    """The learned regrid as its configuration states it: the committed
    network (trained on the generator of amrlbm.learned) deployed on workload
    w, FP32, full regrid every 50 steps with the plain criterion judging every
    step on the device and the network's flags joining every regrid, against
    the same cadence without the network and the plain regrid every 10 steps."""
    import numpy as np
    from amrlbm import learned
    if not os.path.exists(learned.WEIGHTS):
        return None
    ev = learned.evaluate(w, np.load(learned.WEIGHTS), steps, full_every=full_every)
    out = {"workload": f"{w.name} bubble (D={int(w.D)}), {steps} steps, full regrid every {full_every} "
                       f"steps, plain criterion judging every step on the device, the network's flags joining "
                       f"every regrid", "precision": "f32",
           "threshold": learned._threshold()}
    for k, v in ev.items():
        out[k] = {f: (round(x, 4) if isinstance(x, float) else x) for f, x in v.items() if f != "diag"}
    base = ev["plain_every_10"]["diag"]
    out["learned"]["mass_rel_vs_plain"] = abs(ev["learned"]["diag"][3] - base[3]) / abs(base[3])
    out["learned"]["rise_velocity_rel_vs_plain"] = abs(ev["learned"]["diag"][7] - base[7]) / abs(base[7])
    out["gain_mlups_vs_no_network"] = round(ev["learned"]["mlups"] / ev["no_network"]["mlups"], 3)
    out["gain_mlups_vs_plain_every_10"] = round(ev["learned"]["mlups"] / ev["plain_every_10"]["mlups"], 3)
    return out


def fp32_learned(amrlbm, w, steps, warmup, fp64_value, fp64_diag):
    """The north star's pair: FP32 with the learned regrid on the headline
    bubble, timed as measure() times (windows of exactly `steps` steps under
    the learned cadence), with the same cadence without the network beside
    it, and its error against the plain FP64 run after the same steps."""
    import numpy as np
    from amrlbm import learned
    if not os.path.exists(learned.WEIGHTS):
        return None
    wts = np.load(learned.WEIGHTS)
    out = {"workload": describe(w, learned.FULL_EVERY, 32).replace(
               f"regrid every {learned.FULL_EVERY}",
               f"full regrid every {learned.FULL_EVERY}, plain criterion judging every step on the device, "
               f"the network's flags joining every regrid"),
           "unit": "MLUPS (finest-equivalent)"}
    for name, net in (("learned", wts), ("no_network", None)):
        with amrlbm.Mesh(w, precision=32) as m:
            m.setup_rising_bubble()
            m.set_regrid_network(net, learned._threshold())   # once: the windows only step
            m.step_learned(warmup, learned.FULL_EVERY)
            m.synchronize()
            walls, st = [], None
            for _ in range(windows_for(steps)):
                t0 = time.perf_counter()
                s1 = m.step_learned(steps, learned.FULL_EVERY)
                m.synchronize()
                walls.append(time.perf_counter() - t0)
                s1 = {"predicted_regrids": s1["full_regrids"], "predicted_blocks": s1["predicted_blocks"],
                      "misses": s1["misses"], "checks": s1["judged"]}
                st = s1 if st is None else {f: st[f] + s1[f] for f in st}
            walls.sort()
            wall = walls[len(walls) // 2]
            e = {"value": round(w.fine_cells * steps / wall / 1e6, 1), "ms_per_step": round(wall * 1e3 / steps, 4),
                 "windows": len(walls), "leaf_blocks": m.counts()["nleaf"]}
            e.update(st)
            e["missed_refinement_rate"] = st["misses"] / max(st["checks"], 1)
            if name == "learned":
                d = m.output()
                e["error_vs_fp64"] = {"mass_rel": abs(d[3] - fp64_diag[3]) / abs(fp64_diag[3]),
                                      "rise_velocity_rel": abs(d[7] - fp64_diag[7]) / abs(fp64_diag[7])}
                out["diag"] = [float(v) for v in d]   # for the every-step comparison; not printed
            out[name] = e
    out["speedup_vs_plain_fp64"] = round(out["learned"]["value"] / fp64_value, 3)
    out["gain_vs_no_network"] = round(out["learned"]["value"] / out["no_network"]["value"], 3)
    return out


def regrid_interval(amrlbm, w, steps):
    """The regrid-interval knob (amrlbm.tuning): FP64 runs of workload w over
    `steps` steps from the set-up at each interval, their speed and their
    Output diagnostics against the reference's every-step cadence, and the
    fastest interval within the stated bounds."""
    from amrlbm import tuning
    rows = tuning.sweep(w, steps=steps)
    best = tuning.choose(rows)
    keep = ("regrid_every", "us_per_step", "changes", "leaves", "rise_velocity_rel", "mass_rel", "centroid_abs")
    entry = {"workload": describe(w, 1, 64).replace("regrid every 1", "regrid interval swept"), "steps": steps,
             "baseline": "regrid every step (Main_Loop, AMR_Rising_2D_OMP.f90:360)",
             "bounds": {"rise_velocity_rel": tuning.RISE_REL, "centroid_abs": tuning.CENTROID_ABS,
                        "mass_rel": tuning.MASS_REL},
             "rows": [{f: (round(r[f], 9) if isinstance(r[f], float) else r[f]) for f in keep} for r in rows],
             "chosen": best["regrid_every"]}
    return entry, rows[0]["diag"]


def versus(d, ref):
    """Output diagnostics d against ref: relative mass and rise velocity, centroid distance (cells)."""
    return {"mass_rel": abs(d[3] - ref[3]) / abs(ref[3]), "rise_velocity_rel": abs(d[7] - ref[7]) / abs(ref[7]),
            "centroid_abs_cells": max(abs(d[5] - ref[5]), abs(d[6] - ref[6]))}


def perturb_leaves(mesh, rng, noise):
    """Every leaf's interior h and g times (1 + noise * U(-1, 1)), drawn per
    value; C = sum of h again (FP64 mesh)."""
    for b in mesh.leaves():
        d = mesh.get_block(b)
        h, g, C = d["h"], d["g"], d["C"]
        h[:, 1:-1, 1:-1] *= 1.0 + noise * rng.uniform(-1.0, 1.0, h[:, 1:-1, 1:-1].shape)
        g[:, 1:-1, 1:-1] *= 1.0 + noise * rng.uniform(-1.0, 1.0, g[:, 1:-1, 1:-1].shape)
        C[2:-2, 2:-2] = h[:, 1:-1, 1:-1].sum(axis=0)
        mesh.put_block(b, h=h, g=g, C=C)


def sensitivity(amrlbm, w, steps, k, noise=1e-7, every=10, seed=1, check_every=2000):
    """The FP32 path's error against FP64 beside the flow's own sensitivity
    to noise of FP32's size: FP64 runs whose h and g get relative noise of
    `noise` once at set-up, or again before every `every` steps (host-side,
    through the per-block entry points). Every check_every steps, each run's
    Output against the plain FP64 run's."""
    import numpy as np

    def run(precision, inject=0):
        rng = np.random.default_rng(seed)
        rows = []
        with amrlbm.Mesh(w, precision=precision) as m:
            m.setup_rising_bubble()
            if inject < 0:
                perturb_leaves(m, rng, noise)
            t = 0
            while t < steps:
                if inject > 0:
                    perturb_leaves(m, rng, noise)
                n = min(inject if inject > 0 else check_every, check_every - t % check_every, steps - t)
                m.step(n, k)
                t += n
                if t % check_every == 0 or t == steps:
                    rows.append(m.output().tolist())
                    if inject > 0:
                        progress(f"sensitivity: noise every {inject} steps, step {t}")
        return rows

    base = run(64)
    progress("sensitivity: fp64")
    out = {"steps": steps, "regrid_every": k, "noise": noise, "seed": seed,
           "at_steps": [min((i + 1) * check_every, steps) for i in range(len(base))]}
    for name, precision, inject in (("fp32", 32, 0), ("fp64_noise_at_setup", 64, -1),
                                    (f"fp64_noise_every_{every}_steps", 64, every)):
        rows = run(precision, inject)
        out[name] = {f: [versus(d, r)[f] for d, r in zip(rows, base)]
                     for f in ("mass_rel", "rise_velocity_rel", "centroid_abs_cells")}
        progress(f"sensitivity: {name}")
    return out


def multi_bubble(amrlbm, steps, k):
    """The 32-bubble case (amrlbm.multibubble): FP64, FP32 and FP32 with the
    learned regrid (full regrid every learned.FULL_EVERY steps, plain
    criterion judging every step on the device, the network's flags joining
    every regrid) from the same device set-up, and the FP32 runs' total
    mass and per-bubble centroids against FP64's."""
    import numpy as np
    from amrlbm import learned
    from amrlbm import multibubble as mb
    w, bub = mb.WORKLOAD, mb.bubbles()
    out = {"workload": f"{len(bub)} bubbles, diameters 24..96 cells (seeded), {w.fine_shape[0]}x{w.fine_shape[1]} "
                       f"finest-equivalent, levels {w.min_levels}..{w.max_levels}, Eo={w.Bo}, W={w.W:g}, "
                       f"density ratio 1000, viscosity ratio 100, set up on the device",
           "steps": steps, "unit": "MLUPS (finest-equivalent)"}
    runs = {}
    for name, prec, use_net in (("fp64", 64, False), ("fp32", 32, False), ("fp32_learned", 32, True)):
        if use_net and not os.path.exists(learned.WEIGHTS):
            continue
        with amrlbm.Mesh(w, precision=prec) as m:
            mb.setup(m, bub)
            m.synchronize()
            t0 = time.perf_counter()
            try:
                st = learned.run(m, steps, np.load(learned.WEIGHTS)) if use_net else None
                if not use_net:
                    m.step(steps, k)
                m.synchronize()
            except amrlbm.AmrLbmError as exc:   # a state that turned non-finite is reported, not raised
                out[name] = {"error": str(exc)}
                continue
            wall = time.perf_counter() - t0
            runs[name] = {"diag": m.output(), "regions": mb.gas_regions(m)}
            e = {"value": round(w.fine_cells * steps / wall / 1e6, 1), "ms_per_step": round(wall * 1e3 / steps, 4),
                 "leaf_blocks": m.counts()["nleaf"], "regrid_every": learned.FULL_EVERY if use_net else k}
            if st:
                e.update({f: st[f] for f in ("predicted_regrids", "predicted_blocks", "misses", "checks")})
            out[name] = e
    if "fp64" not in runs:
        return out
    ref = runs["fp64"]
    out["fp64"]["gas_regions"] = len(ref["regions"])
    for name in ("fp32", "fp32_learned"):
        if name in runs:
            r = runs[name]
            out[name]["error_vs_fp64"] = {
                "mass_rel": abs(r["diag"][3] - ref["diag"][3]) / abs(ref["diag"][3]),
                "gas_regions": len(r["regions"]),
                "centroid_max_diff_cells": max(mb.match(ref["regions"], r["regions"]),
                                               mb.match(r["regions"], ref["regions"]))}
    return out


def ensemble_run(amrlbm, n, seed, steps, k):
    """The batched ensemble (amrlbm.ensemble): n instances drawn from seed in
    one mesh, FP32 and FP64 over the same steps from the same set-up, and
    each instance's FP32 mass, rise velocity and centroid against FP64."""
    import numpy as np
    from amrlbm.ensemble import ensemble
    e = ensemble(n, seed)
    out = {"workload": f"{n} rising-bubble instances drawn with seed {seed} (SplitMix64): domains L0 x 2 L0 with "
                       f"L0 in {{256, 512, 1024}}, 2-4 levels, D 16..64, Eo 1..100, density ratio 10..1000, "
                       f"viscosity ratio 1..100, W=5; one mesh, every launch over the leaves of all instances",
           "instances": n, "fine_cells": e.fine_cells, "steps": steps, "regrid_every": k,
           "unit": "MLUPS (finest-equivalent, all instances)"}
    diags = {}
    for prec in (32, 64):
        with amrlbm.Mesh(e, precision=prec) as m:
            try:
                m.setup_rising_bubble()
                tuning = tuning_entry(m.autotune(3, k), m.get_tuning()) if TUNE["on"] else None
                m.synchronize()
                c0 = m.counts()
                s0 = m.regrid_stats()
                t0 = time.perf_counter()
                m.step(steps, k)
                m.synchronize()
                wall = time.perf_counter() - t0
                s1 = m.regrid_stats()
                diags[prec] = m.output_instances()
            except amrlbm.AmrLbmError as exc:   # reported, not raised
                out[f"fp{prec}"] = {"error": str(exc)}
                continue
            changes = s1["changes"] - s0["changes"]
            # Refine_Derefine_Mesh, the table and ghost-program upload (programs included) and graph capture
            host = sum(s1["host_ms"][f] - s0["host_ms"][f] for f in ("regrid", "tables", "graphs"))
            out[f"fp{prec}"] = {
                "value": round(e.fine_cells * steps / wall / 1e6, 1), "ms_per_step": round(wall * 1e3 / steps, 4),
                "leaf_blocks": [c0["nleaf"], m.counts()["nleaf"]], "mesh_changes": changes,
                "host_ms_per_change": round(host / max(changes, 1), 2), "tuning": tuning}
    if 32 in diags and 64 in diags:
        d32, d64 = diags[32], diags[64]
        mass = np.abs(d32[:, 3] - d64[:, 3]) / np.abs(d64[:, 3])
        rise = np.abs(d32[:, 7] - d64[:, 7]) / np.abs(d64[:, 7])
        cen = np.maximum(np.abs(d32[:, 5] - d64[:, 5]), np.abs(d32[:, 6] - d64[:, 6]))
        out["fp32_error_vs_fp64"] = {
            "mass_rel_max": float(mass.max()), "mass_rel_median": float(np.median(mass)),
            "rise_velocity_rel_max": float(rise.max()), "rise_velocity_rel_median": float(np.median(rise)),
            "centroid_abs_max_cells": float(cen.max()),
            "nonfinite_instances": int(np.sum(d32[:, 8] != 0) + np.sum(d64[:, 8] != 0))}
    return out


def describe(w, k, precision=64):
    path = "plain FP64 path" if precision == 64 else "FP32 path (deviation storage)"
    what = (f"rising bubble {w.fine_shape[0]}x{w.fine_shape[1]} finest-equivalent, 16x16 blocks, "
            f"levels {w.min_levels}..{w.max_levels}, D={int(w.D)}, regrid every {k}, {path}; "
            f"fluid: the reference's own set (Bo={w.Bo}, Ar={w.Ar:g}, g={w.gy:g}, W={w.W:g}, M0={w.M0:g}, "
            f"density ratio {w.rhoh / w.rhol:g}, viscosity ratio {w.visc_ratio:g}), not the configuration's "
            f"W=4, sigma=0.01, tau_h=0.8, which turn non-finite in the reference")
    if w.instances > 1:
        what = f"{w.instances} independent copies of the {w.L0}x{w.nroots // w.instances * w.L0} " + what
    return what


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000,
                    help="timed steps (the 1024x2048 configuration's full run length)")
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--regrid-every", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=600)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 path's entries")
    ap.add_argument("--large-workload", default="cfg2x16",
                    help="replicated-instance workload measured after the main one ('' to skip)")
    ap.add_argument("--large-steps", type=int, default=200)
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (rocprof counter calibration)")
    ap.add_argument("--multi-steps", type=int, default=3000,
                    help="steps of the 32-bubble case (0 to skip)")
    ap.add_argument("--ensemble-steps", type=int, default=10000,
                    help="steps of the 64-instance ensemble (0 to skip)")
    ap.add_argument("--interval-steps", type=int, default=-1,
                    help="steps of each run of the regrid-interval sweep (default: as many as each headline run "
                         "takes, warmup included; 0 to skip)")
    ap.add_argument("--ensemble-size", type=int, default=64)
    ap.add_argument("--ensemble-seed", type=int, default=2024)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch variant (no autotuner before the timed runs)")
    ap.add_argument("--variant", default="",
                    help="launch variant 'order,depth' for --kernels-only (amrlbm_set_tuning)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="profiling run: time the step kernels on the set-up mesh only (no whole steps)")
    ap.add_argument("--regrid-checks", type=int, default=0,
                    help="with --kernels-only: that many regrid decisions on the set-up state after the kernels")
    ap.add_argument("--profile-steps", action="store_true",
                    help="profiling run: --steps whole steps with regrids, launches issued one by one (no graph)")
    ap.add_argument("--sensitivity", action="store_true",
                    help="accuracy run: the FP32 error against FP64 beside FP64 runs perturbed by FP32-sized "
                         "noise at set-up and every --noise-every steps, over --steps steps (several minutes)")
    ap.add_argument("--noise-every", type=int, default=10)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import amrlbm
    TUNE["on"] = not args.no_autotune
    w = amrlbm.WORKLOADS[args.workload]
    k = args.regrid_every
    if args.kernels_only:
        variant = tuple(int(v) for v in args.variant.split(",")) if args.variant else None
        kernels_only(amrlbm, w, args.calibrate_copy_mb, (64,) if args.no_fp32 else (64, 32), variant, args.regrid_checks)
        return
    if args.profile_steps:
        variant = tuple(int(v) for v in args.variant.split(",")) if args.variant else None
        profile_steps(amrlbm, w, args.steps, k, (64,) if args.no_fp32 else (64, 32), variant)
        return
    if args.sensitivity:
        print(json.dumps({"sensitivity": sensitivity(amrlbm, w, args.steps, k, every=args.noise_every)}), flush=True)
        return
    r = measure(amrlbm, w, args.steps, args.warmup, k, args.calibrate_copy_mb)
    progress("fp64 headline")
    c0, c1 = r["c0"], r["c1"]
    nleaf_mean = 0.5 * (c0["nleaf"] + c1["nleaf"])
    diag = r["diag"]
    out = {
        "metric": METRIC, "value": round(r["value"], 1), "unit": "MLUPS (finest-equivalent)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(r["wall"] * 1e3 / args.steps, 4), "higher_is_better": True,
        "timing": dict(r["timing"], note="value and ms_per_step are the median window; every window is exactly "
                                         "--steps steps between two device synchronisations"),
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: analytic tanh rising-bubble initial state, deterministic",
        "config": {"workload": describe(w, k), "leaf_blocks": [c0["nleaf"], c1["nleaf"]],
                   "leaf_cell_mlups": round(nleaf_mean * 256 * args.steps / r["wall"] / 1e6, 1)},
        "roofline": r["roofline"],
        "kernel_avg_us": r["kernel_us"],
        "kernel_us_per_step": round(r["step_kernel_us"], 3),
        "regrids": r["regrids"],
        "tuning": r["tuning"],
        "diagnostics": {"mass": diag[3], "rise_velocity": diag[7], "centroid_y": diag[6]},
    }
    d32 = None
    if not args.no_fp32:
        r32 = measure(amrlbm, w, args.steps, args.warmup, k, precision=32)
        progress("fp32")
        d32 = r32["diag"]
        # the flow's own sensitivity: FP64 from a set-up whose bubble radius is 1e-7 larger
        rp = measure(amrlbm, w, args.steps, args.warmup, k, precision=64, radius=w.radius * (1.0 + 1e-7))
        dp = rp["diag"]
        progress("fp64 perturbed set-up")
        out["fp32"] = {
            "workload": describe(w, k, 32), "value": round(r32["value"], 1), "unit": "MLUPS (finest-equivalent)",
            "speedup_vs_fp64": round(r32["value"] / r["value"], 3),
            "ms_per_step": round(r32["wall"] * 1e3 / args.steps, 4), "dtype": "f32", "timing": r32["timing"],
            "roofline": r32["roofline"], "kernel_avg_us": r32["kernel_us"], "tuning": r32["tuning"],
            "error_vs_fp64": {
                "after_steps": args.warmup + args.steps * r32["timing"]["windows"],
                "mass_rel": abs(d32[3] - diag[3]) / abs(diag[3]),
                "rise_velocity_rel": abs(d32[7] - diag[7]) / abs(diag[7]),
                "centroid_abs_cells": max(abs(d32[5] - diag[5]), abs(d32[6] - diag[6])),
                "same_mesh": r32["c1"] == r["c1"],
                "fp64_radius_perturbed_1e-7_vs_fp64": {
                    "mass_rel": abs(dp[3] - diag[3]) / abs(diag[3]),
                    "rise_velocity_rel": abs(dp[7] - diag[7]) / abs(diag[7]),
                    "centroid_abs_cells": max(abs(dp[5] - diag[5]), abs(dp[6] - diag[6]))},
                "targets": {"mass_rel": 1e-6, "rise_velocity_rel": 5e-3}},
            "mass_change_since_setup": {"fp64": (diag[3] - r["diag0"][3]) / r["diag0"][3],
                                        "fp32": (d32[3] - r32["diag0"][3]) / r32["diag0"][3]},
            # config 4: the network trained on the generator, deployed on the 128x256 bubble (D=32 at (64, 64))
            "learned_regrid_cfg1": learned_regrid(amrlbm, amrlbm.WORKLOADS["cfg1"]),
        }
        progress("learned regrid on the 128x256 bubble")
        # the FP64 phase-field mass correction on the FP32 path
        rc = measure(amrlbm, w, args.steps, args.warmup, k, precision=32, mass_correction=True)
        dc = rc["diag"]
        out["fp32"]["mass_corrected"] = {
            "what": "FP32 path with amrlbm_set_mass_correction: each segment end (every regrid interval) puts the "
                    "FP64 phase-field mass back to its budget (set-up mass plus each mesh change's restrict/prolong "
                    "change); the reference path itself is not mass-conserving",
            "value": round(rc["value"], 1), "unit": "MLUPS (finest-equivalent)",
            "ms_per_step": round(rc["wall"] * 1e3 / args.steps, 4), "speedup_vs_fp64": round(rc["value"] / r["value"], 3),
            "mass_change_since_setup": (dc[3] - rc["diag0"][3]) / rc["diag0"][3],
            "error_vs_fp64": {"mass_rel": abs(dc[3] - diag[3]) / abs(diag[3]),
                              "rise_velocity_rel": abs(dc[7] - diag[7]) / abs(diag[7]),
                              "centroid_abs_cells": max(abs(dc[5] - diag[5]), abs(dc[6] - diag[6]))}}
        progress("fp32 with the mass correction")
        out["fp32_learned"] = fp32_learned(amrlbm, w, args.steps, args.warmup, r["value"], diag)
        progress("fp32 learned")
    learned_diag = out["fp32_learned"].pop("diag", None) if out.get("fp32_learned") else None
    total = args.warmup + args.steps * windows_for(args.steps)   # the steps every headline run took
    n_int = args.interval_steps if args.interval_steps >= 0 else total
    if n_int > 0:
        out["regrid_interval"], every = regrid_interval(amrlbm, w, n_int)
        progress("regrid interval sweep")
        if n_int == total:   # the same steps: each run against the every-step FP64 run
            cmp = {"fp64_regrid_every_10": versus(diag, every)}
            if d32 is not None:
                cmp["fp32_regrid_every_10"] = versus(d32, every)
            if learned_diag is not None:
                cmp["fp32_learned"] = versus(learned_diag, every)
            out["regrid_interval"]["vs_fp64_every_step"] = cmp
    if args.large_workload:
        wl = amrlbm.WORKLOADS[args.large_workload]
        warm = max(args.warmup // 5, 10)
        entries = {}
        for prec in ((64,) if args.no_fp32 else (64, 32)):
            rl = measure(amrlbm, wl, args.large_steps, warm, k, precision=prec, windows=1)
            ll = rl["c1"]["nleaf"]
            entries[f"fp{prec}"] = {
                "workload": describe(wl, k, prec), "steps": args.large_steps, "warmup": warm,
                "leaf_blocks": ll,
                "leaf_state_mb": round(ll * (2 * 18 * 324 + 400 + 3 * 256) * (prec // 8) / 1e6, 1),
                "value": round(rl["value"], 1), "unit": "MLUPS (finest-equivalent)",
                "ms_per_step": round(rl["wall"] * 1e3 / args.large_steps, 4),
                "roofline": {kk: {f: v[f] for f in ("kernel", "achieved", "frac", "avg_launch_us", "traffic")}
                             for kk, v in rl["roofline_all"].items()},
                "regrids": {f: rl["regrids"][f] for f in ("checks", "changes")},
                "tuning": rl["tuning"],
            }
        out["large_working_set"] = entries
        progress("large working set")
    if args.multi_steps > 0:
        out["multi_bubble"] = multi_bubble(amrlbm, args.multi_steps, k)
        progress("multi bubble")
    if args.ensemble_steps > 0:
        out["ensemble"] = ensemble_run(amrlbm, args.ensemble_size, args.ensemble_seed, args.ensemble_steps, k)
        progress("ensemble")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_steps, k)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
