"""The regrid-interval knob (no counterpart in the reference as a knob).

Main_Loop runs Refine_Derefine_Mesh after every step
(AMR_Rising_2D_OMP.f90:302-364, the call at :360); amrlbm_step takes the
interval k as an argument (the 1024x2048 configuration states k = 10). A
longer interval is cheaper (fewer device decisions, fewer mesh changes, longer
step graphs) but lets the interface run up to k - 1 steps ahead of its
refinement, so the results differ from the every-step run: unlike the launch
variants of amrlbm_autotune, which are bit-exact, this knob is chosen against
an accuracy bound.

sweep times each candidate interval over a run of the workload from its
set-up (the same steps for every candidate, FP64 or FP32) and states its
Output diagnostics against the first candidate's (the reference's every-step
cadence by default); choose keeps the fastest candidate within the bounds.
"""
import time

from . import Mesh

DEFAULT_INTERVALS = (1, 2, 5, 10, 20, 50)
# bounds against the every-step run: the FP32 path's own targets against FP64
RISE_REL, CENTROID_ABS, MASS_REL = 5e-3, 0.1, 1e-6


def sweep(w, intervals=DEFAULT_INTERVALS, steps=20000, precision=64):
    """One row per interval: microseconds per step (wall, device synchronised
    on both sides, set-up untimed), mesh changes, leaves at the end and the
    relative rise velocity and mass and the centroid distance (cells) against
    the run of intervals[0]."""
    rows = []
    for k in intervals:
        with Mesh(w, precision=precision) as m:
            m.setup_rising_bubble()
            m.synchronize()
            t0 = time.perf_counter()
            m.step(steps, k)
            m.synchronize()
            wall = time.perf_counter() - t0
            d = m.output()
            rows.append({"regrid_every": k, "us_per_step": round(wall / steps * 1e6, 3),
                         "changes": m.regrid_stats()["changes"], "leaves": m.counts()["nleaf"],
                         "diag": [float(v) for v in d]})
    b = rows[0]["diag"]
    for r in rows:
        d = r["diag"]
        r["rise_velocity_rel"] = abs(d[7] - b[7]) / abs(b[7])
        r["mass_rel"] = abs(d[3] - b[3]) / abs(b[3])
        r["centroid_abs"] = max(abs(d[5] - b[5]), abs(d[6] - b[6]))
    return rows


def choose(rows, rise_rel=RISE_REL, centroid_abs=CENTROID_ABS, mass_rel=MASS_REL):
    """The fastest row within the bounds (the first row, the baseline, always is)."""
    ok = [r for r in rows
          if r["rise_velocity_rel"] <= rise_rel and r["centroid_abs"] <= centroid_abs and r["mass_rel"] <= mass_rel]
    return min(ok, key=lambda r: r["us_per_step"])
