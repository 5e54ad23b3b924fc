"""Learned regrid: a convolutional network over each leaf block's phase field
and velocity predicts whether the block's region will be refined (or
derefined) HORIZON steps ahead, so that blocks the interface is about to
enter are refined before it arrives and the full regrid can run every
FULL_EVERY steps. The plain gradient criterion (Refine_Derefine.f90:18-54,
:327-352) stays the judge: the library tests it on the device at every
step's state, and the interface on an under-refined block triggers an
immediate plain regrid (a miss). The network never changes the mesh on its
own: its flags join every regrid that runs (full or miss), refining ahead
the blocks the interface is about to reach, which removes the misses they
would have caused. amrlbm_step_learned runs that cadence with every
decision on the device; the host works only when the mesh changes.

The network, its training and inference are the library's HIP kernels
(amrlbm_cnn_*, amrlbm_set_regrid_network); this module builds the training
set from the solver's own runs (the generator below), trains, and evaluates.

Training generator: trajectories of the 1024x2048 configuration's recipe
(two roots stacked in y, 16x16 blocks, four levels, one rising bubble,
regrid every 10 steps) at a quarter of its size (256 x 512, levels 1..4),
with the bubble's diameter uniform in [24, 40] finest cells and its centre
uniform in [R + 8, L0 - R - 8] in x (the set-up's profile does not wrap
around the periodic x boundary) and in [L0/4, L0/2] in y, drawn from a seed
with SplitMix64,
and the gravity of the time steps uniform between the reference's and
GEN_GY_MAX (eight times it), so that slow interfaces and interfaces that
cross into new blocks within the 400 steps of a trajectory are both in the
set. The trajectories run 64 at a time in one ensemble
mesh; an instance that turns non-finite is left out of the set.

    python -m amrlbm.learned train    # writes regrid_cnn_weights.npy beside this file
"""
import dataclasses
import os
import sys
import time

import numpy as np

from . import Mesh, cnn_forward, cnn_init, cnn_train
from .ensemble import Ensemble, Stream
from .workloads import WORKLOADS, Workload

HORIZON = 50       # labels: the topology this many steps after the input state
FULL_EVERY = 20    # the cadence's full-regrid interval (amrlbm_step_learned)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "regrid_cnn_weights.npy")
# the decision threshold on logit 0 chosen on held-out trajectories, with their scores
THRESHOLD_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "regrid_cnn_threshold.json")


def _threshold():
    """The committed network's decision threshold (0 without the file)."""
    import json
    if not os.path.exists(THRESHOLD_FILE):
        return 0.0
    return float(json.load(open(THRESHOLD_FILE))["threshold"])

# bubble set-ups of the cfg2 geometry used for training (radius, centre) and
# the one held out for evaluation
TRAIN_VARIANTS = [(56.0, 448.0, 448.0), (64.0, 512.0, 512.0), (72.0, 576.0, 480.0), (48.0, 512.0, 384.0)]
EVAL_VARIANT = (60.0, 480.0, 544.0)


def variant(w, radius, x0, y0):
    return dataclasses.replace(w, radius=radius, x0=x0, y0=y0)


def _positions(rows):
    leaf, inner = set(), set()
    for r in rows:
        if r[0] >= 0:
            (leaf if r[5] else inner).add((int(r[0]), int(r[2]), int(r[3])))
    return leaf, inner


def labels(rows_now, ids_now, rows_later):
    """[n, 2] targets of the leaves ids_now: their region is refined (0) or
    covered by a coarser leaf (1) in the later topology."""
    leaf, inner = _positions(rows_later)
    y = np.zeros((len(ids_now), 2), dtype=np.float32)
    for k, n in enumerate(ids_now):
        r = rows_now[n - 1]
        L, I, J = int(r[0]), int(r[2]), int(r[3])
        if (L, I, J) in inner:
            y[k, 0] = 1.0
        for d in range(1, L + 1):
            if (L - d, I >> d, J >> d) in leaf:
                y[k, 1] = 1.0
                break
    return y


def generate(w, variants, steps, every=HORIZON, horizon=HORIZON, regrid_every=10, precision=32):
    """Network inputs and targets from plain runs: every `every` steps the
    input of every leaf, labelled by the topology `horizon` steps later."""
    X, Y = [], []
    for (R, x0, y0) in variants:
        wv = variant(w, R, x0, y0)
        with Mesh(wv, precision=precision) as m:
            m.setup_rising_bubble()
            snaps = []
            for t in range(0, steps + 1, every):
                if t:
                    m.step(every, regrid_every)
                x, ids = m.cnn_features()
                rows, _ = m.topology()
                snaps.append((x, ids, rows))
        lag = horizon // every
        for k in range(len(snaps) - lag):
            x, ids, rows = snaps[k]
            X.append(x)
            Y.append(labels(rows, ids, snaps[k + lag][2]))
    return np.concatenate(X), np.concatenate(Y)


def run(mesh, nsteps, weights=None, full_every=FULL_EVERY, threshold=None):
    """Advance the mesh nsteps under the learned regrid's cadence
    (amrlbm_step_learned): the full regrid every full_every steps, the plain
    criterion judging every step on the device, and every regrid refining the
    blocks the network flags too (weights; none: the plain criterion alone)."""
    mesh.set_regrid_network(weights, _threshold() if threshold is None else threshold)
    st = mesh.step_learned(nsteps, full_every)
    return {"predicted_regrids": st["full_regrids"], "predicted_blocks": st["predicted_blocks"],
            "misses": st["misses"], "checks": st["judged"]}


def evaluate(w, weights, steps, precision=32, variant_=None, full_every=FULL_EVERY):
    """The learned regrid against the same cadence without the network, and
    against the plain regrid every 10 steps, on workload w (a bubble variant
    of it with variant_ = (radius, x0, y0))."""
    wv = variant(w, *variant_) if variant_ else w
    out = {}
    for name, wts in (("learned", weights), ("no_network", None)):
        with Mesh(wv, precision=precision) as m:
            m.setup_rising_bubble()
            m.step(1, 1 << 30)   # C's ghost ring and U, P of a stepped state, as every regrid sees them
            m.synchronize()
            t0 = time.perf_counter()
            st = run(m, steps, wts, full_every)
            m.synchronize()
            st["wall_s"] = time.perf_counter() - t0
            st["mlups"] = wv.fine_cells * steps / st["wall_s"] / 1e6
            st["leaves"] = m.counts()["nleaf"]
            st["missed_refinement_rate"] = st["misses"] / max(st["checks"], 1)
            st["diag"] = m.output().tolist()
            out[name] = st
    with Mesh(wv, precision=precision) as m:
        m.setup_rising_bubble()
        m.step(1, 1 << 30)
        m.synchronize()
        t0 = time.perf_counter()
        m.step(steps, 10)
        m.synchronize()
        wall = time.perf_counter() - t0
        out["plain_every_10"] = {"wall_s": wall, "mlups": wv.fine_cells * steps / wall / 1e6,
                                 "leaves": m.counts()["nleaf"], "diag": m.output().tolist(),
                                 "changes": m.regrid_stats()["changes"]}
    return out


# ---- training set from the generator -------------------------------------
GEN_L0, GEN_MIN_LEVELS, GEN_GY_MAX, GEN_STEPS = 256, 1, 4.0e-5, 400
GEN_SNAPS = range(100, GEN_STEPS - HORIZON + 1, 10)   # inputs at these steps, labels HORIZON later


def generator_instances(n, seed):
    """n trajectories of the generator (module docstring) drawn from seed."""
    s = Stream(seed)
    out = []
    for q in range(n):
        R = 0.5 * (24.0 + 16.0 * s.uniform())
        x0 = R + 8.0 + (GEN_L0 - 2.0 * R - 16.0) * s.uniform()
        y0 = GEN_L0 * (0.25 + 0.25 * s.uniform())
        gy = 5.0e-6 + (GEN_GY_MAX - 5.0e-6) * s.uniform()
        out.append(Workload(f"g{q}", L0=GEN_L0, min_levels=GEN_MIN_LEVELS, radius=R, x0=x0, y0=y0, gy_step=gy))
    return out


def generator_ensemble(n, seed):
    members = tuple(generator_instances(n, seed))
    return Ensemble(f"gen{n}_s{seed}", members, block_cap=-(-n * 400 // 4096) * 4096)


def generate_ensembles(count, seed0, per=64, keep_near=0.25, keep_far=0.01, precision=32, log=None):
    """Inputs and targets of count trajectories (per at a time in one
    ensemble mesh, seeds seed0, seed0 + 1, ...): at each of GEN_SNAPS steps
    the input of every leaf, labelled by the topology HORIZON steps later.
    Kept, from the instances that stay finite: every positive, keep_near of
    the negatives whose phase field varies across the block (max - min of C
    > 0.01: the interface or near it), and keep_far of the others (uniform
    bulk liquid or gas)."""
    X, Y = [], []
    rng = np.random.default_rng(seed0)
    done = 0
    e = 0
    while done < count:
        ens = generator_ensemble(min(per, count - done), seed0 + e)
        with Mesh(ens, precision=precision) as m:
            m.setup_rising_bubble()
            snaps = {}
            t = 0
            for t_next in sorted(set(GEN_SNAPS) | {t + HORIZON for t in GEN_SNAPS}):
                m.step(t_next - t, 10)
                t = t_next
                x, ids = m.cnn_features()
                rows, _ = m.topology()
                snaps[t] = (x, ids, rows)
            finite = m.output_instances(allow_nonfinite=True)[:, 8] == 0
        bad = int(np.sum(~finite))
        for t in GEN_SNAPS:
            x, ids, rows = snaps[t]
            y = labels(rows, ids, snaps[t + HORIZON][2])
            lev, J = rows[ids - 1, 0], rows[ids - 1, 3]
            inst = (J >> lev) // 2
            c = x[:, 0]
            near = (c.max(axis=(1, 2)) - c.min(axis=(1, 2))) > 0.01
            u = rng.random(len(ids))
            keep = ((y[:, 0] > 0.5) | (near & (u < keep_near)) | (u < keep_far)) & finite[inst]
            X.append(x[keep])
            Y.append(y[keep])
        done += len(ens.members)
        e += 1
        if log:
            log(f"[generate] {done}/{count} trajectories, {sum(len(v) for v in Y)} samples, "
                f"non-finite instances {bad}")
    return np.concatenate(X), np.concatenate(Y)


def scores(x, y, weights, threshold=0.0):
    """Refinement recall and precision of the network on (x, y)."""
    pred = cnn_forward(x, weights)[:, 0] > threshold
    pos = y[:, 0] > 0.5
    tp = int(np.sum(pred & pos))
    return {"samples": int(len(y)), "positives": int(pos.sum()), "threshold": float(threshold),
            "refine_recall": tp / max(int(pos.sum()), 1), "refine_precision": tp / max(int(pred.sum()), 1)}


def best_threshold(x, y, weights):
    """The threshold on logit 0 with the best F1 score of the refinement flag on (x, y)."""
    z = cnn_forward(x, weights)[:, 0]
    pos = y[:, 0] > 0.5
    best, best_f1 = 0.0, -1.0
    for t in np.quantile(z, np.linspace(0.5, 0.9999, 400)):
        pred = z > t
        tp = float(np.sum(pred & pos))
        f1 = 2.0 * tp / max(float(pred.sum() + pos.sum()), 1.0)
        if f1 > best_f1:
            best, best_f1 = float(t), f1
    return best


TRAIN_TRAJECTORIES, TRAIN_SEED, HELDOUT_SEED = 5000, 1000, 9000


def train(trajectories=TRAIN_TRAJECTORIES, epochs=20, seed=1, pos_weight=10.0, log=print):
    X, Y = generate_ensembles(trajectories, TRAIN_SEED, log=log)
    pos = Y.mean(axis=0)
    pw = float(pos_weight)
    weights, hist = cnn_train(X, Y, cnn_init(seed), epochs=epochs, batch=256, lr=2e-3, pos_weight=pw, seed=seed)
    info = {"trajectories": trajectories, "positive_rate": pos.tolist(), "pos_weight": pw,
            "loss": [float(v) for v in hist], "train": scores(X, Y, weights)}
    del X, Y
    # the threshold from one held-out draw, its scores on another
    Xv, Yv = generate_ensembles(256, HELDOUT_SEED, log=log)
    thr = best_threshold(Xv, Yv, weights)
    info["threshold"] = thr
    info["validation"] = scores(Xv, Yv, weights, thr)
    del Xv, Yv
    Xh, Yh = generate_ensembles(256, HELDOUT_SEED + 1000, log=log)
    info["held_out"] = scores(Xh, Yh, weights, thr)
    info["held_out_threshold_0"] = scores(Xh, Yh, weights, 0.0)
    return weights, info


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "train":
        import json
        wts, info = train()
        np.save(WEIGHTS, wts)
        json.dump({"threshold": info["threshold"], "held_out": info["held_out"]}, open(THRESHOLD_FILE, "w"), indent=1)
        print(info, flush=True)
        print(evaluate(WORKLOADS["cfg1"], wts, 2000), flush=True)
