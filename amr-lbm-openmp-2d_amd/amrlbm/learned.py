"""Learned regrid: a convolutional network over each leaf block's phase field
and velocity predicts whether the block's region will be refined (or
derefined) HORIZON steps ahead, so that blocks the interface is about to
enter are refined before it arrives and the full regrid can run every
HORIZON steps. The plain gradient criterion (Refine_Derefine.f90:18-54,
:327-352) stays the judge: it is evaluated on the device every `judge_every`
steps, and any interface cell on an under-refined block triggers an
immediate plain regrid (a miss).

The network, its training and inference are the library's HIP kernels
(amrlbm_cnn_*); this module builds the training set from the solver's own
runs, trains, and runs the learned regrid.

    python -m amrlbm.learned train    # writes regrid_cnn_weights.npy beside this file
"""
import dataclasses
import os
import sys
import time

import numpy as np

from . import Mesh, cnn_forward, cnn_init, cnn_train
from .workloads import WORKLOADS

HORIZON = 50
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "regrid_cnn_weights.npy")

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


def run(mesh, nsteps, weights=None, predict_every=HORIZON, judge_every=10, threshold=0.0):
    """Advance the mesh nsteps with the full regrid every predict_every steps
    (the plain criterion plus, with weights, the blocks the network flags for
    refinement) and the plain criterion judging every judge_every steps.

    Two regrids never follow each other without a step between them.
    Err_Interface reads C's ghost ring, which only the step fills, and a
    block just derefined holds a ring of zeros (Restrict.f90:23-30), so a
    second regrid at once would refine every such block again from the
    spurious jump and prolong C from the zero ring into its children. The
    judge is therefore not consulted at a multiple of predict_every, where
    the full regrid comes next (in this call or at the start of the next),
    and a call that starts at the step of the last regrid this function ran
    on the mesh (a judge's, ending the previous call) skips its first full
    regrid."""
    st = {"predicted_regrids": 0, "predicted_blocks": 0, "misses": 0, "checks": 0}
    t = 0
    while t < nsteps:
        if t % predict_every == 0 and getattr(mesh, "_last_regrid_step", None) != mesh.counts()["T"]:
            ids = np.zeros(0, dtype=np.int32)
            if weights is not None:
                x, all_ids = mesh.cnn_features()
                ids = all_ids[cnn_forward(x, weights)[:, 0] > threshold]
            mesh.refine_derefine_predicted(ids)
            mesh._last_regrid_step = mesh.counts()["T"]
            st["predicted_regrids"] += 1
            st["predicted_blocks"] += int(ids.size)
        n = min(judge_every, nsteps - t, predict_every - t % predict_every)
        mesh.step(n, 1 << 30)
        t += n
        if t % predict_every == 0:
            continue
        _, refine = mesh.regrid_check()
        st["checks"] += 1
        if refine:
            mesh.refine_derefine_mesh()
            mesh._last_regrid_step = mesh.counts()["T"]
            st["misses"] += 1
    return st


def evaluate(w, weights, steps, precision=32, variant_=EVAL_VARIANT):
    """The learned regrid against the same cadence without the network, and
    against the plain regrid every 10 steps, on a held-out bubble."""
    wv = variant(w, *variant_)
    out = {}
    for name, wts in (("learned", weights), ("no_network", None)):
        with Mesh(wv, precision=precision) as m:
            m.setup_rising_bubble()
            t0 = time.perf_counter()
            st = run(m, steps, wts)
            m.synchronize()
            st["wall_s"] = time.perf_counter() - t0
            st["mlups"] = wv.fine_cells * steps / st["wall_s"] / 1e6
            st["leaves"] = m.counts()["nleaf"]
            st["diag"] = m.output().tolist()
            out[name] = st
    with Mesh(wv, precision=precision) as m:
        m.setup_rising_bubble()
        t0 = time.perf_counter()
        m.step(steps, 10)
        m.synchronize()
        wall = time.perf_counter() - t0
        out["plain_every_10"] = {"wall_s": wall, "mlups": wv.fine_cells * steps / wall / 1e6,
                                 "leaves": m.counts()["nleaf"], "diag": m.output().tolist(),
                                 "changes": m.regrid_stats()["changes"]}
    return out


def train(steps=3000, epochs=40, seed=1, pos_weight=None):
    w = WORKLOADS["cfg2"]
    X, Y = generate(w, TRAIN_VARIANTS, steps)
    pos = Y.mean(axis=0)
    pw = pos_weight if pos_weight is not None else float(min(50.0, (1 - pos[0]) / max(pos[0], 1e-4)))
    weights, hist = cnn_train(X, Y, cnn_init(seed), epochs=epochs, batch=256, lr=2e-3, pos_weight=pw, seed=seed)
    pred = cnn_forward(X, weights) > 0
    tp = int(np.sum(pred[:, 0] & (Y[:, 0] > 0.5)))
    info = {"samples": int(X.shape[0]), "positive_rate": pos.tolist(), "pos_weight": pw,
            "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
            "refine_recall": tp / max(int(np.sum(Y[:, 0] > 0.5)), 1),
            "refine_precision": tp / max(int(np.sum(pred[:, 0])), 1)}
    return weights, info


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "train":
        wts, info = train()
        np.save(WEIGHTS, wts)
        print(info)
        print(evaluate(WORKLOADS["cfg2"], wts, 2000))
