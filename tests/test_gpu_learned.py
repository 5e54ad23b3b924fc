"""The learned regrid on the solver: training data from the solver's own runs
(the generator of amrlbm.learned), training on the GPU, the committed
network's recall and precision on held-out trajectories, and the learned
cadence with every decision on the device (amrlbm_step_learned): the plain
criterion judges every step and leaves no interface on an under-refined
block, calls split anywhere give the same run, and on the 128x256 bubble (the
network trained on the generator, deployed there) the missed-refinement rate
and the gain over the cadence without the network are reported."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm as mod
    mod.lib()
    return mod


def test_labels_from_topology(amrlbm, workloads):
    """A leaf refined later is labelled (1, 0); the children of a block that is
    a leaf later are labelled (0, 1)."""
    from amrlbm import learned
    # 1 root split into 4 children (ids 2..5); later: child 2 refined, children 3..5 unchanged
    now = np.zeros((5, 21), dtype=np.int32)
    now[0, :6] = [0, 16, 0, 0, -1, 0]
    for k, (I, J) in enumerate(((0, 0), (1, 0), (0, 1), (1, 1))):
        now[1 + k, :6] = [1, 8, I, J, 1, 1]
    later = now.copy()
    later[1, 5] = 0   # block 2 became a parent
    y = learned.labels(now, [2, 3, 4, 5], later)
    assert y.tolist() == [[1, 0], [0, 0], [0, 0], [0, 0]]
    coarse = now.copy()
    coarse[0, 5] = 1
    coarse[1:, 0] = -1   # children freed: the root is a leaf again
    y = learned.labels(now, [2, 3, 4, 5], coarse)
    assert y.tolist() == [[0, 1]] * 4


def test_generator_data_and_training(amrlbm):
    """The generator's trajectories stay finite and hold refinement
    positives; training on them lowers the loss."""
    from amrlbm import learned
    X, Y = learned.generate_ensembles(128, 4242)
    npos = int(np.sum(Y[:, 0] > 0.5))
    print(f"[learned] generator: {X.shape[0]} samples, {npos} refinement positives", flush=True)
    assert X.shape[1:] == (3, 18, 18) and np.isfinite(X).all()
    assert npos >= 20
    w, hist = amrlbm.cnn_train(X, Y, amrlbm.cnn_init(2), epochs=4, batch=256, lr=2e-3, pos_weight=10.0)
    print(f"[learned] loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    assert hist[-1] < hist[0]


def test_committed_network_recall_on_held_out_trajectories(amrlbm):
    """The committed network at its committed threshold, on generator
    trajectories of a seed no training or threshold choice used: refinement
    recall and precision (held-out 0.58 and 0.49 when committed)."""
    from amrlbm import learned
    assert os.path.exists(learned.WEIGHTS) and os.path.exists(learned.THRESHOLD_FILE)
    X, Y = learned.generate_ensembles(256, 7777)
    sc = learned.scores(X, Y, np.load(learned.WEIGHTS), learned._threshold())
    print(f"[learned] held-out: {sc}", flush=True)
    assert sc["positives"] >= 40
    assert sc["refine_recall"] >= 0.35
    assert sc["refine_precision"] >= 0.25


def test_learned_cadence_judges_every_step(amrlbm, workloads):
    """On the faster-rising 128x256 bubble the judge regrids at once when the
    interface reaches an under-refined block: after the run the plain
    criterion asks for no refinement, every step was judged, and the mass
    stays within FP32 rounding of the plain regrid every 10 steps."""
    from amrlbm import learned
    wts = np.load(learned.WEIGHTS)
    w = workloads["cfg1_rise"]
    with amrlbm.Mesh(w, precision=32) as m, amrlbm.Mesh(w, precision=32) as p:
        m.setup_rising_bubble()
        p.setup_rising_bubble()
        st = learned.run(m, 2200, wts)
        p.step(2200, 10)
        print(f"[learned] cfg1_rise 2200 steps: {st}", flush=True)
        assert st["checks"] == 2200 and st["predicted_regrids"] == 2200 // learned.FULL_EVERY
        assert m.regrid_check()[1] is False
        dm, dp = m.output(), p.output()
        assert dm[8] == 0 and abs(dm[3] - dp[3]) <= 1e-5 * abs(dp[3])


def test_network_flags_join_the_regrid(amrlbm, workloads):
    """The network's flags reach the regrid: with a threshold every logit
    passes and no Err_Interface gate, every leaf below the finest level that
    the interface's profile touches is flagged at each mesh change, so the
    run refines blocks the plain criterion leaves alone and ends with more
    leaves than the same cadence without the network."""
    from amrlbm import learned
    wts = np.load(learned.WEIGHTS)
    w = workloads["cfg1_rise"]
    with amrlbm.Mesh(w, precision=32) as m, amrlbm.Mesh(w, precision=32) as p:
        m.setup_rising_bubble()
        p.setup_rising_bubble()
        m.set_regrid_network(wts, threshold=-1e30, gate=0.0)
        p.set_regrid_network(None)
        st = m.step_learned(1000, learned.FULL_EVERY)
        sp = p.step_learned(1000, learned.FULL_EVERY)
        print(f"[learned] every logit passing: {st}, {m.counts()['nleaf']} leaves; "
              f"no network: {sp}, {p.counts()['nleaf']} leaves", flush=True)
        assert sp["predicted_blocks"] == 0 and st["predicted_blocks"] > 0
        assert m.counts()["nleaf"] > p.counts()["nleaf"]
        assert m.output()[8] == 0


def test_learned_cadence_split_calls(amrlbm, workloads):
    """The cadence counts steps as amrlbm_step does (T), so calls split
    anywhere give the same run, bit for bit."""
    from amrlbm import learned
    wts = np.load(learned.WEIGHTS)
    w = workloads["cfg1_rise"]
    states = []
    for chunks in ([900], [337, 63, 500], [50, 1, 849]):
        with amrlbm.Mesh(w) as m:
            m.setup_rising_bubble()
            for n in chunks:
                learned.run(m, n, wts)
            states.append((m.topology()[0], {n: m.get_block(n) for n in m.leaves()}))
    for rows, arr in states[1:]:
        assert np.array_equal(rows, states[0][0])
        for n in arr:
            for f in ("C", "P", "Ux", "Uy"):
                assert np.array_equal(arr[n][f], states[0][1][n][f]), (n, f)


def test_deployed_on_cfg1_bubble(amrlbm, workloads):
    """Deployed on the 128x256 bubble (D = 32) as the learned-regrid
    configuration states it: full regrid every 50 steps, the plain criterion
    judging every step and the network's flags joining every regrid, over
    10,000 steps (the mesh changes as the bubble rises): the missed-refinement
    rate (judge regrids per judged step) and the throughput against the same
    cadence without the network and the plain regrid every 10 steps are
    reported; the misses stay under 1 % of the judged steps."""
    from amrlbm import learned
    ev = learned.evaluate(workloads["cfg1"], np.load(learned.WEIGHTS), 10000, full_every=50)
    for k, v in ev.items():
        print(f"[learned] cfg1 {k}: " + ", ".join(f"{f} {x:.4g}" if isinstance(x, float) else f"{f} {x}"
                                                for f, x in v.items() if f != "diag"), flush=True)
    assert ev["learned"]["missed_refinement_rate"] <= 0.01
    assert ev["learned"]["checks"] == 10000
    base = np.array(ev["plain_every_10"]["diag"])
    d = np.array(ev["learned"]["diag"])
    assert d[8] == 0 and abs(d[3] - base[3]) <= 1e-5 * abs(base[3])
