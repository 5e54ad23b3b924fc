"""The learned regrid on the solver: training data built from the solver's own
runs, training on the GPU, and the learned regrid with the plain criterion as
judge. What the network gains over the same regrid cadence without it is
printed (misses: regrids the judge had to trigger; throughput); the test
asserts the parts a user relies on: labels, training progress, and that the
judge leaves no interface on an under-refined block."""
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


def test_training_on_solver_data(amrlbm, workloads):
    from amrlbm import learned
    X, Y = learned.generate(workloads["cfg1"], [(16.0, 64.0, 64.0), (14.0, 60.0, 56.0)], 1500, precision=32)
    print(f"[learned] cfg1 data: {X.shape[0]} samples, positive rate {Y.mean(axis=0)}")
    assert X.shape[1:] == (3, 18, 18) and np.isfinite(X).all()
    w, hist = amrlbm.cnn_train(X, Y, amrlbm.cnn_init(2), epochs=10, batch=64, lr=2e-3, pos_weight=5.0)
    print(f"[learned] loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    assert hist[-1] < hist[0]


def test_learned_regrid_with_judge(amrlbm, workloads):
    from amrlbm import learned
    assert os.path.exists(learned.WEIGHTS)
    wts = np.load(learned.WEIGHTS)
    ev = learned.evaluate(workloads["cfg2"], wts, 600)
    for k, v in ev.items():
        print(f"[learned] {k}: " + ", ".join(f"{f} {x:.4g}" if isinstance(x, float) else f"{f} {x}"
                                           for f, x in v.items() if f != "diag"))
    base = np.array(ev["plain_every_10"]["diag"])
    for k in ("learned", "no_network"):
        d = np.array(ev[k]["diag"])
        assert d[8] == 0
        assert abs(d[3] - base[3]) <= 1e-5 * abs(base[3])
        assert ev[k]["checks"] == 48     # every 10 steps but at the 12 full regrids (every 50)
    # after the last judge check (at 290) and one step to fill C's ghost ring,
    # the plain criterion asks for no refinement
    with amrlbm.Mesh(learned.variant(workloads["cfg2"], *learned.EVAL_VARIANT), precision=32) as m:
        m.setup_rising_bubble()
        learned.run(m, 290, wts)
        m.step(1, 1 << 30)
        assert m.regrid_check()[1] is False
