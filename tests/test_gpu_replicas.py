"""Independent copies of a rising-bubble domain in one mesh
(amrlbm_setup_replicas): every step launch covers the leaves of all copies,
and walls separate them, so each copy must evolve exactly as the single
domain does. The single domain itself matches the oracle bit for bit
(test_gpu_parity.py); here every leaf of every copy is compared bitwise with
the leaf at the same (level, I, J) of a single-domain run on the same GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm as mod
    mod.lib()
    return mod


def _leaves_by_position(mesh):
    rows, _ = mesh.topology()
    out = {}
    for n in range(1, rows.shape[0] + 1):
        r = rows[n - 1]
        if r[0] >= 0 and r[5]:
            out[(int(r[0]), int(r[2]), int(r[3]))] = n
    return out


@pytest.mark.parametrize("name,single,steps,k", [("cfg1x2", "cfg1", 200, 5), ("cfg2x16", "cfg2", 60, 10)])
def test_replicas_match_single_instance(amrlbm, workloads, name, single, steps, k):
    w, w1 = amrlbm.WORKLOADS[name], amrlbm.WORKLOADS[single]
    print(f"[replicas] {name}: {w.instances} copies of {single}, {steps} steps, regrid every {k}", flush=True)
    with amrlbm.Mesh(w1) as a, amrlbm.Mesh(w) as b:
        a.setup_rising_bubble()
        b.setup_rising_bubble()
        a.step(steps, k)
        b.step(steps, k)
        ca, cb = a.counts(), b.counts()
        print(f"[replicas] single: {ca}  copies: {cb}", flush=True)
        assert cb["nleaf"] == w.instances * ca["nleaf"]
        assert (cb["levelmin"], cb["levelmax"], cb["T"]) == (ca["levelmin"], ca["levelmax"], ca["T"])
        la, lb = _leaves_by_position(a), _leaves_by_position(b)
        roots = w.nroots // w.instances
        nbits = 0
        for (lev, I, J), n in la.items():
            ref = a.get_block(n)
            for t in range(w.instances):
                m = lb[(lev, I, J + t * (roots << lev))]
                got = b.get_block(m)
                for f in ("C", "P", "Ux", "Uy"):
                    assert np.array_equal(got[f], ref[f]), (t, lev, I, J, f)
                for f in ("h", "g"):   # the fused stream leaves the previous ghost ring: interiors
                    assert np.array_equal(got[f][:, 1:-1, 1:-1], ref[f][:, 1:-1, 1:-1]), (t, lev, I, J, f)
                nbits += 1
        print(f"[replicas] {name}: {nbits} leaf copies bitwise equal to the single domain")
        sa, sb = a.regrid_stats(), b.regrid_stats()
        assert sb["changes"] == sa["changes"] and sb["checks"] == sa["checks"]
