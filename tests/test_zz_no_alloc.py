"""amrlbm_output_noalloc (the diagnostics with their per-leaf partials in the
library's static device scratch, no allocation) against amrlbm_output on the
same state: bit for bit, on two calls in a row, the second on a smaller mesh,
FP64 and FP32."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm as mod
    mod.lib()
    return mod


@pytest.mark.parametrize("precision", [64, 32])
def test_output_noalloc_equals_output(amrlbm, workloads, precision):
    for name, steps in (("cfg2", 60), ("tiny", 30)):   # the second call on fewer leaves
        w = workloads[name]
        with amrlbm.Mesh(w, precision=precision) as m:
            m.setup_rising_bubble()
            m.step(steps, 10)
            ref = m.output()
            got = m.output_noalloc()
            print(f"[noalloc] FP{precision} {name}: {m.counts()['nleaf']} leaves, mass {got[3]:.15e}", flush=True)
            assert np.array_equal(got, ref), (got, ref)
            again = m.output_noalloc()
            assert np.array_equal(again, ref)


def test_output_noalloc_after_mesh_change(amrlbm, workloads):
    """Across a mesh change (leaf list and tables replaced) and after another
    mesh's call, the scratch's old contents do not leak into the result."""
    w = workloads["tiny_rise"]
    with amrlbm.Mesh(w) as m:
        m.setup_rising_bubble()
        m.step(1100, 1)   # the first refinement is at step 1077
        assert m.regrid_stats()["changes"] >= 1
        assert np.array_equal(m.output_noalloc(), m.output())
