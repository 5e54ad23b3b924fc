"""Ghost programs are kept by block id and rebuilt after a regrid only where a
block they were built from changed. With AMRLBM_VERIFY_PROGRAMS set, every
table build also builds every leaf's programs afresh and fails if a reused
one differs. These runs cross several mesh changes (refinements and
derefinements) with the check on and end in the same state as the run
without it (the fused-step parity tests hold such runs to the oracle)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amrlbm():
    import amrlbm as mod
    mod.lib()
    return mod


def _setup(m, w, bubbles):
    if bubbles:
        m.setup_bubbles(*zip(*bubbles))
    else:
        m.setup_rising_bubble()


@pytest.mark.parametrize("name,steps,k,bubbles", [
    ("cfg2", 2000, 10, None),
    ("cfg2", 1500, 5, [(48.0, 352.0, 400.0), (56.0, 672.0, 560.0)]),   # (R, x0, y0)
])
def test_reused_programs_equal_fresh_builds(amrlbm, workloads, monkeypatch, name, steps, k, bubbles):
    w = workloads[name]
    print(f"[program cache] {name} {bubbles or 'rising bubble'}: {steps} steps, regrid every {k}, "
          "verifying reused programs", flush=True)
    with amrlbm.Mesh(w) as a:
        _setup(a, w, bubbles)
        monkeypatch.setenv("AMRLBM_VERIFY_PROGRAMS", "1")
        a.step(steps, k)
        monkeypatch.delenv("AMRLBM_VERIFY_PROGRAMS")
        st = a.regrid_stats()
        print(f"[program cache] {st['changes']} mesh changes, {a.counts()['nleaf']} leaves at the end")
        assert st["changes"] >= 2
        da = a.output()
    # the same run with no check: the programs are the same, so the state is too
    with amrlbm.Mesh(w) as b:
        _setup(b, w, bubbles)
        b.step(steps, k)
        assert np.array_equal(da, b.output())
