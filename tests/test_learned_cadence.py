"""The learned regrid's cadence (amrlbm.learned.run), driven on the host with
a stand-in mesh that records its calls: a full regrid every 50 steps, the
plain judge every 10 steps in between, and never two regrids without a step
between them, also when the run is split into several calls. Err_Interface
reads C's ghost ring, which only the step fills, and a block just derefined
holds a ring of zeros (Restrict.f90:23-30), so a regrid right after another
would refine every such block again."""
import pytest


class RecordingMesh:
    """Records step / check / regrid calls; the judge asks for a regrid at
    the steps in refine_at."""

    def __init__(self, refine_at=()):
        self.t = 0
        self.calls = []
        self.refine_at = set(refine_at)

    def counts(self):
        return {"T": self.t}

    def step(self, n, regrid_every):
        assert n > 0 and regrid_every >= 1 << 30
        self.t += n
        self.calls.append(("step", self.t))

    def regrid_check(self):
        self.calls.append(("check", self.t))
        hit = self.t in self.refine_at
        return hit, hit

    def refine_derefine_mesh(self):
        self.calls.append(("regrid", self.t))

    def refine_derefine_predicted(self, ids):
        assert ids.size == 0
        self.calls.append(("predicted", self.t))


def back_to_back(calls):
    """Whether two regrids follow each other with no step between them."""
    prev = None
    for kind, _ in calls:
        if kind == "check":
            continue
        if kind != "step" and prev is not None and prev != "step":
            return True
        prev = kind
    return False


def test_back_to_back_detector():
    assert back_to_back([("step", 10), ("check", 10), ("regrid", 10), ("predicted", 10)])
    assert not back_to_back([("predicted", 0), ("step", 10), ("check", 10), ("regrid", 10), ("step", 20)])


def test_single_call_cadence():
    from amrlbm import learned
    m = RecordingMesh(refine_at=range(0, 601, 10))
    st = learned.run(m, 600, None)
    predicted = [t for k, t in m.calls if k == "predicted"]
    checks = [t for k, t in m.calls if k == "check"]
    assert predicted == list(range(0, 600, 50))
    assert checks == [t for t in range(10, 601, 10) if t % 50]
    assert st["checks"] == 48 and st["misses"] == 48 and st["predicted_regrids"] == 12
    assert not back_to_back(m.calls)


@pytest.mark.parametrize("chunks", [[300, 300], [500, 100], [290, 310], [55, 45, 500], [7, 3, 40, 550]])
@pytest.mark.parametrize("every", [1, 10, 7])
def test_split_calls_never_regrid_twice_in_a_row(chunks, every):
    """With the judge asking for a regrid at every check (every step, or some
    of them), calls of any length never put two regrids next to each other,
    and every call's steps add up."""
    from amrlbm import learned
    m = RecordingMesh(refine_at=range(0, 601, every))
    for n in chunks:
        learned.run(m, n, None)
    assert m.t == 600
    assert not back_to_back(m.calls)
    assert m.calls[0] == ("predicted", 0)
