import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "amr-lbm-openmp-2d_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libamrlbm.so)")


def snapshot_rows(snap):
    """21-int topology rows of a Snapshot, the C-ABI's block table format."""
    return snap.topology().astype(np.int32)


def load_into_mesh(mesh, snap):
    """Copy an oracle/reference snapshot into a library mesh."""
    mesh.set_topology(snapshot_rows(snap), snap.kk, snap.free, snap.T)
    for n, b in snap.blocks.items():
        if b.h is not None:
            mesh.put_block(n, h=b.h, g=b.g, C=b.C, P=b.P, Ux=b.Ux, Uy=b.Uy)


def mesh_state(mesh):
    """Topology rows, free list, kk and leaf arrays of a library mesh."""
    rows, free = mesh.topology()
    c = mesh.counts()
    arrays = {}
    for n in range(1, rows.shape[0] + 1):
        if rows[n - 1, 0] >= 0 and rows[n - 1, 5]:
            arrays[n] = mesh.get_block(n)
    return rows, free, c["kk"], arrays


def compare_state(mesh, snap, pop_ghosts=True, c_ring2=True, tol=1e-12, bit_exact=True):
    """Compare a library mesh with a Snapshot: topology exactly, arrays to a
    relative tolerance |a-b| <= tol * max(|b|, 1) and, with bit_exact (the FP64
    path's claim), every value bit for bit. Returns (max_rel, nbitdiff)."""
    rows, free, kk, arrays = mesh_state(mesh)
    ref_rows = snapshot_rows(snap)
    assert rows.shape == ref_rows.shape, f"Nblks {rows.shape[0]} != {ref_rows.shape[0]}"
    bad = np.argwhere(rows != ref_rows)
    assert bad.size == 0, f"topology differs at (block-1, field) {bad[:5].tolist()}"
    assert list(free) == list(snap.free), "free list differs"
    assert kk == snap.kk, "KK differs"
    worst, nbit = 0.0, 0
    for n, a in arrays.items():
        b = snap.blocks[n]
        for f in ("h", "g", "C", "P", "Ux", "Uy"):
            x, y = a[f], getattr(b, f)
            if f in ("h", "g") and not pop_ghosts:
                x, y = x[:, 1:-1, 1:-1], y[:, 1:-1, 1:-1]
            if f == "C" and not c_ring2:
                x, y = x[1:-1, 1:-1], y[1:-1, 1:-1]
            d = np.abs(x - y) / np.maximum(np.abs(y), 1.0)
            m = float(d.max())
            assert m <= tol, f"block {n} field {f}: max rel diff {m:.3e}"
            worst = max(worst, m)
            nb = int(np.count_nonzero(np.ascontiguousarray(x).view(np.int64) !=
                                      np.ascontiguousarray(y).view(np.int64)))
            assert not (bit_exact and nb), f"block {n} field {f}: {nb} values differ in some bit"
            nbit += nb
    return worst, nbit


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def workloads():
    from amrlbm.workloads import WORKLOADS
    return WORKLOADS
