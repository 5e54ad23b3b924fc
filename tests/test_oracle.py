"""The oracle (oracle/amroracle.c) pinned against the reference build's golden
vectors and against closed forms. CPU only.

The golden vectors are SHA-256 hashes of every snapshot the reference build
wrote (tests/golden/make_golden.py); the oracle writes the same file layout,
so equal hashes mean bit-identical topology, free list and arrays, ghosts
included.
"""
import ctypes
import hashlib
import json
import os
import tempfile

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HASHES = json.load(open(os.path.join(GOLDEN, "reference_hashes.json")))
PHASES = ["c2", "collide", "hg", "stream", "macro_h", "c", "macro_g"]


def snap_hash(o):
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
        path = f.name
    try:
        o.save(path)
        return hashlib.sha256(open(path, "rb").read()).hexdigest()
    finally:
        os.unlink(path)


def test_tiny_every_phase_matches_reference(oracle_mod, workloads):
    """tiny: every phase of steps 0, 1 and 10 (regrid every step), bit for bit."""
    g = HASHES["tiny"]
    o = oracle_mod.Oracle(workloads["tiny"])
    o.setup()
    assert snap_hash(o) == g["init"]
    for t in range(g["steps"]):
        o.set_step(t)
        o.phase("levels")
        dump = t in g["phase_steps"]
        if dump:
            assert snap_hash(o) == g["phases"][f"s{t}_p0"], f"step {t} start"
        for k, ph in enumerate(PHASES, start=1):
            o.phase(ph)
            if dump:
                assert snap_hash(o) == g["phases"][f"s{t}_p{k}"], f"step {t} after {ph}"
        o.regrid()
        if dump:
            assert snap_hash(o) == g["phases"][f"s{t}_p8"], f"step {t} after regrid"
    o.set_step(g["steps"])
    o.phase("levels")
    assert snap_hash(o) == g["final"]


@pytest.mark.parametrize("name", ["cfg1", "cfg2"])
def test_runs_match_reference(oracle_mod, workloads, name):
    g = HASHES[name]
    o = oracle_mod.Oracle(workloads[name])
    o.setup()
    assert snap_hash(o) == g["init"]
    o.run(g["steps"], g["regrid_every"])
    assert snap_hash(o) == g["final"]


def replay_with_phase_hashes(oracle_mod, w, g, check=True):
    """Run sequence g of the reference build on the oracle: the set-up, whole
    steps, and phase by phase at each step whose regrid changes the mesh,
    comparing the snapshot hashes there. Returns the oracle mesh at the end
    and the (Nblks, nLeaf, KK) before and after each such regrid."""
    o = oracle_mod.Oracle(w)
    o.setup()
    if check:
        assert snap_hash(o) == g["init"]
    k = g["regrid_every"]
    moves = []
    for t in range(g["steps"]):
        if t not in g["change_steps"]:
            o.step(k)
            continue
        o.set_step(t)
        o.phase("levels")
        if check:
            assert snap_hash(o) == g["phases"][f"s{t}_p0"], f"step {t} start"
        for p, ph in enumerate(PHASES, start=1):
            o.phase(ph)
            if check:
                assert snap_hash(o) == g["phases"][f"s{t}_p{p}"], f"step {t} after {ph}"
        before = o.snapshot()
        if (t + 1) % k == 0:
            o.regrid()
        after = o.snapshot()
        if check:
            assert snap_hash(o) == g["phases"][f"s{t}_p8"], f"step {t} after regrid"
        moves.append(((before.nblks, before.nleaf, before.kk), (after.nblks, after.nleaf, after.kk),
                      before.leaves != after.leaves))
        o.set_step(t + 1)
    o.phase("levels")
    if check:
        assert snap_hash(o) == g["final"]
    return o, moves


@pytest.mark.parametrize("name", ["tiny_rise", "cfg1_rise", "cfg2_regrid"])
def test_mesh_changing_runs_match_reference(oracle_mod, workloads, name):
    """Runs whose mesh changes while stepping: every phase of each step whose
    regrid changes the mesh (refinement, derefinement, reuse of freed ids,
    prolong and restrict of stepped fields, Err_Interface from the C fill),
    ghosts included, and the final state, bit for bit."""
    g = HASHES["rise"][name]
    w = workloads[g["workload"]]
    assert (w.gy_step or None) == g["gy_step"]
    _, moves = replay_with_phase_hashes(oracle_mod, w, g)
    assert all(changed for _, _, changed in moves), moves
    print(name, moves)


def test_mesh_changing_runs_cover_refine_derefine_and_reuse(oracle_mod, workloads):
    """The golden mesh-changing runs hold at least one refinement, one
    derefinement and one refinement that takes freed ids from the FIFO."""
    kinds = set()
    for name in ("tiny_rise", "cfg1_rise"):
        g = HASHES["rise"][name]
        _, moves = replay_with_phase_hashes(oracle_mod, workloads[g["workload"]], g, check=False)
        for (nb0, nl0, kk0), (nb1, nl1, kk1), _ in moves:
            if nl1 > nl0 and kk1 < kk0:
                kinds.add("reuse")
            elif nl1 > nl0:
                kinds.add("refine")
            elif nl1 < nl0:
                kinds.add("derefine")
    assert kinds == {"refine", "derefine", "reuse"}, kinds


def test_regrid_walk_matches_reference(oracle_mod, workloads):
    """Refine, derefine and free-list reuse on every call, prolong/restrict data."""
    g = HASHES["walk"]
    w = workloads[g["workload"]]
    o = oracle_mod.Oracle(w)
    o.setup()
    for k in range(1, g["moves"] + 1):
        o.set_bubble_c(w.L0 // 2 + k * g["step"][0], w.L0 // 2 + k * g["step"][1])
        o.phase("levels")
        o.regrid()
        assert snap_hash(o) == g["snapshots"][k - 1], f"move {k}"


def test_loaded_reference_snapshot_continues_bit_exact(oracle_mod, workloads):
    """A reference snapshot loaded into the oracle (alias structure rebuilt
    from the quadtree) continues exactly as the reference did."""
    o = oracle_mod.Oracle(workloads["tiny"])
    o.load(os.path.join(GOLDEN, "tiny_s1_p3.bin"))
    o.phase("stream")
    assert snap_hash(o) == HASHES["tiny"]["phases"]["s1_p4"]


def _diag_rows(path):
    rows = [l.split() for l in open(path) if not l.startswith("#")]
    return {int(r[0]): np.array([float(v) for v in r[4:12]]) for r in rows}


@pytest.mark.parametrize("name,every", [("cfg1", 50), ("cfg2", 10)])
def test_diagnostics_match_reference(oracle_mod, workloads, name, every):
    """Mass, rho range, max|Uy|, bubble area, centroid and rise velocity of the
    oracle against the reference build's printout (full precision)."""
    ref = _diag_rows(os.path.join(GOLDEN, f"{name}_diag.txt"))
    g = HASHES[name]
    o = oracle_mod.Oracle(workloads[name])
    o.setup()
    for t in range(0, g["steps"] + 1, every):
        d = o.diagnostics()
        # reference order: rho_min rho_max umax mass area xc yc vy
        mine = np.array([d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7]])
        np.testing.assert_allclose(mine, ref[t], rtol=1e-13, atol=1e-300, err_msg=f"T={t}")
        if t < g["steps"]:
            o.run(every, g["regrid_every"])


# ---------------------------------------------------------------------------
# closed forms (pins that do not use the reference build)
# ---------------------------------------------------------------------------

def _P(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def test_mrt_with_unit_rates_returns_equilibrium(oracle_mod):
    """MRT_Collision_AMR with s9 = 1 relaxes every moment fully: g -> geq,
    which holds only if Reconvt_D2Q9_standard inverts Convert_D2Q9_standard."""
    rng = np.random.default_rng(1)
    L = oracle_mod.lib()
    for _ in range(20):
        g = rng.normal(size=9)
        geq = rng.normal(size=9)
        L.orc_mrt_collision(1.0, _P(geq), _P(g))
        np.testing.assert_allclose(g, geq, rtol=0, atol=1e-14)


def test_mrt_conserves_first_moments(oracle_mod):
    """Moments 1-7 relax at rate 1 and 8-9 at s9: the density moment of g-geq
    is removed exactly, whatever s9."""
    rng = np.random.default_rng(2)
    L = oracle_mod.lib()
    for s9 in (0.3, 0.9, 1.7):
        g = rng.normal(size=9)
        geq = g + rng.normal(size=9) * 1e-3
        geq += (g.sum() - geq.sum()) / 9.0          # same zeroth moment
        g0 = g.copy()
        L.orc_mrt_collision(s9, _P(geq), _P(g))
        assert abs(g.sum() - g0.sum()) < 1e-14


def _poly2(x, y, c):
    """biquadratic: sum c[i,j] x^i y^j, i,j <= 2"""
    return sum(c[i, j] * x ** i * y ** j for i in range(3) for j in range(3))


def test_biquadratic_from_fine_exact_on_biquadratics(oracle_mod):
    """Biquadratic_from_Fine reproduces biquadratic data at the coarse centres."""
    rng = np.random.default_rng(3)
    c = rng.normal(size=(3, 3))
    L = oracle_mod.lib()
    i2, j2, X2, Y2 = 4, 32, 2, 16
    # fine cells i=1..i2 at x = i - 0.5, coarse cells X at x = 2X - 1
    xf = np.arange(1, i2 + 1) - 0.5
    yf = np.arange(1, j2 + 1) - 0.5
    fin = np.asfortranarray(_poly2(xf[:, None], yf[None, :], c))
    out = np.zeros((X2, Y2), order="F")
    L.orc_biquadratic_from_fine(i2, j2, X2, Y2, _P(fin), _P(out))
    xc = 2.0 * np.arange(1, X2 + 1) - 1.0
    yc = 2.0 * np.arange(1, Y2 + 1) - 1.0
    np.testing.assert_allclose(out, _poly2(xc[:, None], yc[None, :], c), rtol=1e-12, atol=1e-12)


def test_biquadratic_from_coarse_exact_on_biquadratics(oracle_mod):
    """Biquadratic_from_Coarse reproduces biquadratic data at the fine centres
    a quarter coarse cell off the coarse centres."""
    rng = np.random.default_rng(4)
    c = rng.normal(size=(3, 3))
    L = oracle_mod.lib()
    I2, J2, x2, y2 = 3, 10, 2, 16
    # coarse cell I at X = I (unit spacing); fine cell x sits at (x+1)//2 + 1 -/+ 1/4
    XC = np.arange(1, I2 + 1, dtype=float)
    YC = np.arange(1, J2 + 1, dtype=float)
    fin = np.asfortranarray(_poly2(XC[:, None], YC[None, :], c))
    out = np.zeros((x2, y2), order="F")
    L.orc_biquadratic_from_coarse(I2, J2, x2, y2, _P(fin), _P(out))
    xs = np.array([(x + 1) // 2 + 1 + (-0.25 if x % 2 else 0.25) for x in range(1, x2 + 1)])
    ys = np.array([(y + 1) // 2 + 1 + (-0.25 if y % 2 else 0.25) for y in range(1, y2 + 1)])
    np.testing.assert_allclose(out, _poly2(xs[:, None], ys[None, :], c), rtol=1e-12, atol=1e-12)


def test_prolong_interpolation_exact_on_bilinear_data(oracle_mod):
    """Bilinear_Interpolate_Extrapolate (Prolong.f90:43-70) is exact for
    bilinear data, edge extrapolation included."""
    L = oracle_mod.lib()
    n = 16
    X = np.arange(1, n + 1) - 0.5          # coarse centres, unit cells
    fin = np.asfortranarray(0.3 + 0.7 * X[:, None] - 0.2 * X[None, :] + 0.05 * X[:, None] * X[None, :])
    out = np.zeros((2 * n, 2 * n), order="F")
    L.orc_bilinear_interp_extrap(n, n, _P(fin), _P(out))
    x = (np.arange(1, 2 * n + 1) - 0.5) / 2.0
    ref = 0.3 + 0.7 * x[:, None] - 0.2 * x[None, :] + 0.05 * x[:, None] * x[None, :]
    np.testing.assert_allclose(out, ref, rtol=1e-13, atol=1e-13)


def test_restriction_is_the_2x2_mean(oracle_mod):
    """Bilinear_from_Fine (Interpolations.f90:142-166): exact mass conservation
    on data with a power-of-two denominator."""
    rng = np.random.default_rng(5)
    L = oracle_mod.lib()
    fin = np.asfortranarray(rng.integers(-1000, 1000, size=(32, 32)).astype(float) / 8.0)
    out = np.zeros((16, 16), order="F")
    L.orc_bilinear_from_fine(32, 32, 16, 16, _P(fin), _P(out))
    ref = fin.reshape(16, 2, 16, 2).mean(axis=(1, 3))
    assert np.array_equal(out, ref)
    assert out.sum() * 4 == fin.sum()


def test_lax_wendroff_limits(oracle_mod, workloads):
    """Streaming_LaxWendroff_AMR: cfl = 1 is the exact shift; linear data is
    translated exactly by cfl cells; the rest population is never moved."""
    w = workloads["cfg1"]
    P = oracle_mod.make_params(w)
    L = oracle_mod.lib()
    rng = np.random.default_rng(6)
    f = np.asfortranarray(rng.normal(size=(9, 18, 18)))
    a, b = f.copy(order="F"), f.copy(order="F")
    L.orc_lax_wendroff_block(ctypes.byref(P), 1.0, _P(a))
    L.orc_perfect_shift_block(ctypes.byref(P), _P(b))
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-15)
    ex = [0, 1, 0, -1, 0, 1, -1, -1, 1]
    ey = [0, 0, 1, 0, -1, 1, 1, -1, -1]
    x = np.arange(18)[:, None] * np.ones((1, 18))
    y = np.ones((18, 1)) * np.arange(18)[None, :]
    lin = np.asfortranarray(np.stack([0.5 + 0.1 * q + 0.03 * x - 0.07 * y for q in range(9)]))
    for cfl in (0.5, 0.25, 0.125):
        s = lin.copy(order="F")
        L.orc_lax_wendroff_block(ctypes.byref(P), cfl, _P(s))
        for q in range(9):
            want = 0.5 + 0.1 * q + 0.03 * (x - cfl * ex[q]) - 0.07 * (y - cfl * ey[q])
            np.testing.assert_allclose(s[q, 1:17, 1:17], want[1:17, 1:17], rtol=1e-13, atol=1e-14)
        assert np.array_equal(s[0], lin[0])


def test_uniform_liquid_is_a_fixed_point_of_the_collision(oracle_mod, workloads):
    """C = 1, U = 0, P constant, h = w, g = P w: mu = 0, grad C = 0, Fb = 0,
    so neither population changes."""
    w = workloads["cfg1"]
    P = oracle_mod.make_params(w)
    L = oracle_mod.lib()
    wa = np.array([16, 4, 4, 4, 4, 1, 1, 1, 1]) / 36.0
    C = np.ones((20, 20), order="F")
    Pp = np.full((16, 16), 0.013, order="F")
    U = np.zeros((16, 16), order="F")
    V = np.zeros((16, 16), order="F")
    h = np.asfortranarray(np.broadcast_to(wa[:, None, None], (9, 18, 18)).copy())
    g = np.asfortranarray(np.broadcast_to(0.013 * wa[:, None, None], (9, 18, 18)).copy())
    h0, g0 = h.copy(), g.copy()
    for dx in (1, 2, 8):
        L.orc_collision_block(ctypes.byref(P), dx, _P(C), _P(Pp), _P(U), _P(V), _P(h), _P(g))
        np.testing.assert_allclose(h, h0, rtol=0, atol=1e-16)
        np.testing.assert_allclose(g, g0, rtol=0, atol=1e-16)


def test_h_collision_conserves_phase_field(oracle_mod, workloads):
    """sum_i heq_i = C (sum Gamma = 1, sum w e.n = 0), so the BGK h-collision
    keeps sum_i h_i when it equals C."""
    w = workloads["cfg1"]
    P = oracle_mod.make_params(w)
    L = oracle_mod.lib()
    o = oracle_mod.Oracle(w)
    o.setup()
    s = o.snapshot()
    b = s.blocks[s.leaves[5]]
    h = np.asfortranarray(b.h.copy())
    g = np.asfortranarray(b.g.copy())
    C = np.asfortranarray(b.C.copy())
    before = h[:, 1:17, 1:17].sum(axis=0)
    np.testing.assert_allclose(before, C[2:18, 2:18], rtol=1e-14)
    L.orc_collision_block(ctypes.byref(P), b.dx, _P(C), _P(np.asfortranarray(b.P)), _P(np.asfortranarray(b.Ux)),
                          _P(np.asfortranarray(b.Uy)), _P(h), _P(g))
    np.testing.assert_allclose(h[:, 1:17, 1:17].sum(axis=0), before, rtol=1e-13, atol=1e-15)


def test_topology_invariants_after_regrids(oracle_mod, workloads):
    """2:1 balance between leaf neighbours, and same-level neighbours see each
    other from the opposite side, after every move of the regrid walk."""
    w = workloads["cfg1"]
    o = oracle_mod.Oracle(w)
    o.setup()
    opp = {1: 2, 2: 1, 3: 4, 4: 3, 5: 8, 8: 5, 6: 7, 7: 6}
    for k in range(1, 15):
        o.set_bubble_c(w.L0 // 2 + 3.0 * k, w.L0 // 2 + 5.0 * k)
        o.phase("levels")
        o.regrid()
        s = o.snapshot()
        for n in s.leaves:
            b = s.blocks[n]
            for d in range(1, 9):
                m = b.nghbr[d - 1]
                if m < 0:
                    assert m == -d and d in (3, 4, 5, 6, 7, 8)
                    continue
                nb = s.blocks[m]
                assert b.level - nb.level in (0, 1), (n, d, m)
                if nb.level == b.level:
                    assert nb.nghbr[opp[d] - 1] == n, (n, d, m)
