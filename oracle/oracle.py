"""ctypes binding of the C restatement in amroracle.c, plus a reader for the
binary snapshots the reference harness writes (oracle/ref_harness.f90).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
MAGIC = 1095586380


def build(force=False):
    """Compile amroracle.c (no FMA, no reassociation) into oracle/_build."""
    os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
    src = os.path.join(HERE, "amroracle.c")
    if (not force and os.path.exists(LIB_PATH)
            and os.path.getmtime(LIB_PATH) >= os.path.getmtime(src)):
        return LIB_PATH
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
                           "-shared", "-o", LIB_PATH, src, "-lm"])
    return LIB_PATH


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("nx", "ny", "L0", "nroots", "min_levels", "max_levels", "max_blocks")] + \
               [(n, ctypes.c_double) for n in
                ("M0", "W", "sigma", "beta", "kappa", "taul", "tauh", "rhol", "rhoh", "gy")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.POINTER(Params)]
        L.orc_destroy.argtypes = [vp]
        L.orc_setup_rising_bubble.argtypes = [vp, dp, dp, dp]
        for f in ("orc_determine_min_max_levels", "orc_fill_ghost_cells_c2", "orc_collide_all",
                  "orc_fill_ghost_cells_hg", "orc_stream_all", "orc_macroscopic_h_all",
                  "orc_fill_ghost_cells_c", "orc_macroscopic_g_all"):
            getattr(L, f).argtypes = [vp]
        L.orc_refine_derefine_mesh.argtypes = [vp, dp, dp]
        L.orc_set_bubble_c.argtypes = [vp, dp, dp, dp]
        L.orc_step.argtypes = [vp, ip]
        L.orc_run.argtypes = [vp, ip, ip]
        L.orc_set_step.argtypes = [vp, ip]
        L.orc_diagnostics.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.orc_load_snapshot.argtypes = [vp, ctypes.c_char_p]
        L.orc_save_snapshot.argtypes = [vp, ctypes.c_char_p]
        L.orc_nghbr_value.argtypes = [vp, ip, ip]
        L.orc_reference_physics.argtypes = [ctypes.POINTER(Params), dp, ip, dp, dp, dp, dp, dp, dp]
        P = ctypes.POINTER(ctypes.c_double)
        L.orc_collision_block.argtypes = [ctypes.POINTER(Params), ip, P, P, P, P, P, P]
        L.orc_macro_h_block.argtypes = [ctypes.POINTER(Params), P, P]
        L.orc_macro_g_block.argtypes = [ctypes.POINTER(Params), ip, P, P, P, P, P]
        L.orc_perfect_shift_block.argtypes = [ctypes.POINTER(Params), P]
        L.orc_lax_wendroff_block.argtypes = [ctypes.POINTER(Params), dp, P]
        L.orc_err_interface.restype = dp
        L.orc_err_interface.argtypes = [ctypes.POINTER(Params), ip, P]
        L.orc_biquadratic_from_fine.argtypes = [ip, ip, ip, ip, P, P]
        L.orc_biquadratic_from_coarse.argtypes = [ip, ip, ip, ip, P, P]
        L.orc_bilinear_from_fine.argtypes = [ip, ip, ip, ip, P, P]
        L.orc_bilinear_interp_extrap.argtypes = [ip, ip, P, P]
        L.orc_mrt_collision.argtypes = [dp, P, P]
        _lib = L
    return _lib


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous or a.flags.f_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(w):
    """Params from a workload description (amrlbm.workloads.Workload)."""
    p = Params()
    p.nx = p.ny = w.nx
    p.L0, p.nroots, p.min_levels, p.max_levels = w.L0, w.nroots, w.min_levels, w.max_levels
    p.max_blocks = w.max_blocks
    p.M0, p.W, p.sigma, p.beta, p.kappa = w.M0, w.W, w.sigma, w.beta, w.kappa
    p.taul, p.tauh, p.rhol, p.rhoh, p.gy = w.taul, w.tauh, w.rhol, w.rhoh, w.gy_run
    return p


class Oracle:
    """One mesh of the C restatement."""

    def __init__(self, workload):
        self.w = workload
        self.params = make_params(workload)
        self.h = lib().orc_create(ctypes.byref(self.params))

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setup(self):
        rc = lib().orc_setup_rising_bubble(self.h, self.w.radius, self.w.x0, self.w.y0)
        if rc != 0:
            raise RuntimeError("oracle set-up failed (MaxBlks exceeded)")

    def load(self, path):
        rc = lib().orc_load_snapshot(self.h, path.encode())
        if rc != 0:
            raise RuntimeError(f"orc_load_snapshot({path}) = {rc}")

    def save(self, path):
        if lib().orc_save_snapshot(self.h, path.encode()) != 0:
            raise RuntimeError("orc_save_snapshot failed")

    def phase(self, name):
        L = lib()
        {"levels": L.orc_determine_min_max_levels, "c2": L.orc_fill_ghost_cells_c2,
         "collide": L.orc_collide_all, "hg": L.orc_fill_ghost_cells_hg,
         "stream": L.orc_stream_all, "macro_h": L.orc_macroscopic_h_all,
         "c": L.orc_fill_ghost_cells_c, "macro_g": L.orc_macroscopic_g_all}[name](self.h)

    def set_bubble_c(self, x0, y0):
        lib().orc_set_bubble_c(self.h, self.w.radius, x0, y0)

    def regrid(self, refin_tol=0.002, deref_tol=0.001):
        lib().orc_refine_derefine_mesh(self.h, refin_tol, deref_tol)

    def step(self, regrid_every=1):
        lib().orc_step(self.h, regrid_every)

    def set_step(self, T):
        lib().orc_set_step(self.h, T)

    def run(self, nsteps, regrid_every=1):
        lib().orc_run(self.h, nsteps, regrid_every)

    def diagnostics(self):
        out = (ctypes.c_double * 9)()
        lib().orc_diagnostics(self.h, out)
        return np.array(out[:])

    def snapshot(self):
        """Current state as a Snapshot (through a temporary file)."""
        import tempfile
        with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
            path = f.name
        try:
            self.save(path)
            return read_snapshot(path)
        finally:
            os.unlink(path)


PHASES = ["start", "c2", "collide", "hg", "stream", "macro_h", "c", "macro_g", "regrid"]


class Block:
    __slots__ = ("id", "level", "dx", "I", "J", "parent", "isleaf", "hasleaf", "refin", "deref",
                 "child", "nghbr", "h", "g", "C", "P", "Ux", "Uy")


class Snapshot:
    """Mesh state in the reference layouts. Arrays use reference index order:
    h[q, x, y] with x,y in 0..nx+1; C[x+2, y+2] for x,y in -1..nx+2;
    P[i-1, j-1] for i,j in 1..nx."""

    def __init__(self, hdr, free, blocks):
        self.hdr = hdr
        self.nx, self.ny, self.L0 = int(hdr[2]), int(hdr[3]), int(hdr[4])
        self.min_levels, self.max_levels = int(hdr[5]), int(hdr[6])
        self.nblks, self.nleaf, self.T = int(hdr[7]), int(hdr[8]), int(hdr[9])
        self.levelmin, self.levelmax, self.kk = int(hdr[10]), int(hdr[11]), int(hdr[12])
        self.max_blocks = int(hdr[14])
        self.free = list(free)
        self.blocks = blocks          # dict id -> Block (all ids 1..nblks)

    @property
    def leaves(self):
        return [n for n in range(1, self.nblks + 1) if self.blocks[n].isleaf]

    def topology(self):
        """Integer description of every block: level, dx, I, J, parent, flags, child, nghbr."""
        rows = []
        for n in range(1, self.nblks + 1):
            b = self.blocks[n]
            rows.append([b.level, b.dx, b.I, b.J, b.parent, b.isleaf, b.hasleaf, b.refin,
                         b.deref] + list(b.child) + list(b.nghbr))
        return np.array(rows, dtype=np.int64)


def read_snapshot(path):
    with open(path, "rb") as f:
        data = f.read()
    hdr = np.frombuffer(data, dtype=np.int32, count=16, offset=0)
    if hdr[0] != MAGIC:
        raise ValueError(f"{path}: not a snapshot")
    nx, ny = int(hdr[2]), int(hdr[3])
    off = 64
    nfree = int(hdr[13])
    free = np.frombuffer(data, dtype=np.int32, count=nfree, offset=off)
    off += 4 * nfree
    nf = 9 * (nx + 2) * (ny + 2)
    nc = (nx + 4) * (ny + 4)
    npp = nx * ny
    blocks = {}
    for n in range(1, int(hdr[7]) + 1):
        iv = np.frombuffer(data, dtype=np.int32, count=22, offset=off)
        off += 88
        b = Block()
        b.id = n
        (b.level, b.dx, b.I, b.J, b.parent, b.isleaf, b.hasleaf, b.refin,
         b.deref) = (int(v) for v in iv[:9])
        b.child = tuple(int(v) for v in iv[9:13])
        b.nghbr = tuple(int(v) for v in iv[13:21])
        if iv[21]:
            arr = np.frombuffer(data, dtype=np.float64, count=2 * nf + nc + 3 * npp, offset=off)
            off += 8 * (2 * nf + nc + 3 * npp)
            b.h = arr[:nf].reshape((9, nx + 2, ny + 2), order="F")
            b.g = arr[nf:2 * nf].reshape((9, nx + 2, ny + 2), order="F")
            b.C = arr[2 * nf:2 * nf + nc].reshape((nx + 4, ny + 4), order="F")
            k = 2 * nf + nc
            b.P = arr[k:k + npp].reshape((nx, ny), order="F")
            b.Ux = arr[k + npp:k + 2 * npp].reshape((nx, ny), order="F")
            b.Uy = arr[k + 2 * npp:k + 3 * npp].reshape((nx, ny), order="F")
        else:
            b.h = b.g = b.C = b.P = b.Ux = b.Uy = None
        blocks[n] = b
    return Snapshot(hdr, free, blocks)


def compare(a, b, fields=("h", "g", "C", "P", "Ux", "Uy"), leaves_only=True):
    """Max abs and max relative (|x-y| / max(|y|, 1e-30) over |y|>1e-14 ... ) differences per field;
    raises on topology mismatch. Returns dict field -> (max_abs, n_bitdiff)."""
    ta, tb = a.topology(), b.topology()
    if ta.shape != tb.shape or not np.array_equal(ta, tb):
        raise AssertionError("topology differs")
    if a.free != b.free or a.kk != b.kk:
        raise AssertionError("free list differs")
    res = {}
    ids = a.leaves if leaves_only else range(1, a.nblks + 1)
    for f in fields:
        mx, nbit = 0.0, 0
        for n in ids:
            x, y = getattr(a.blocks[n], f), getattr(b.blocks[n], f)
            if x is None or y is None:
                if (x is None) != (y is None):
                    raise AssertionError(f"block {n}: {f} allocated in one snapshot only")
                continue
            d = np.abs(x - y)
            mx = max(mx, float(np.nanmax(d)) if d.size else 0.0)
            nbit += int(np.count_nonzero(x.view(np.int64) != y.view(np.int64)))
        res[f] = (mx, nbit)
    return res
