"""Python host binding of libamrlbm.so (the MI355X two-phase AMR-LBM step).

The product is the C-ABI library (include/amrlbm.h); this module is a thin
ctypes layer over it for tests, benchmarks and scripting. It mirrors the
reference's subroutine surface: Mesh.fill_ghost_cells_c2() is
Fill_Ghost_Cells_C2, Mesh.refine_derefine_mesh() is Refine_Derefine_Mesh, and
so on. There is no CPU fallback: if the library is missing or no GPU is
present, constructing a Mesh raises.
"""
import ctypes
import os

import numpy as np

from .workloads import WORKLOADS, Workload  # noqa: F401

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "lib", "libamrlbm.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "amrlbm.h")

NX = NY = 16
ERRORS = {1: "HIP", 2: "ARG", 3: "CAPACITY", 4: "TOPOLOGY", 5: "NONFINITE", 6: "STATE"}


class AmrLbmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"amrlbm error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("nx", "ny", "L0", "nroots", "min_levels", "max_levels", "max_blocks")] + \
               [(n, ctypes.c_double) for n in
                ("M0", "W", "sigma", "beta", "kappa", "taul", "tauh", "rhol", "rhoh", "gy")]


class Instance(ctypes.Structure):
    """amrlbm_instance: one instance of an ensemble mesh."""
    _fields_ = [(n, ctypes.c_int) for n in ("L0", "min_levels", "max_levels")] + \
               [(n, ctypes.c_double) for n in
                ("M0", "W", "sigma", "beta", "kappa", "taul", "tauh", "rhol", "rhoh", "gy", "R", "x0", "y0")]


def instance_from_workload(w):
    d = Instance()
    d.L0, d.min_levels, d.max_levels = w.L0, w.min_levels, w.max_levels
    d.M0, d.W, d.sigma, d.beta, d.kappa = w.M0, w.W, w.sigma, w.beta, w.kappa
    d.taul, d.tauh, d.rhol, d.rhoh, d.gy = w.taul, w.tauh, w.rhol, w.rhoh, w.gy_run
    d.R, d.x0, d.y0 = w.radius, w.x0, w.y0
    return d


def params_from_workload(w):
    p = Params()
    p.nx, p.ny = w.nx, w.ny
    p.L0, p.nroots, p.min_levels, p.max_levels = w.L0, w.nroots, w.min_levels, w.max_levels
    p.max_blocks = w.max_blocks
    p.M0, p.W, p.sigma, p.beta, p.kappa = w.M0, w.W, w.sigma, w.beta, w.kappa
    p.taul, p.tauh, p.rhol, p.rhoh, p.gy = w.taul, w.tauh, w.rhol, w.rhoh, w.gy_run
    return p


_lib = None
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)


def exported_functions():
    """Names of the functions include/amrlbm.h declares."""
    import re
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(amrlbm_\w+)\s*\(", src, re.M)))


def lib():
    """Load libamrlbm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built; run amr-lbm-openmp-2d_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        PP = ctypes.POINTER(Params)
        sig = {
            "amrlbm_create": [PP, ctypes.POINTER(vp)],
            "amrlbm_create_fp32": [PP, ctypes.POINTER(vp)],
            "amrlbm_destroy": [vp],
            "amrlbm_setup_rising_bubble": [vp, dp, dp, dp],
            "amrlbm_setup_replicas": [vp, ip, dp, dp, dp],
            "amrlbm_setup_bubbles": [vp, ip, _DP, _DP, _DP],
            "amrlbm_setup_ensemble": [vp, ip, ctypes.POINTER(Instance)],
            "amrlbm_output_instances": [vp, ip, _DP],
            "amrlbm_tecplot_cell_center": [vp, ctypes.c_char_p, dp],
            "amrlbm_get_counts": [vp, _IP],
            "amrlbm_set_step": [vp, ip],
            "amrlbm_get_topology": [vp, _IP, _IP],
            "amrlbm_set_topology": [vp, ip, _IP, ip, ip, _IP, ip],
            "amrlbm_put_block": [vp, ip, _DP, _DP, _DP, _DP, _DP, _DP],
            "amrlbm_get_block": [vp, ip, _DP, _DP, _DP, _DP, _DP, _DP],
            "amrlbm_determine_min_max_levels": [vp, _IP, _IP],
            "amrlbm_fill_ghost_cells_c2": [vp],
            "amrlbm_collide": [vp],
            "amrlbm_fill_ghost_cells_hg": [vp],
            "amrlbm_stream": [vp],
            "amrlbm_phase_field_lbm_h_g": [vp, _DP],
            "amrlbm_macroscopic_values_h": [vp],
            "amrlbm_fill_ghost_cells_c": [vp],
            "amrlbm_macroscopic_values_g": [vp],
            "amrlbm_refine_derefine_mesh": [vp, dp, dp],
            "amrlbm_step": [vp, ip, ip],
            "amrlbm_profile_steps": [vp, ip, ip, _DP],
            "amrlbm_output": [vp, _DP],
            "amrlbm_output_noalloc": [vp, _DP],
            "amrlbm_synchronize": [vp],
            "amrlbm_last_step_ms": [vp, _DP],
            "amrlbm_regrid_stats": [vp, _DP],
            "amrlbm_time_kernels": [vp, ip, _DP],
            "amrlbm_set_tuning": [vp, _IP],
            "amrlbm_get_tuning": [vp, _IP],
            "amrlbm_autotune": [vp, ip, ip, _DP],
            "amrlbm_set_mass_correction": [vp, ip],
            "amrlbm_get_mass_budget": [vp, ip, _DP],
            "amrlbm_save_snapshot": [vp, ctypes.c_char_p],
            "amrlbm_cnn_forward": [ctypes.POINTER(ctypes.c_float), ip, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
            "amrlbm_cnn_gradient": [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ip, ctypes.POINTER(ctypes.c_float), ctypes.c_float, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
            "amrlbm_cnn_train": [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ip, ctypes.POINTER(ctypes.c_float), ip, ip, ctypes.c_float, ctypes.c_float, ctypes.c_uint, ctypes.POINTER(ctypes.c_float)],
            "amrlbm_cnn_features": [vp, ctypes.POINTER(ctypes.c_float), _IP],
            "amrlbm_regrid_check": [vp, _IP],
            "amrlbm_set_regrid_network": [vp, ctypes.POINTER(ctypes.c_float), dp, dp],
            "amrlbm_step_learned": [vp, ip, ip, ctypes.POINTER(ctypes.c_longlong)],
            "amrlbm_refine_derefine_predicted": [vp, _IP, ip],
            "amrlbm_load_snapshot": [vp, ctypes.c_char_p],
            "amrlbm_collision_block": [PP, ip, _DP, _DP, _DP, _DP, _DP, _DP],
            "amrlbm_macro_g_block": [PP, ip, _DP, _DP, _DP, _DP, _DP],
            "amrlbm_stream_block": [PP, ip, _DP],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.amrlbm_last_error.restype = ctypes.c_char_p
        L.amrlbm_last_error.argtypes = []
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise AmrLbmError(rc, lib().amrlbm_last_error().decode())


def _d(a):
    return None if a is None else a.ctypes.data_as(_DP)


def _i(a):
    return None if a is None else a.ctypes.data_as(_IP)


def pop_array():
    """Empty array in the reference layout h(9, 0:nx+1, 0:ny+1)."""
    return np.zeros((9, NX + 2, NY + 2), order="F")


CNN_NPARAM = 3058


def _f(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def cnn_forward(x, weights):
    """Logits [n, 2] of the regrid predictor for inputs x [n, 3, 18, 18]."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    out = np.zeros((x.shape[0], 2), dtype=np.float32)
    _check(lib().amrlbm_cnn_forward(_f(x), x.shape[0], _f(w), _f(out)))
    return out


def cnn_gradient(x, y, weights, pos_weight=1.0):
    """(gradient [CNN_NPARAM], loss) of the weighted logistic loss."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    g = np.zeros(CNN_NPARAM, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    _check(lib().amrlbm_cnn_gradient(_f(x), _f(y), x.shape[0], _f(w), float(pos_weight), _f(g), _f(loss)))
    return g, float(loss[0])


def cnn_train(x, y, weights, epochs, batch=256, lr=1e-3, pos_weight=1.0, seed=1):
    """Train on the GPU; returns (weights, per-epoch mean loss)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    w = np.array(weights, dtype=np.float32)
    hist = np.zeros(max(epochs, 1), dtype=np.float32)
    _check(lib().amrlbm_cnn_train(_f(x), _f(y), x.shape[0], _f(w), int(epochs), int(batch), float(lr),
                                  float(pos_weight), int(seed), _f(hist)))
    return w, hist[:epochs]


def cnn_init(seed=0):
    """He-uniform initial weights, zero biases."""
    rng = np.random.default_rng(seed)
    w = np.zeros(CNN_NPARAM, dtype=np.float32)
    o = 0
    for cout, fan in ((16, 27), (16, 144), (2, 144)):
        lim = np.sqrt(6.0 / fan)
        w[o:o + cout * fan] = rng.uniform(-lim, lim, cout * fan)
        o += cout * fan + cout
    return w


class Mesh:
    """One AMR mesh resident on the GPU. precision=64 is the reference's FP64
    arithmetic; precision=32 the FP32 production form (amrlbm_create_fp32),
    with the same host-side interface (FP64 arrays in, FP64 arrays out)."""

    def __init__(self, workload, precision=64):
        if precision not in (32, 64):
            raise ValueError("precision is 32 or 64")
        self.w = workload
        self.precision = precision
        self.params = params_from_workload(workload)
        h = ctypes.c_void_p()
        create = lib().amrlbm_create_fp32 if precision == 32 else lib().amrlbm_create
        _check(create(ctypes.byref(self.params), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().amrlbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- set-up and state exchange -------------------------------------
    def setup_rising_bubble(self, radius=None, x0=None, y0=None):
        """The reference's set-up; a workload of several instances gets one
        bubble per instance (amrlbm_setup_replicas)."""
        w = self.w
        if hasattr(w, "members"):   # an ensemble: each instance its own bubble
            self.setup_ensemble(w.members)
            return
        R = w.radius if radius is None else radius
        x = w.x0 if x0 is None else x0
        y = w.y0 if y0 is None else y0
        if w.instances > 1:
            _check(lib().amrlbm_setup_replicas(self.h, w.instances, R, x, y))
        else:
            _check(lib().amrlbm_setup_rising_bubble(self.h, R, x, y))

    def write_tecplot(self, path, D=None):
        """The reference's Tecplot file of the leaves (TecPlot_CellCenter);
        D defaults to the workload's bubble diameter."""
        _check(lib().amrlbm_tecplot_cell_center(self.h, os.fsencode(path), float(D if D is not None else self.w.D)))

    def setup_ensemble(self, members):
        """Instances of uneven size, level range and fluid in one mesh
        (amrlbm_setup_ensemble); members are one-domain Workloads."""
        arr = (Instance * len(members))(*[instance_from_workload(w) for w in members])
        _check(lib().amrlbm_setup_ensemble(self.h, len(members), arr))

    def output_instances(self, allow_nonfinite=False):
        """[n, 9] per-instance diagnostics of an ensemble mesh, each in its
        instance's frame (amrlbm_output_instances)."""
        n = self.w.instances
        out = np.zeros((n, 9))
        rc = lib().amrlbm_output_instances(self.h, n, _d(out))
        if rc != 0 and not (allow_nonfinite and rc == 5):
            _check(rc)
        return out

    def setup_bubbles(self, R, x0, y0):
        """Device-resident set-up of several bubbles (amrlbm_setup_bubbles)."""
        R, x0, y0 = (np.ascontiguousarray(v, dtype=np.float64) for v in (R, x0, y0))
        _check(lib().amrlbm_setup_bubbles(self.h, int(R.size), _d(R), _d(x0), _d(y0)))

    def counts(self):
        c = np.zeros(7, dtype=np.int32)
        _check(lib().amrlbm_get_counts(self.h, _i(c)))
        keys = ("nblks", "nleaf", "levelmin", "levelmax", "kk", "nfree", "T")
        return dict(zip(keys, (int(v) for v in c)))

    def set_step(self, T):
        """The step counter T (amrlbm_set_step): the phase entry points leave it
        to the caller, as the reference's subroutines leave it to Main_Loop."""
        _check(lib().amrlbm_set_step(self.h, int(T)))

    def topology(self):
        """(blocks[nblks,21], free list) in the reference's terms."""
        c = self.counts()
        t = np.zeros((max(c["nblks"], 1), 21), dtype=np.int32)
        f = np.zeros(max(c["nfree"], 1), dtype=np.int32)
        _check(lib().amrlbm_get_topology(self.h, _i(t), _i(f)))
        return t[:c["nblks"]], list(f[:c["nfree"]])

    def set_topology(self, blocks, kk, free, T=0):
        blocks = np.ascontiguousarray(blocks, dtype=np.int32)
        free = np.ascontiguousarray(np.asarray(free, dtype=np.int32).reshape(-1))
        _check(lib().amrlbm_set_topology(self.h, blocks.shape[0], _i(blocks), int(kk), len(free),
                                         _i(free) if len(free) else None, int(T)))

    def put_block(self, bid, h=None, g=None, C=None, P=None, Ux=None, Uy=None):
        conv = [None if a is None else np.asfortranarray(a, dtype=np.float64) for a in (h, g, C, P, Ux, Uy)]
        _check(lib().amrlbm_put_block(self.h, int(bid), *[_d(a) for a in conv]))

    def get_block(self, bid):
        h, g = pop_array(), pop_array()
        C = np.zeros((NX + 4, NY + 4), order="F")
        P, Ux, Uy = (np.zeros((NX, NY), order="F") for _ in range(3))
        _check(lib().amrlbm_get_block(self.h, int(bid), _d(h), _d(g), _d(C), _d(P), _d(Ux), _d(Uy)))
        return {"h": h, "g": g, "C": C, "P": P, "Ux": Ux, "Uy": Uy}

    def leaves(self):
        t, _ = self.topology()
        return [n + 1 for n in range(t.shape[0]) if t[n, 5]]

    # ---- the reference's per-step subroutines ----------------------------
    def determine_min_max_levels(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(lib().amrlbm_determine_min_max_levels(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def fill_ghost_cells_c2(self):
        _check(lib().amrlbm_fill_ghost_cells_c2(self.h))

    def collide(self):
        _check(lib().amrlbm_collide(self.h))

    def fill_ghost_cells_hg(self):
        _check(lib().amrlbm_fill_ghost_cells_hg(self.h))

    def stream(self):
        _check(lib().amrlbm_stream(self.h))

    def phase_field_lbm_h_g(self):
        ms = (ctypes.c_double * 3)()
        _check(lib().amrlbm_phase_field_lbm_h_g(self.h, ms))
        return list(ms)

    def macroscopic_values_h(self):
        _check(lib().amrlbm_macroscopic_values_h(self.h))

    def fill_ghost_cells_c(self):
        _check(lib().amrlbm_fill_ghost_cells_c(self.h))

    def macroscopic_values_g(self):
        _check(lib().amrlbm_macroscopic_values_g(self.h))

    def refine_derefine_mesh(self, refin_tol=0.002, deref_tol=0.001):
        _check(lib().amrlbm_refine_derefine_mesh(self.h, refin_tol, deref_tol))

    def step(self, nsteps=1, regrid_every=1):
        _check(lib().amrlbm_step(self.h, int(nsteps), int(regrid_every)))

    def profile_steps(self, nsteps, regrid_every=1):
        """Per-kernel-class device times (HIP events) over nsteps steps."""
        st = (ctypes.c_double * 12)()
        _check(lib().amrlbm_profile_steps(self.h, int(nsteps), int(regrid_every), st))
        names = ("c2_collide", "hg_stream", "c_macro_g")
        out = {n: {"ms": st[3 * k], "launches": int(st[3 * k + 1]), "cells": int(st[3 * k + 2])}
               for k, n in enumerate(names)}
        out["regrid"] = {"ms": st[9], "count": int(st[10])}
        return out

    def last_step_ms(self):
        v = ctypes.c_double()
        _check(lib().amrlbm_last_step_ms(self.h, ctypes.byref(v)))
        return v.value

    def save_snapshot(self, path):
        """Restart file in the reference dump harness's layout."""
        _check(lib().amrlbm_save_snapshot(self.h, os.fsencode(path)))

    def load_snapshot(self, path):
        """Replace the mesh by a snapshot file (the library's or the reference harness's)."""
        _check(lib().amrlbm_load_snapshot(self.h, os.fsencode(path)))

    def regrid_check(self):
        """(would change the mesh, would refine a block) at the current state."""
        out = np.zeros(2, dtype=np.int32)
        _check(lib().amrlbm_regrid_check(self.h, _i(out)))
        return bool(out[0]), bool(out[1])

    def refine_derefine_predicted(self, ids):
        """Refine_Derefine_Mesh with the given blocks flagged for refinement too."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        _check(lib().amrlbm_refine_derefine_predicted(self.h, _i(ids), int(ids.size)))

    def set_regrid_network(self, weights, threshold=0.0, gate=0.25):
        """The learned predictor amrlbm_step_learned uses (None: none)."""
        if weights is None:
            _check(lib().amrlbm_set_regrid_network(self.h, None, 0.0, 0.0))
            return
        w = np.ascontiguousarray(weights, dtype=np.float32)
        assert w.size == CNN_NPARAM
        _check(lib().amrlbm_set_regrid_network(self.h, _f(w), float(threshold), float(gate)))

    def step_learned(self, nsteps, full_every=20):
        """nsteps steps under the learned regrid's cadence (amrlbm_step_learned):
        {full_regrids, predicted_blocks, misses, judged}."""
        st = (ctypes.c_longlong * 4)()
        _check(lib().amrlbm_step_learned(self.h, int(nsteps), int(full_every), st))
        return {"full_regrids": st[0], "predicted_blocks": st[1], "misses": st[2], "judged": st[3]}

    def cnn_features(self):
        """Network input of every leaf: (x [nLeaf, 3, 18, 18] float32, block ids)."""
        n = self.counts()["nleaf"]
        x = np.zeros((n, 3, 18, 18), dtype=np.float32)
        ids = np.zeros(n, dtype=np.int32)
        _check(lib().amrlbm_cnn_features(self.h, _f(x), _i(ids)))
        return x, ids

    def time_kernels(self, reps=200):
        """Average ms per launch of the three step kernels (back-to-back launches,
        HIP events on the mesh's stream); the state is left unchanged."""
        out = (ctypes.c_double * 3)()
        _check(lib().amrlbm_time_kernels(self.h, int(reps), out))
        return {"c2_collide": out[0], "hg_stream": out[1], "c_macro_g": out[2]}

    def regrid_stats(self):
        """Device regrid decisions taken by step(), how many changed the mesh,
        how many segments launched ahead of a change were taken back, and the
        host time spent on regrids, tables, ghost programs and graphs."""
        out = (ctypes.c_double * 9)()
        _check(lib().amrlbm_regrid_stats(self.h, out))
        return {"checks": int(out[0]), "changes": int(out[1]), "segments_taken_back": int(out[2]),
                "host_ms": {"regrid": round(out[3], 3), "tables": round(out[4], 3),
                            "programs": round(out[5], 3), "graphs": round(out[6], 3)},
                "graphs_built": int(out[7]), "device_flag_changes": int(out[8])}

    def set_tuning(self, order, depth, execution=0):
        """Launch variant (amrlbm_set_tuning): leaf order 0-3, hg operand depth
        0-2, execution 0 (step graph) or 1 (persistent segment kernel)."""
        _check(lib().amrlbm_set_tuning(self.h, _i(np.array([order, depth, execution], dtype=np.int32))))

    def get_tuning(self):
        k = np.zeros(3, dtype=np.int32)
        _check(lib().amrlbm_get_tuning(self.h, _i(k)))
        return int(k[0]), int(k[1]), int(k[2])

    def autotune(self, reps=10, regrid_every=10):
        """Time the 24 launch variants on the current state and keep the
        fastest (amrlbm_autotune): [(order, depth, execution, us per step)],
        us -1 where the segment kernel does not fit."""
        t = np.zeros(96)
        _check(lib().amrlbm_autotune(self.h, int(reps), int(regrid_every), _d(t)))
        return [(int(t[4 * v]), int(t[4 * v + 1]), int(t[4 * v + 2]), float(t[4 * v + 3])) for v in range(24)]

    def set_mass_correction(self, on=True):
        """Hold each instance's FP64 phase-field mass at its budget (its value
        now plus each later mesh change's restrict/prolong change), corrected
        at every segment end (amrlbm_set_mass_correction)."""
        _check(lib().amrlbm_set_mass_correction(self.h, 1 if on else 0))

    def mass_budget(self):
        """[instances, 4]: target, mass before the last correction, interface
        weight, lambda (amrlbm_get_mass_budget)."""
        n = self.w.instances if hasattr(self.w, "members") else 1
        out = np.zeros((n, 4))
        _check(lib().amrlbm_get_mass_budget(self.h, n, _d(out)))
        return out

    def output(self, allow_nonfinite=False):
        """rho_min, rho_max, max|Uy|, mass, bubble area, centroid x, y, rise velocity, non-finite."""
        out = (ctypes.c_double * 9)()
        rc = lib().amrlbm_output(self.h, out)
        if rc != 0 and not (allow_nonfinite and rc == 5):
            _check(rc)
        return np.array(out[:])

    def output_noalloc(self, allow_nonfinite=False):
        """output() through amrlbm_output_noalloc (partials in the library's
        static device scratch; no allocation)."""
        out = (ctypes.c_double * 9)()
        rc = lib().amrlbm_output_noalloc(self.h, out)
        if rc != 0 and not (allow_nonfinite and rc == 5):
            _check(rc)
        return np.array(out[:])

    def synchronize(self):
        _check(lib().amrlbm_synchronize(self.h))

    def phase(self, name):
        {"c2": self.fill_ghost_cells_c2, "collide": self.collide, "hg": self.fill_ghost_cells_hg,
         "stream": self.stream, "macro_h": self.macroscopic_values_h, "c": self.fill_ghost_cells_c,
         "macro_g": self.macroscopic_values_g, "levels": self.determine_min_max_levels,
         "regrid": self.refine_derefine_mesh}[name]()


# ---- per-block kernels on host arrays (reference layouts) ---------------
def collision_block(workload, dx, C, P, Ux, Uy, h, g):
    """Collision_h_g_AMR on one block; h, g are updated in place (Fortran order)."""
    p = params_from_workload(workload)
    _check(lib().amrlbm_collision_block(ctypes.byref(p), int(dx), _d(C), _d(P), _d(Ux), _d(Uy), _d(h), _d(g)))


def macro_g_block(workload, dx, g, C, P, Ux, Uy):
    p = params_from_workload(workload)
    _check(lib().amrlbm_macro_g_block(ctypes.byref(p), int(dx), _d(g), _d(C), _d(P), _d(Ux), _d(Uy)))


def stream_block(workload, dx, f):
    p = params_from_workload(workload)
    _check(lib().amrlbm_stream_block(ctypes.byref(p), int(dx), _d(f)))
