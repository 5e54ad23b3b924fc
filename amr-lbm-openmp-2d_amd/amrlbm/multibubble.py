"""The 32-bubble case: 2048 x 4096 finest-equivalent, levels 3..7, 32
non-overlapping bubbles with diameters uniform in [24, 96] finest cells at
seeded random positions, density ratio 1000, viscosity ratio 100, Eo = 10,
interface width W = 5 (the reference's; at W = 4 single bubbles of these
sizes run, but the 32-bubble field turns non-finite within about 1000 steps,
FP64 and FP32 alike, with the smallest bubbles left out too). The fluid parameters (surface tension, viscosities) are derived as in
Set_Rising_Bubble_Parameters (AMR_Rising_2D_OMP.f90:249-259) from the mean
diameter, 60 cells, with the Archimedes number scaled so that the relaxation
times equal the 1024 x 2048 case's (Ar = 40 (60/128)^1.5: tau_h = 0.243,
tau_l = 2.43; the reference's Ar = 40 at D = 60 gives tau_h = 0.078, which
does not stay stable at this density ratio). The bubbles are set up on the device
(amrlbm_setup_bubbles); per-bubble centroids come from the gas regions
(C < 1/2) labelled on the finest-equivalent raster, so that bubbles that
merge are followed as one region."""
import math
import random

import numpy as np

from .workloads import Workload

WORKLOAD = Workload("multi32", L0=2048, min_levels=3, radius=30.0, x0=1024.0, y0=2048.0, Bo=10, W=5.0,
                    Ar=40.0 * (60.0 / 128.0) ** 1.5, block_cap=32768)


def bubbles(n=32, seed=2024, dmin=24.0, dmax=96.0, gap=8.0, w=WORKLOAD):
    """n bubbles (R, x0, y0), diameters uniform in [dmin, dmax], at least `gap`
    cells apart and from the domain's edges (the set-up measures plain, not
    periodic, distances)."""
    rng = random.Random(seed)
    lx, ly = w.fine_shape
    out = []
    for _ in range(200000):
        if len(out) == n:
            return out
        r = 0.5 * rng.uniform(dmin, dmax)
        x = rng.uniform(r + gap, lx - r - gap)
        y = rng.uniform(r + gap, ly - r - gap)
        if all(math.hypot(x - x2, y - y2) >= r + r2 + gap for r2, x2, y2 in out):
            out.append((r, x, y))
    raise RuntimeError("cannot place the bubbles")


def setup(mesh, bub):
    r, x, y = zip(*bub)
    mesh.setup_bubbles(r, x, y)


def gas_regions(mesh):
    """Connected regions of C < 1/2 (4-connected on the finest-equivalent
    raster): [(area, xc, yc)] with area = sum((1 - C) dx^2) over the region's
    cells and the centroid weighted the same way, cell centres as in the
    diagnostics (AMR_Rising_2D_OMP.f90:469-478)."""
    from scipy import ndimage
    x, ids = mesh.cnn_features()
    rows, _ = mesh.topology()
    lx, ly = mesh.w.fine_shape
    nx = mesh.w.nx
    # the network input holds the signed distance d / 16 = (W / 4) ln(C / (1 - C)) / 16: C back from it
    d = x[:, 0, 1:nx + 1, 1:nx + 1].astype(np.float64) * 16.0   # [leaf, y, x]
    c = 1.0 / (1.0 + np.exp(-4.0 * d / mesh.w.W))
    gas = c < 0.5
    raster = np.zeros((ly, lx), dtype=bool)
    origin = []
    for k, n in enumerate(ids):
        _, dx, bi, bj = (int(v) for v in rows[n - 1][:4])
        x0, y0 = bi * dx * nx, bj * dx * nx
        origin.append((x0, y0, dx))
        if gas[k].any():
            raster[y0:y0 + dx * nx, x0:x0 + dx * nx] = np.repeat(np.repeat(gas[k], dx, axis=0), dx, axis=1)
    labels, nreg = ndimage.label(raster)
    area = np.zeros(nreg + 1)
    sx = np.zeros(nreg + 1)
    sy = np.zeros(nreg + 1)
    idx = np.arange(1, nx + 1, dtype=np.float64) - 0.5
    for k, (x0, y0, dx) in enumerate(origin):
        if not gas[k].any():
            continue
        lab = labels[y0:y0 + dx * nx:dx, x0:x0 + dx * nx:dx]
        wgt = np.where(gas[k], (1.0 - c[k]) * dx * dx, 0.0)
        xc = x0 + dx * idx[None, :]
        yc = y0 + dx * idx[:, None]
        np.add.at(area, lab, wgt)
        np.add.at(sx, lab, wgt * xc)
        np.add.at(sy, lab, wgt * yc)
    return [(area[i], sx[i] / area[i], sy[i] / area[i]) for i in range(1, nreg + 1) if area[i] > 0]


def match(a, b):
    """Greatest distance between a region of a and the nearest region of b
    (the centroid difference of each bubble between two runs)."""
    pb = np.array([(x, y) for _, x, y in b])
    return max(float(np.min(np.hypot(pb[:, 0] - x, pb[:, 1] - y))) for _, x, y in a)
