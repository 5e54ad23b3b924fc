"""Rising-bubble workloads: mesh geometry and fluid parameters.

The fluid parameters are the reference's own set (Set_Rising_Bubble_Parameters,
AMR_Rising_2D_OMP.f90:241-259: Bo=40, Ar=40, g=5e-6, W=5, rho_h=1, rho_l=0.001,
viscosity ratio 100) with M0=0.03 (D2Q9_Mod.f90:10). The geometries are
16x16-cell blocks under two root blocks stacked in y (per instance), x periodic,
bounce-back walls in y. The derived quantities are computed in the reference's operation
order so that they agree bit for bit with the reference build.
"""
import math
from dataclasses import dataclass, field


@dataclass(frozen=True)
class Workload:
    name: str
    L0: int                 # root block extent in finest cells (Parameters.f90:15)
    min_levels: int         # Parameters.f90:17
    radius: float           # bubble radius in finest cells (AMR_Rising_2D_OMP.f90:154)
    x0: float
    y0: float
    nx: int = 16
    nroots: int = 2
    Bo: int = 40
    Ar: float = 40.0
    gy: float = 5.0e-6
    W: float = 5.0
    M0: float = 0.03
    rhol: float = 0.001
    rhoh: float = 1.0
    visc_ratio: float = 100.0
    refin_tol: float = 0.002
    deref_tol: float = 0.001
    instances: int = 1      # independent copies stacked in y (amrlbm_setup_replicas)
    gy_step: float = 0.0    # when set, the gravity of the time steps (the set-up's sigma, tau and
                            # pressure keep the ones gy gives): a faster-rising bubble whose
                            # interface crosses blocks within a few thousand steps
    block_cap: int = 0      # MaxBlks when set; else the reference's formula

    @property
    def ny(self):
        return self.nx

    @property
    def max_levels(self):      # Parameters.f90:16
        return int(round(math.log(self.L0 // self.nx) / math.log(2.0)))

    @property
    def max_blocks(self):      # Parameters.f90:19 (per root pair: 6 (L0/nx)^2)
        return self.block_cap or 6 * (self.L0 // self.nx) ** 2

    @property
    def fine_shape(self):      # finest-equivalent lattice (x, y)
        return (self.L0, self.nroots * self.L0)

    @property
    def fine_cells(self):
        return self.fine_shape[0] * self.fine_shape[1]

    @property
    def gy_run(self):          # the gravity the time steps use (D2Q9_Mod's Gy)
        return self.gy_step or self.gy

    @property
    def D(self):
        return 2.0 * self.radius

    # Set_Rising_Bubble_Parameters (AMR_Rising_2D_OMP.f90:249-259)
    @property
    def sigma(self):
        D = self.D
        return self.gy * (self.rhoh - self.rhol) * (D * D) / self.Bo

    @property
    def muh(self):
        D = self.D
        return math.sqrt(self.gy * self.rhoh * (self.rhoh - self.rhol) * (D * D * D)) / self.Ar

    @property
    def tauh(self):
        return 3.0 * (self.muh / self.rhoh)

    @property
    def taul(self):
        return 3.0 * ((self.muh / self.visc_ratio) / self.rhol)

    @property
    def beta(self):
        return 12.0 * self.sigma / self.W

    @property
    def kappa(self):
        return 1.5 * self.sigma * self.W


WORKLOADS = {
    # 64 x 128 finest-equivalent, levels 1..2: the small parity case
    "tiny": Workload("tiny", L0=64, min_levels=1, radius=10.0, x0=32.0, y0=32.0),
    # 128 x 256, levels 1..3, D=32 at (64,64): the CPU-runnable case
    "cfg1": Workload("cfg1", L0=128, min_levels=1, radius=16.0, x0=64.0, y0=64.0),
    # 1024 x 2048, levels 3..6 (dx = 8,4,2,1), D=128 at (512,512): the benchmark case
    "cfg2": Workload("cfg2", L0=1024, min_levels=3, radius=64.0, x0=512.0, y0=512.0),
    # 16 independent copies of cfg2 in one mesh (1024 x 32768 finest-equivalent, ~4,600 leaves, ~0.5 GB
    # of leaf state): the step kernels at a working set beyond the 256 MB Infinity Cache
    "cfg2x16": Workload("cfg2x16", L0=1024, min_levels=3, radius=64.0, x0=512.0, y0=512.0, nroots=32,
                        instances=16, block_cap=16384),
    # the tiny and 128 x 256 bubbles with 20x and 10x the gravity in the time steps: the mesh
    # changes within 2,400 steps (refinement, derefinement, reuse of freed ids); parity cases
    # across mesh changes
    "tiny_rise": Workload("tiny_rise", L0=64, min_levels=1, radius=10.0, x0=32.0, y0=32.0, gy_step=1.0e-4),
    "cfg1_rise": Workload("cfg1_rise", L0=128, min_levels=1, radius=16.0, x0=64.0, y0=64.0, gy_step=5.0e-5),
    # 2 copies of cfg1: the replicated-instance parity case
    "cfg1x2": Workload("cfg1x2", L0=128, min_levels=1, radius=16.0, x0=64.0, y0=64.0, nroots=4, instances=2),
}
