"""Ensembles of independent rising-bubble instances in one mesh
(amrlbm_setup_ensemble): every instance is a domain of two roots stacked in
y with its own size, level range, fluid and bubble, and every step launch
covers the leaves of all of them.

An ensemble is drawn from a seed with SplitMix64 (Steele, Lea and Flood,
"Fast splittable pseudorandom number generators", OOPSLA 2014; the
generator of java.util.SplittableRandom): draw k of the stream is
mix(seed + (k + 1) * 0x9E3779B97F4A7C15), and a uniform double in [0, 1) is
the top 53 bits of a draw times 2^-53. Per instance, in this order:

    domain      L0 from {256, 512, 1024} (L0 x 2 L0 finest cells)
    levels      2, 3 or 4 (min_levels = max_levels - levels + 1, max_levels = log2(L0 / 16))
    D           uniform in [16, 64] finest cells, centred at (L0/2, L0/2) (the reference's convention)
    Eo          uniform in [1, 100]
    rho ratio   log-uniform in [10, 1000] (rho_h = 1, rho_l = 1 / ratio)
    mu ratio    log-uniform in [1, 100]

W, M0 and g are the reference's (Set_Rising_Bubble_Parameters,
AMR_Rising_2D_OMP.f90:249-259: W = 5, g = 5e-6; M0 = 0.03). The configuration
names W = 4 and leaves Ar open: the reference's W = 5 is kept, as for every
other workload here, because the 32-bubble case turns non-finite at W = 4.
Ar is set per instance so that the lighter of the two relaxation times
is cfg2's (0.243): tau_l / tau_h = (rho ratio) / (mu ratio), so tau_h =
0.243 max(1, mu ratio / rho ratio), and Ar = sqrt(g rho_h (rho_h - rho_l)
D^3) / mu_h with mu_h = tau_h rho_h / 3. The reference's formulas then give
sigma, mu_h, tau_h, tau_l, beta and kappa of each instance
(workloads.Workload), in the reference's operation order.
"""
import math
from dataclasses import dataclass

from .workloads import Workload

MASK = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def splitmix64(seed, k):
    """Draw k (0-based) of the SplitMix64 stream of seed."""
    z = (seed + (k + 1) * GOLDEN) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


class Stream:
    def __init__(self, seed):
        self.seed, self.k = seed & MASK, 0

    def uniform(self):
        u = (splitmix64(self.seed, self.k) >> 11) * 2.0 ** -53
        self.k += 1
        return u

    def choice(self, options):
        return options[min(int(self.uniform() * len(options)), len(options) - 1)]


TAU_MIN = 0.243


def member(name, L0, levels, D, Eo, rho_ratio, mu_ratio, W=5.0, gy=5.0e-6):
    """One instance as a Workload (its own one-domain description)."""
    max_levels = int(round(math.log2(L0 // 16)))
    rhol = 1.0 / rho_ratio
    tauh = TAU_MIN * max(1.0, mu_ratio / rho_ratio)
    muh = tauh * 1.0 / 3.0
    Ar = math.sqrt(gy * 1.0 * (1.0 - rhol) * D ** 3) / muh
    return Workload(name, L0=L0, min_levels=max_levels - levels + 1, radius=D / 2.0, x0=L0 / 2.0, y0=L0 / 2.0,
                    Bo=Eo, Ar=Ar, gy=gy, W=W, rhol=rhol, visc_ratio=mu_ratio)


def draw(n, seed):
    """n instances of the ensemble of seed (the table above)."""
    s = Stream(seed)
    out = []
    for q in range(n):
        L0 = s.choice((256, 512, 1024))
        levels = s.choice((2, 3, 4))
        D = 16.0 + 48.0 * s.uniform()
        Eo = 1.0 + 99.0 * s.uniform()
        rho_ratio = 10.0 ** (1.0 + 2.0 * s.uniform())
        mu_ratio = 10.0 ** (2.0 * s.uniform())
        out.append(member(f"i{q}", L0, levels, D, Eo, rho_ratio, mu_ratio))
    return out


@dataclass(frozen=True)
class Ensemble:
    """A mesh of several instances: the mesh parameters cover them all
    (nroots = 2 n, the deepest level range, block capacity block_cap); the
    fluid constants of the mesh are the first instance's, and each kernel
    takes its block's instance's."""
    name: str
    members: tuple
    block_cap: int
    nx: int = 16

    @property
    def ny(self):
        return self.nx

    @property
    def instances(self):
        return len(self.members)

    @property
    def nroots(self):
        return 2 * len(self.members)

    @property
    def max_levels(self):
        return max(m.max_levels for m in self.members)

    @property
    def min_levels(self):
        return min(m.min_levels for m in self.members)

    @property
    def L0(self):
        return self.nx << self.max_levels

    @property
    def max_blocks(self):
        return self.block_cap

    @property
    def fine_cells(self):
        """Finest-equivalent cells of all instances together."""
        return sum(m.fine_cells for m in self.members)

    def __getattr__(self, name):   # M0, W, sigma, ...: the first instance's
        if name in ("M0", "W", "sigma", "beta", "kappa", "taul", "tauh", "rhol", "rhoh", "gy", "gy_run"):
            return getattr(self.members[0], name)
        raise AttributeError(name)


def base_blocks(members):
    """Blocks of the instances' base meshes (two roots refined uniformly to
    min_levels): 2 (4^(min+1) - 1) / 3 each."""
    return sum(2 * (4 ** (w.min_levels + 1) - 1) // 3 for w in members)


def ensemble(n=64, seed=2024, block_cap=None, name=None):
    """The batched ensemble: n instances drawn from seed; by default room for
    twice the base meshes' blocks (the interfaces' refinement and the free
    list), in multiples of 4096."""
    members = tuple(draw(n, seed))
    if block_cap is None:
        block_cap = -(-2 * base_blocks(members) // 4096) * 4096
    return Ensemble(name or f"ens{n}_s{seed}", members, block_cap)


def instance_of(level, J):
    """(instance, J in the instance's frame) of a block at (level, J)."""
    q = (J >> level) // 2
    return q, J - q * (2 << level)
