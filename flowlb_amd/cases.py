"""Synthetic culvert workloads (lattice units; y = 0 is the bed).

single_phase_512x128   : channel with bed/top walls, velocity inlet u = 0.05,
                         density-1 outlet, embankment x = 224..288 above a 32-cell
                         opening, tau = 0.56, rho0 = 1 + U(-1e-4, 1e-4), 10^4 steps.
free_surface_4096x1024 : closed box, water x <= 1799 to depth 600, embankment
                         x = 2000..2200 above a 120-cell barrel, gravity 2e-5,
                         tau = 0.51, 5*10^4 steps.
free_surface_8192x2048 : the same scaled x2 (water x < 3600 to depth 1200,
                         embankment 4000..4400, barrel 240).
free_surface_1024x256  : the same scaled /4 (comparison run of the large one).
adjoint_1024x256       : fully submerged culvert, opening 64 at x = 400..480,
                         inlet u = 0.04, tau = 0.6, body force 1e-6 along x, a
                         64x64 Brinkman window x = 336..399, y = 1..64.
"""
from dataclasses import dataclass, field, replace

import numpy as np

from . import core


@dataclass
class Case:
    name: str
    nx: int
    ny: int
    free_surface: bool
    bc: tuple
    emb: tuple
    opening: int
    water: tuple = (0, 0)
    u_inlet: float = 0.0
    rho_noise: float = 0.0
    seed: int = 20261202
    tau: float = 0.6
    force: tuple = (0.0, 0.0)
    steps: int = 1000
    extra: dict = field(default_factory=dict)

    @property
    def N(self):
        return self.nx * self.ny

    def build(self):
        """flags, rho0, u0 (host arrays)."""
        return core.build_culvert(self.nx, self.ny, self.free_surface, self.bc, self.emb,
                                  self.opening, self.water, self.u_inlet, self.rho_noise,
                                  self.seed)

    def options(self, storage="fp32", **kw):
        use_force = self.force != (0.0, 0.0)
        return core.default_options(tau=self.tau, fx=self.force[0], fy=self.force[1],
                                    use_force=int(use_force), problem=int(self.free_surface),
                                    storage=storage, **kw)

    def solver(self, storage="fp32", **kw):
        flags, rho0, u0 = self.build()
        s = core.Solver(flags, self.nx, self.ny, self.options(storage, **kw), rho0=rho0, u0=u0)
        if self.free_surface:
            s.set_fs_fields(*core.init_free_surface_fields(flags))
        return s


W = core.WALL
SINGLE_PHASE = Case("single_phase_512x128", 512, 128, False, (core.INLET, core.OUTLET, W, W),
                    (224, 288), 32, u_inlet=0.05, rho_noise=1e-4, tau=0.56, steps=10000)
FS_4096 = Case("free_surface_4096x1024", 4096, 1024, True, (W, W, W, W), (2000, 2200), 120,
               water=(1799, 600), tau=0.51, force=(0.0, -2e-5), steps=50000)
FS_8192 = Case("free_surface_8192x2048", 8192, 2048, True, (W, W, W, W), (4000, 4400), 240,
               water=(3599, 1200), tau=0.51, force=(0.0, -2e-5), steps=50000)
FS_1024 = Case("free_surface_1024x256", 1024, 256, True, (W, W, W, W), (500, 550), 30,
               water=(449, 150), tau=0.51, force=(0.0, -2e-5), steps=50000)
ADJOINT = Case("adjoint_1024x256", 1024, 256, False, (core.INLET, core.OUTLET, W, W), (400, 480),
               64, u_inlet=0.04, tau=0.6, force=(1e-6, 0.0), steps=2000,
               extra=dict(window=(336, 399, 1, 64), t_avg=(1500, 1999), ckpt=100))

# The 512 x 128 single-phase culvert scaled by 16 (fully wet apart from the
# embankment): the single-phase stream-collide kernel at HBM scale.
SINGLE_PHASE_8192 = Case("single_phase_8192x2048", 8192, 2048, False,
                         (core.INLET, core.OUTLET, W, W), (3584, 4608), 512, u_inlet=0.05,
                         rho_noise=1e-4, tau=0.56, steps=10000)

CASES = {c.name: c for c in (SINGLE_PHASE, FS_4096, FS_8192, FS_1024, ADJOINT, SINGLE_PHASE_8192)}


def scaled(case: Case, div: int, name=None) -> Case:
    """The same geometry with every length divided by div (for quick parity runs)."""
    def s(v):
        return max(1, int(round(v / div)))
    return replace(case, name=name or f"{case.name}/div{div}", nx=s(case.nx), ny=s(case.ny),
                   emb=(s(case.emb[0]), s(case.emb[1])), opening=s(case.opening),
                   water=(s(case.water[0]), s(case.water[1])) if case.water[0] else (0, 0))


def batch_instances(count=256, seed=20261202):
    """Random single-phase culverts: Nx ~ U{256..2048}, Ny ~ U{64..512}, opening
    round(U[0.08, 0.30] Ny), embankment of max(8, Nx/16) columns centred at
    U[Nx/3, 2Nx/3], inlet u ~ U[0.02, 0.08], tau ~ U[0.52, 0.8]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        nx = int(rng.integers(256, 2049))
        ny = int(rng.integers(64, 513))
        opening = int(round(rng.uniform(0.08, 0.30) * ny))
        thick = max(8, nx // 16)
        xc = int(rng.uniform(nx / 3, 2 * nx / 3))
        x0 = max(1, xc - thick // 2)
        u_in = float(rng.uniform(0.02, 0.08))
        tau = float(rng.uniform(0.52, 0.8))
        out.append(Case(f"instance{k}", nx, ny, False, (core.INLET, core.OUTLET, W, W),
                        (x0, x0 + thick - 1), opening, u_inlet=u_in, tau=tau, steps=5000))
    return out
