"""Workload definitions and a run builder that sets up the objects in the
order of the reference driver (main/main.py:37-166).

Physics of every workload: x in [-L/2, L/2) with L = 2 pi / k0; velocity box
[-8.5, 8.5]^2 (thermal units, vt = 1); flux_coefficients = -1,
poisson_coefficient = -1 (exact Debye units); magnetic_field = omega_c/omega_p.
Initial condition: ring distribution of index j times
(1 + 1e-3 cos(k0 x) + sum_{m=2..4} 1e-6 cos(m k0 x + phi_m)), phi_m seeded.
"""
import dataclasses

import numpy as np

from . import _lib

from .basis import Basis3D
from .elliptic import Elliptic
from .fluxes import DGFlux
from .grid import Distribution, Grid3D, seeded_phases
from .timestep import Stepper


@dataclasses.dataclass
class Workload:
    name: str
    n: int
    nx: int
    nu: int
    nv: int
    steps: int
    dt: float                 # stable step used for parity / error runs
    dt_configured: float      # the step quoted with the workload (above the stability limit)
    k0: float = 0.1
    b_field: float = 0.1
    ring_j: int = 6
    vmax: float = 8.5
    seed: int = 0

    @property
    def dof(self):
        return self.nx * self.nu * self.nv * self.n ** 3


CONFIGS = {
    # 2.1M DOF reference-size case
    "ring16": Workload("ring16_n8_16x16x16", n=8, nx=16, nu=16, nv=16, steps=50, dt=0.008,
                       dt_configured=0.02),
    # 116.64M DOF single-GPU case
    "ring64": Workload("ring64_n9_64x50x50", n=9, nx=64, nu=50, nv=50, steps=200, dt=0.002,
                       dt_configured=0.01),
}


class Run:
    """Objects of one seeded run, built as main/main.py:45-154 builds them."""

    def __init__(self, n, nx, nu, nv, k0=0.1, b_field=0.1, ring_j=6, vmax=8.5, vt=1.0, acc=-1.0,
                 poisson=-1.0, amp1=1e-3, amps=(1e-6, 1e-6, 1e-6), seed=0, phases=None, dt=0.002,
                 device="cuda", storage="f64", ic="modulated", om=0.9, om_pc=1.0, delta_n=0.01):
        """``ic="modulated"``: the workloads' ring x density modulation;
        ``ic="eigen"``: the reference driver's ring + Bernstein eigenfunction
        (main/main.py:73-77) with ``om``, ``om_pc``, ``delta_n``."""
        self.basis = Basis3D(np.array([n, n, n]))
        L = 2.0 * np.pi / k0
        lows = np.array([-L / 2.0, -vmax * vt, -vmax * vt])
        highs = np.array([L / 2.0, vmax * vt, vmax * vt])
        self.grids = Grid3D(self.basis, lows, highs, np.array([nx, nu, nv]))
        self.phases = seeded_phases(seed, len(amps)) if phases is None else np.asarray(phases)
        self.dist = Distribution(vt, ring_j, [nx, nu, nv], [n, n, n], om=om, om_pc=om_pc, delta_n=delta_n,
                                 device=device)
        self.dist.initialize_quad_weights(self.grids)
        if ic == "eigen":
            self.dist.initialize_gpu(self.grids)
        elif ic == "modulated":
            self.dist.initialize_ring_modulated(self.grids, amp1, amps, self.phases)
        else:
            raise ValueError(f"unknown initial condition {ic!r}")
        self.elliptic = Elliptic(poisson)
        self.elliptic.set_magnetic_field(b_field)
        self.flux = DGFlux(np.array([nx + 2, nu + 2, nv + 2]), np.array([n, n, n]), acc)
        self.flux._geom(self.grids, self.elliptic)
        n0 = self.dist.moment_zero()
        self.elliptic.poisson2(_lib.density_deviation(self.grids.geometry(), n0, sign=-1),
                               self.grids.x, self.basis.b1)                   # main/main.py:131
        self.storage = storage
        if storage != "f64":
            self.dist.to_storage(self.grids, storage)
        self.stepper = Stepper(3, n, write_time=1e30, final_time=1e30, dt=dt)

    @classmethod
    def from_workload(cls, w, dt=None, device="cuda", **kw):
        return cls(w.n, w.nx, w.nu, w.nv, k0=w.k0, b_field=w.b_field, ring_j=w.ring_j, vmax=w.vmax,
                   seed=w.seed, dt=w.dt if dt is None else dt, device=device, **kw)

    @property
    def dof(self):
        g = self.grids
        return g.x.res * g.u.res * g.v.res * g.x.order ** 3

    def step(self, energy_out=None):
        self.stepper.nonlinear_ssp_rk(self.dist, self.basis, self.elliptic, self.grids, self.flux,
                                      energy_out=energy_out)

    def resident_begin(self, layout=None):
        self.stepper.resident_begin(self.dist, self.grids, self.flux, self.elliptic, layout=layout)

    def resident_end(self):
        self.stepper.resident_end(self.dist)

    def run(self, steps, graph_steps=0):
        self.stepper.main_loop(self.dist, self.basis, self.elliptic, self.grids, self.flux, steps=steps,
                               graph_steps=graph_steps)
        return self.stepper

    def capture(self, steps):
        """A HIP graph of ``steps`` steps of this run (see StepGraph)."""
        from .timestep import StepGraph
        return StepGraph(self.stepper, self.dist, self.basis, self.elliptic, self.grids, self.flux, steps)
