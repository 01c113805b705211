"""MI355X-native 1D2V nodal-DG Vlasov-Poisson SSP-RK3 path.

Drop-ins for the reference's object API (crewsdw/Vlasov1D2V):
``Basis1D/Basis3D`` (main/basis.py), ``Grid1D/Grid3D/Distribution``
(main/grid.py), ``DGFlux`` (main/fluxes.py), ``Elliptic`` (main/elliptic.py)
and ``Stepper`` (main/timestep.py).  Device work runs in the hand-written
gfx950 HIP library ``libvlasov1d2v.so`` through a ctypes C ABI
(include/vlasov1d2v.h); torch tensors are the device buffers.
"""
from .basis import Basis1D, Basis3D, lobatto  # noqa: F401
from .grid import Grid1D, Grid3D, Distribution, seeded_phases  # noqa: F401
from .fluxes import DGFlux  # noqa: F401
from .elliptic import Elliptic  # noqa: F401
from .timestep import Stepper, StepGraph, get_max_speeds  # noqa: F401
from .workload import Run, CONFIGS  # noqa: F401
from .ensemble import Ensemble, Member, draw_members  # noqa: F401
