"""CPU oracle for the 1D2V nodal-DG Vlasov-Poisson SSP-RK3 path.

TEST INFRASTRUCTURE ONLY.  This package is a plain NumPy FP64 restatement of
the reference's hot path (crewsdw/Vlasov1D2V, main/basis.py, main/grid.py,
main/fluxes.py, main/elliptic.py, main/timestep.py) with NumPy in place of
CuPy.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it, and there only as the checker or as the
timed CPU baseline.  The product package ``vlasov1d2v_amd`` never imports it.

Pinning status of each function is listed in the header of
``oracle/vlasov_oracle.py``.
"""
from .vlasov_oracle import *  # noqa: F401,F403
