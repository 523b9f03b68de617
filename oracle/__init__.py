"""CPU FP64 restatement of SpecMC's time-step hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker the HIP path is measured against.  It restates, in
plain numpy/scipy and in the reference's own complex arithmetic, the algorithm
of the reference package ``specmc`` for the stepper behind
``api.march_adaptive``:

* ``oracle.spectral``  -- CGL nodes, Chebyshev bases, the x-FFT transforms and
  the Chebyshev-tau solver (rbc_galerkin.py:63-185, 317-339);
* ``oracle.krylov``    -- right-preconditioned restarted FGMRES
  (krylov_fgmres.py:42-179);
* ``oracle.laws``      -- viscosity laws, the eta window, the sources
  (viscosity.py:157-189, arrhenius.py:208-264, eta_bounds.py:199-480,
  heating.py:168-230, dissipation.py:234-487);
* ``oracle.models``    -- the isoviscous ``RBC`` step, the BDF2 implicit theta
  step, the depth-only and the coupled eta(T) Stokes solves (BFBT route with
  the legacy / direct rescue, the legacy lagged-LU route), CFL and diagnostics;
* ``oracle.budget``    -- the energy budget (heating.py:269-306,
  dissipation.py:365-571).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this package, and only as the checker or the timed CPU baseline.  The
product package ``specmc_amd`` never imports it.

Parity status of every function is listed in the repository's design notes;
each function here cites the reference file:line it follows.
"""
from . import spectral, krylov, laws, models, budget  # noqa: F401
from .models import build_oracle, march_adaptive, march  # noqa: F401

__all__ = ["spectral", "krylov", "laws", "models", "budget", "build_oracle",
           "march_adaptive", "march"]
