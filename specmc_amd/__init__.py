"""specmc_amd: SpecMC's time-step hot path on AMD Instinct MI355X (gfx950).

    specmc_amd.ConvectionConfig   the scenario (same fields as specmc's)
    specmc_amd.api                config -> device model, march helpers
    specmc_amd.batch              many runs advancing together on one GPU
    specmc_amd.serve              continuous-batching request loop

The stepper's arithmetic runs in hand-written HIP kernels
(``specmc_amd/csrc``) behind the C ABI of ``include/specmc_hip.h``.
"""
from __future__ import annotations

import importlib

from .config import ConvectionConfig
from . import _numerics

__all__ = ["ConvectionConfig", "api", "_numerics"]
__version__ = "0.1.0"


def __getattr__(name):
    if name in ("api", "batch", "serve", "models", "engine"):
        return importlib.import_module("." + name, __name__)
    raise AttributeError("module %r has no attribute %r" % (__name__, name))
