"""galic_amd: GALIC's iterative equilibrium hot path on MI355X (gfx950).

The compute path is the HIP library libgalic_hip.so behind the C ABI in
include/galic_hip.h; this package loads it (runtime.py), prepares model
inputs (model.py) and drives the optimisation loop the way the reference's
main.c does.  Importing the package does not touch the GPU.
"""
from . import _abi, model, rng  # noqa: F401

__all__ = ["_abi", "model", "rng"]
