"""MI355X (gfx950) implementation of the implicit elastodynamics step of the higher-order
embedded FEM reference: Stable Neo-Hookean assembly, colored CSR assembly, Schwarz-PCG and
Backward-Euler Newton with line search, as HIP kernels behind the C ABI of include/hoefem.h."""

import importlib

from . import scenes  # noqa: F401
from ._lib import HoefemError, HoefemLibraryMissing, LIB_PATH  # noqa: F401

__all__ = ["scenes", "HoefemError", "HoefemLibraryMissing", "LIB_PATH"]

_FEM_NAMES = ("StableNeoHookeanMaterial", "ElasticityModel", "SchwarzPreconditioner", "conjugate_gradient",
              "NewtonCgIntegrator")


def __getattr__(name):
    # the torch-backed interface is imported lazily so that scene generation works without torch
    if name == "fem" or name in _FEM_NAMES:
        fem = importlib.import_module(".fem", __name__)
        return fem if name == "fem" else getattr(fem, name)
    raise AttributeError(name)
