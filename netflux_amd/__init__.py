"""MI355X-native explicit Euler step of the NetFlux finite-volume solver.

HIP kernels for gfx950 behind the C ABI of include/netflux.h
(libnetflux.so, built by netflux_amd/Makefile); this package is the Python
mirror of the reference's Config / EulerSolver / SolverData / Driver.
"""
from ._abi import LIB_PATH, lib  # noqa: F401
from .solver import BatchSolver, Config, Driver, EulerSolver, Mesh, SolverData  # noqa: F401
