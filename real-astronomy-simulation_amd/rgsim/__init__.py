"""rgsim -- MI355X-native gravity hot path for the real-astronomy-simulation
reference: direct N^2 and LBVH Barnes-Hut forces with a leapfrog KDK driver,
hand-written HIP kernels behind the C ABI in include/rgsim.h.

Module map to the reference (core/...):
    rgsim.engine.Engine                 core/engine.py:32
    rgsim.solvers.gravity               core/solvers/gravity.py
    rgsim.solvers.barnes_hut            core/solvers/barnes_hut.py
    rgsim.integrators.leapfrog          core/integrators/leapfrog.py
    rgsim.state.State                   core/state.py:25
    rgsim.units.G                       core/units.py:21
    rgsim.ic                            core/ic/plummer.py, core/ic/disk.py
"""
from .units import G  # noqa: F401
from .state import State  # noqa: F401

__all__ = ["G", "State"]
