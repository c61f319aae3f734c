"""Engine factory for a level's solver name, as core/simulation.py:20-33
(``_make_engine``), with the tree solver reachable.

The reference's factory builds ``Engine(softening)`` for ``"direct"`` and
raises for any other self-gravity solver, although configs/levels.yaml
names ``tree`` and Engine already takes ``gravity_mode="bh"``.  Here
``"tree"`` (and ``"bh"``) build the Barnes-Hut engine.  ``"test_particle"``
(massless tracers in an analytic potential, core/solvers/test_particle.py)
is not part of this package and raises.
"""
from __future__ import annotations

from typing import Optional

from .engine import Engine

SOLVERS = ("direct", "tree", "bh")


def make_engine(solver: str, softening: float, params: Optional[dict] = None, *,
                precision: str = "fp64") -> Engine:
    """An Engine for ``solver``; ``params`` may carry ``theta`` for the tree
    (default 0.6, core/engine.py:36) and ``G``."""
    params = params or {}
    extra = {"precision": precision}
    if "G" in params:
        extra["G"] = float(params["G"])
    if solver == "direct":
        return Engine(softening=softening, **extra)
    if solver in ("tree", "bh"):
        return Engine(softening=softening, gravity_mode="bh",
                      theta=float(params.get("theta", 0.6)), **extra)
    if solver == "test_particle":
        raise NotImplementedError("the test-particle engine is not part of rgsim")
    raise ValueError(f"solver '{solver}' not implemented")
