"""Seeded synthetic initial conditions (restated generators, G explicit)."""
from .plummer import make_plummer, random_directions  # noqa: F401
from .disk import make_disk_galaxy, make_disk_bulge  # noqa: F401
