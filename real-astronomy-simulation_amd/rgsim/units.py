"""Galactic unit system (kpc, km/s, 1e10 Msun), as core/units.py:1-37."""

# Newton's constant in kpc (km/s)^2 / (1e10 Msun)  (core/units.py:21).
G = 43009.1725

KPC_PER_KM = 3.240779289e-17
GYR_PER_TIMEUNIT = 0.97781
MSUN_PER_MASSUNIT = 1.0e10


def time_to_gyr(t: float) -> float:
    return t * GYR_PER_TIMEUNIT
