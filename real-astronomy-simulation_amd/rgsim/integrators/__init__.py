"""Leapfrog KDK pieces over the HIP C ABI."""
