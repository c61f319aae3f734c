"""Gravity solvers (direct N^2, Barnes-Hut) over the HIP C ABI."""
