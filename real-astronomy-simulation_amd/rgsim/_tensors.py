"""Validation of the torch device buffers handed to the C ABI."""
from __future__ import annotations


def vec3(t, n: int, name: str):
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise TypeError(f"{name} must be a contiguous float64 CUDA tensor")
    if t.numel() != 3 * n:
        raise ValueError(f"{name} must hold ({n}, 3) values, has {tuple(t.shape)}")
    return t


def vec1(t, n: int, name: str):
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise TypeError(f"{name} must be a contiguous float64 CUDA tensor")
    if t.numel() != n:
        raise ValueError(f"{name} must hold ({n},) values, has {tuple(t.shape)}")
    return t


def vec3_f32(t, n: int, name: str):
    """(n,3) contiguous float32 CUDA tensor: the FP32 offsets of hi/lo positions."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor")
    if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
        raise TypeError(f"{name} must be a contiguous float32 CUDA tensor")
    if t.numel() != 3 * n:
        raise ValueError(f"{name} must hold ({n}, 3) values, has {tuple(t.shape)}")
    return t
