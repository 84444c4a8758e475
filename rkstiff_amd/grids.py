"""Uniform Fourier grids (grids.py:40-131 of the reference), as float64 torch tensors.

``dx = (b - a)/n``, ``x = arange(a, b, dx)``, ``kx = 2 pi (r)fftfreq(n, dx)``;
n must be an even integer greater than 2.  Host-side setup, not a hot path.
"""
import numpy as np
import torch


def _check(n):
    if not isinstance(n, (int, np.integer)):
        raise TypeError("n must be an integer.")
    if n <= 2 or (n % 2) != 0:
        raise ValueError("n must be an even integer greater than 2.")


def construct_x_kx_rfft(n, a=0.0, b=2 * np.pi, device=None):
    """(x[n], kx[n//2+1]) for real fields."""
    _check(n)
    dx = (b - a) / n
    x = torch.as_tensor(np.arange(a, b, dx), dtype=torch.float64)
    kx = torch.as_tensor(2 * np.pi * np.fft.rfftfreq(n, d=dx), dtype=torch.float64)
    return x.to(device) if device else x, kx.to(device) if device else kx


def construct_x_kx_fft(n, a=0.0, b=2 * np.pi, device=None):
    """(x[n], kx[n]) for complex fields."""
    _check(n)
    dx = (b - a) / n
    x = torch.as_tensor(np.arange(a, b, dx), dtype=torch.float64)
    kx = torch.as_tensor(2 * np.pi * np.fft.fftfreq(n, d=dx), dtype=torch.float64)
    return x.to(device) if device else x, kx.to(device) if device else kx
