"""Grids of the reference (grids.py:40-219), as float64 torch tensors.

Fourier: ``dx = (b - a)/n``, ``x = arange(a, b, dx)``, ``kx = 2 pi (r)fftfreq(n, dx)``;
n must be an even integer greater than 2.  Chebyshev-Gauss-Lobatto points and
their differentiation matrix for the dense / diagonalized strategies.
Host-side setup, not a hot path.
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


def construct_x_cheb(n, a=-1.0, b=1.0, device=None):
    """n + 1 Chebyshev-Gauss-Lobatto points cos(pi (j - n)/n), j = 0..n (ascending from a to b,
    as grids.py:134-175 computes them)."""
    if not isinstance(n, (int, np.integer)):
        raise TypeError("n must be an integer.")
    if n < 2:
        raise ValueError("n must be ≥ 2.")
    x = np.cos(np.linspace(-np.pi, 0.0, n + 1))
    x = a + (b - a) * (x + 1) / 2.0
    t = torch.as_tensor(x, dtype=torch.float64)
    return t.to(device) if device else t


def construct_x_dx_cheb(n, a=-1.0, b=1.0, device=None):
    """(x, D) with D the Chebyshev differentiation matrix on those points (grids.py:178-219):
    D_ij = c_i / (c_j (x_i - x_j)) off the diagonal, c = (2, 1, ..., 1, 2) (-1)^i, and each
    diagonal entry minus its row's off-diagonal sum, so D 1 = 0."""
    x = construct_x_cheb(n, a, b).numpy()
    c = np.ones(n + 1)
    c[0] = c[-1] = 2.0
    c = c * (-1.0) ** np.arange(n + 1)
    dx = x[:, None] - x[None, :]
    d = np.outer(c, 1.0 / c) / (dx + np.eye(n + 1))
    d = d - np.diag(d.sum(axis=1))
    xt = torch.as_tensor(x, dtype=torch.float64)
    dt = torch.as_tensor(d, dtype=torch.float64)
    if device:
        return xt.to(device), dt.to(device)
    return xt, dt
