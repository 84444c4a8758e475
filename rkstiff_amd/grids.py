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


def construct_r_kr_hankel(nr, rmax, device=None):
    """(r, kr, bessel_zeros, jN) of grids.py:222-264: with j_1 < ... < j_{nr+1} the zeros of
    J0, r_i = j_i rmax / j_{nr+1} and kr_i = j_i / rmax (i <= nr); jN = j_{nr+1}."""
    import scipy.special as sp
    if nr < 4:
        raise ValueError("nr must be ≥ 4.")
    if rmax <= 0:
        raise ValueError("rmax must be positive.")
    z = sp.jn_zeros(0, nr + 1)
    zeros, jN = z[:-1], float(z[-1])
    r = torch.as_tensor(zeros * rmax / jN, dtype=torch.float64)
    kr = torch.as_tensor(zeros / rmax, dtype=torch.float64)
    zt = torch.as_tensor(zeros, dtype=torch.float64)
    if device:
        r, kr, zt = r.to(device), kr.to(device), zt.to(device)
    return r, kr, zt, jN


def mirror_grid(r, u=None, axis=-1):
    """Mirror a radial grid (and optionally a function on it) onto [-rmax, rmax]
    (grids.py:267-321): rnew = [-flip(r), r]; u is mirrored along ``axis``
    (-1: 1-D, 0: rows, 1: columns).  Accepts NumPy arrays or torch tensors."""
    if isinstance(r, torch.Tensor):
        rnew = torch.cat([-torch.flip(r, dims=[0]), r])
        if u is None:
            return rnew, None
        if axis == -1:
            return rnew, torch.cat([torch.flip(u, dims=[-1]), u], dim=-1)
        if axis == 0:
            return rnew, torch.cat([torch.flip(u, dims=[0]), u], dim=0)
        if axis == 1:
            return rnew, torch.cat([torch.flip(u, dims=[1]), u], dim=1)
        raise ValueError("axis must be -1, 0, or 1.")
    rnew = np.hstack([-np.flipud(r), r])
    if u is None:
        return rnew, None
    if axis == -1:
        return rnew, np.hstack([np.flipud(u), u])
    if axis == 0:
        return rnew, np.vstack([np.flipud(u), u])
    if axis == 1:
        return rnew, np.hstack([np.fliplr(u), u])
    raise ValueError("axis must be -1, 0, or 1.")
