"""Uniform Fourier grids (oracle).

Follows reference ``rkstiff/grids.py:40-85`` (rfft) and ``grids.py:88-131``
(fft): ``dx = (b-a)/n``, ``x = arange(a, b, dx)``, ``kx = 2*pi*(r)fftfreq(n, dx)``.
Pinned by: golden fixtures (the KS/NLS problems are built from these grids and
their trajectories match the reference bitwise) and by the spectral-derivative
identities in ``tests/test_oracle_fft.py``.
"""
import numpy as np


def _check_n(n):
    if not isinstance(n, (int, np.integer)):
        raise TypeError("n must be an integer.")
    if n <= 2 or (n % 2) != 0:
        raise ValueError("n must be an even integer greater than 2.")


def x_kx_rfft(n, a=0.0, b=2 * np.pi):
    """Grid ``x`` of ``n`` points on ``[a, b)`` and the ``n//2+1`` rfft wavenumbers."""
    _check_n(n)
    dx = (b - a) / n
    x = np.arange(a, b, dx)
    kx = 2 * np.pi * np.fft.rfftfreq(n, d=dx)
    return x, kx


def x_kx_fft(n, a=0.0, b=2 * np.pi):
    """Grid ``x`` of ``n`` points on ``[a, b)`` and the ``n`` fft wavenumbers."""
    _check_n(n)
    dx = (b - a) / n
    x = np.arange(a, b, dx)
    kx = 2 * np.pi * np.fft.fftfreq(n, d=dx)
    return x, kx
