"""Model operators under the reference's names (models.py:32-247).

Fourier models return the built-in (device-fused) nonlinear terms; the
Chebyshev Allen-Cahn model returns a dense lin_op for the dense or
diagonalized strategies and an nl_func on device tensors.
"""
import numpy as np
import torch

from .nonlinear import AdvectiveNL, _as_f64


def _sech2_profile(x, ampl, x0, t):
    return 0.5 * ampl**2 / np.cosh(ampl * (x - x0 - ampl**2 * t) / 2) ** 2


def kdv_soliton(x, ampl=0.5, x0=0.0, t=0.0):
    """u = (a^2/2) sech^2(a (x - x0 - a^2 t) / 2) (models.py:32-73)."""
    xs = x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
    u = _sech2_profile(xs, ampl, x0, t)
    return torch.as_tensor(u, dtype=torch.float64) if isinstance(x, torch.Tensor) else u


def kdv_multi_soliton(x, ampl, x0, t=0.0):
    """Sum of sech^2 solitons (models.py:76-110)."""
    if len(x0) != len(ampl):
        raise ValueError("Lengths of ampl and x0 must match.")
    xs = x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
    u = sum(_sech2_profile(xs, a, c, t) for a, c in zip(ampl, x0))
    return torch.as_tensor(u, dtype=torch.float64) if isinstance(x, torch.Tensor) else u


def kdv_ops(kx):
    """L = i kx^3, N = -6 rfft(u u_x) (models.py:113-150)."""
    kx = _as_f64(kx)
    return torch.complex(torch.zeros_like(kx), kx**3), AdvectiveNL(kx, -6.0)


def burgers_ops(kx, mu):
    """L = -mu kx^2, N = -rfft(u u_x) (models.py:153-194)."""
    kx = _as_f64(kx)
    return -mu * kx**2, AdvectiveNL(kx, -1.0)


def allen_cahn_ops(x, d_cheb_matrix, epsilon=0.01):
    """Chebyshev Allen-Cahn (models.py:202-247): lin_op = epsilon D^2 + I on the interior
    points; nl_func(w) = x - (w + x)^3 on the interior (w = u - x carries the boundary values)."""
    x = _as_f64(x)
    d = d_cheb_matrix.detach().cpu().to(torch.float64) if isinstance(d_cheb_matrix, torch.Tensor) \
        else torch.as_tensor(np.asarray(d_cheb_matrix), dtype=torch.float64)
    lin_op = epsilon * (d @ d) + torch.eye(d.shape[0], dtype=torch.float64)
    lin_op = lin_op[1:-1, 1:-1].contiguous()
    xi = x[1:-1]
    cache = {}

    def nl_func(w):
        key = (w.device, w.dtype)
        if key not in cache:
            cache[key] = xi.to(device=w.device, dtype=w.dtype)
        xs = cache[key]
        v = w + xs
        return xs - v * v * v
    return lin_op, nl_func
