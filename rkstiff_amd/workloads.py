"""Seeded synthetic ensembles (host NumPy arrays, spectral space).

Each builder returns one tuple per trajectory, ready for ``BatchSolver``
problems: the wavenumbers, the diagonal lin_op, the nonlinear term and the
initial spectrum.  The draws are made in the order stated, so a seed fixes
every trajectory.
"""
import numpy as np

from .nonlinear import AdvectiveNL, CubicNL


def ks_ensemble(count, seed, sizes=(256, 512, 1024, 2048, 4096), domain=(22.0, 200.0)):
    """Kuramoto-Sivashinsky u_t = -u_xx - u_xxxx - u u_x (README.md:89-97) on [0, D).

    Per trajectory, in this order: n ~ U{sizes}, D ~ U[domain], a_k ~ N(0, 0.1^2)
    and theta_k ~ U[0, 2 pi) for k = 1..8; u0 = sum_k a_k cos(2 pi k x / D + theta_k).
    Returns [(n, kx, L, NL, u0_hat)] with L = kx^2 (1 - kx^2), NL = -rfft(u u_x).
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(sizes))
        dom = float(rng.uniform(*domain))
        amp = rng.normal(0.0, 0.1, 8)
        th = rng.uniform(0.0, 2 * np.pi, 8)
        x = np.arange(n) * (dom / n)
        kx = 2 * np.pi * np.fft.rfftfreq(n, d=dom / n)
        k = np.arange(1, 9)
        u0 = np.sum(amp[:, None] * np.cos(2 * np.pi * k[:, None] * x[None, :] / dom + th[:, None]), axis=0)
        out.append((n, kx, kx**2 * (1 - kx**2), AdvectiveNL(kx, -1.0), np.fft.rfft(u0)))
    return out


def nls_ensemble(count, seed, sizes=(1024, 2048, 4096, 8192)):
    """Cubic NLS u_t = (i/2) u_xx + i |u|^2 u on [-64, 64) with 1-3 solitons.

    Per trajectory, in this order: n ~ U{sizes}, m ~ U{1, 2, 3}; then per
    soliton: eta ~ U[1, 2.5], v ~ U[-2, 2], x0 ~ U[-40, 40], phase ~ U[0, 2 pi).
    u0 = sum eta sech(eta (x - x0)) exp(i (v x + phase)).
    Returns [(n, kx, L, NL, u0_hat)] with L = -(i/2) kx^2, NL = i fft(|v|^2 v).
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(sizes))
        m = int(rng.integers(1, 4))
        x = np.arange(-64.0, 64.0, 128.0 / n)[:n]
        u0 = np.zeros(n, dtype=np.complex128)
        for _s in range(m):
            eta = rng.uniform(1.0, 2.5)
            v = rng.uniform(-2.0, 2.0)
            x0 = rng.uniform(-40.0, 40.0)
            ph = rng.uniform(0.0, 2 * np.pi)
            u0 += eta / np.cosh(eta * (x - x0)) * np.exp(1j * (v * x + ph))
        kx = 2 * np.pi * np.fft.fftfreq(n, d=128.0 / n)
        out.append((n, kx, -0.5j * kx**2, CubicNL(n, 1.0), np.fft.fft(u0)))
    return out
