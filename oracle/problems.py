"""Nonlinear terms and benchmark problems (oracle, NumPy FP64).

Nonlinear terms follow the reference's user ``nl_func`` patterns:

* ``rfft_advect_nl(kx, coef)``: ``coef * rfft(irfft(u) * irfft(1j*kx*u))``.
  KS and Burgers use coef=-1 (README.md:94-97, models.py:189-192), KdV coef=-6
  (models.py:140-143).
* ``nls_nl(gamma)``: ``1j*gamma*fft(|v|^2 v)``, ``v = ifft(u)``, with
  ``|v|^2 = v.real**2 + v.imag**2`` (demos/nls.ipynb cell 29, 1-D form).
* ``ac2d_nl(ny, nx)``: ``rfft2(w - w**3)``, ``w = irfft2(u)`` on a flattened
  ``ny x (nx//2+1)`` spectrum (the 2-D Fourier Allen-Cahn of the 512x512
  batch workload; the reference ships only a Chebyshev Allen-Cahn).

Problem builders restate the workloads: KS on [0, 32 pi) with n=1024
(README.md:80-107), two-soliton NLS on [-64, 64) with n=8192, the seeded KS
ensemble, the 2-D Allen-Cahn batch, the KdV soliton and Burgers fixtures
(tests/testing_util.py:28-56).
Pinned by: golden trajectories (KS/NLS/KdV/Burgers), analytic KdV and NLS
soliton solutions, and DFT brute force in ``tests/test_oracle_fft.py``.
"""
import numpy as np

from .grids import x_kx_fft, x_kx_rfft


# ----------------------------------------------------------------------------
# nonlinear terms
# ----------------------------------------------------------------------------
def rfft_advect_nl(kx, coef=-1.0):
    """coef * rfft(u * u_x) evaluated spectrally (KS / Burgers / KdV form)."""
    def nl(uf):
        u = np.fft.irfft(uf)
        ux = np.fft.irfft(1j * kx * uf)
        if coef == -1.0:
            return -np.fft.rfft(u * ux)
        return coef * np.fft.rfft(u * ux)
    return nl


def nls_nl(gamma=1.0):
    """1j*gamma*fft(|v|^2 v) with v = ifft(u) (cubic NLS, focusing for gamma>0)."""
    def nl(uf):
        v = np.fft.ifft(uf)
        v2 = v.real**2 + v.imag**2
        return 1j * gamma * np.fft.fft(v2 * v)
    return nl


def ac2d_nl(ny, nx):
    """rfft2(w - w**3), w = irfft2 of the flattened ny x (nx//2+1) spectrum."""
    m = nx // 2 + 1

    def nl(uf):
        w = np.fft.irfft2(uf.reshape(ny, m), s=(ny, nx))
        return np.fft.rfft2(w - w**3).reshape(-1)
    return nl


# ----------------------------------------------------------------------------
# problem builders: each returns dict(L, nl, u0, plus grid metadata)
# ----------------------------------------------------------------------------
def ks_readme(n=1024, a=0.0, b=32.0 * np.pi):
    """Kuramoto-Sivashinsky quickstart: L = kx^2 (1 - kx^2), u0 = cos(x/16)(1+sin(x/16))."""
    x, kx = x_kx_rfft(n, a, b)
    L = kx**2 * (1 - kx**2)
    u0 = np.cos(x / 16) * (1.0 + np.sin(x / 16))
    return dict(kind="rfft", n=n, x=x, kx=kx, L=L, nl=rfft_advect_nl(kx, -1.0),
                u0=np.fft.rfft(u0), nl_coef=-1.0)


def nls_two_soliton(n=8192, a=-64.0, b=64.0):
    """Two colliding NLS solitons: u_t = (i/2) u_xx + i |u|^2 u, L = -(i/2) kx^2."""
    x, kx = x_kx_fft(n, a, b)
    L = -0.5j * kx**2
    u0 = (2.0 / np.cosh(2.0 * (x + 10.0)) * np.exp(1j * x)
          + 1.5 / np.cosh(1.5 * (x - 10.0)) * np.exp(-0.5j * x))
    return dict(kind="fft", n=n, x=x, kx=kx, L=L, nl=nls_nl(1.0), u0=np.fft.fft(u0),
                nl_coef=1.0)


def nls_soliton_exact(x, t, eta=1.0, v=0.0, x0=0.0):
    """Single NLS soliton eta sech(eta (x-x0-v t)) exp(i (v x + (eta^2 - v^2) t / 2))."""
    return (eta / np.cosh(eta * (x - x0 - v * t))
            * np.exp(1j * (v * x + 0.5 * (eta**2 - v**2) * t)))


def kdv_soliton(x, ampl=0.5, x0=0.0, t=0.0):
    """KdV single soliton, models.py:32-73."""
    return 0.5 * ampl**2 / (np.cosh(ampl * (x - x0 - ampl**2 * t) / 2) ** 2)


def kdv_fixture():
    """tests/testing_util.py:42-56: n=256 on [-30,30), ampl 1, x0=-5, h=0.025, 200 steps."""
    x, kx = x_kx_rfft(256, -30, 30)
    h, steps = 0.025, 200
    tf = h * steps
    u0 = np.fft.rfft(kdv_soliton(x, ampl=1.0, x0=-5.0, t=0))
    exact = np.fft.rfft(kdv_soliton(x, ampl=1.0, x0=-5.0, t=tf))
    L = 1j * kx**3
    return dict(kind="rfft", n=256, x=x, kx=kx, L=L, nl=rfft_advect_nl(kx, -6.0),
                u0=u0, exact=exact, h=h, steps=steps, tf=tf, nl_coef=-6.0)


def burgers_fixture(n=1024, mu=0.0005):
    """tests/testing_util.py:28-39: n=1024 on [-pi,pi), u0 = exp(-10 sin^2(x/2))."""
    x, kx = x_kx_rfft(n, -np.pi, np.pi)
    L = -mu * kx**2
    u0 = np.exp(-10 * np.sin(x / 2) ** 2)
    return dict(kind="rfft", n=n, x=x, kx=kx, L=L, nl=rfft_advect_nl(kx, -1.0),
                u0=np.fft.rfft(u0), nl_coef=-1.0)


KS_SIZES = (256, 512, 1024, 2048, 4096)


def ks_ensemble(count, seed, sizes=KS_SIZES):
    """Seeded KS ensemble: per instance draw n, domain length, 8 amplitudes, 8 phases.

    Draw order per instance: n ~ U{sizes}, Lx ~ U[22, 200], a_k ~ N(0, 0.1^2)
    (k=1..8), theta_k ~ U[0, 2 pi) (k=1..8).  u0 = sum_k a_k cos(2 pi k x / Lx + theta_k).
    """
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(sizes))
        Lx = float(rng.uniform(22.0, 200.0))
        amp = rng.normal(0.0, 0.1, 8)
        th = rng.uniform(0.0, 2 * np.pi, 8)
        out.append(ks_instance(n, Lx, amp, th))
    return out


def ks_instance(n, Lx, amp, th):
    x, kx = x_kx_rfft(n, 0.0, Lx)
    k = np.arange(1, 9)
    u0 = np.sum(amp[:, None] * np.cos(2 * np.pi * k[:, None] * x[None, :] / Lx + th[:, None]),
                axis=0)
    return dict(kind="rfft", n=n, Lx=Lx, x=x, kx=kx, L=kx**2 * (1 - kx**2),
                nl=rfft_advect_nl(kx, -1.0), u0=np.fft.rfft(u0), nl_coef=-1.0)


NLS_SIZES = (1024, 2048, 4096, 8192)


def nls_ensemble(count, seed, sizes=NLS_SIZES, a=-64.0, b=64.0):
    """Seeded NLS ensemble: n ~ U{sizes}, m ~ U{1,2,3} solitons, eta ~ U[1,2.5],
    v ~ U[-2,2], x0 ~ U[-40,40], phase ~ U[0,2 pi) (draw order as listed, per soliton)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(sizes))
        m = int(rng.integers(1, 4))
        x, kx = x_kx_fft(n, a, b)
        u0 = np.zeros(n, dtype=np.complex128)
        for _j in range(m):
            eta = rng.uniform(1.0, 2.5)
            v = rng.uniform(-2.0, 2.0)
            x0 = rng.uniform(-40.0, 40.0)
            ph = rng.uniform(0.0, 2 * np.pi)
            u0 = u0 + nls_soliton_exact(x, 0.0, eta, v, x0) * np.exp(1j * ph)
        out.append(dict(kind="fft", n=n, x=x, kx=kx, L=-0.5j * kx**2, nl=nls_nl(1.0),
                        u0=np.fft.fft(u0), nl_coef=1.0))
    return out


def allen_cahn_2d(ny=512, nx=512, eps=0.01, kcut=32.0, seed=0, batch=1):
    """2-D periodic Allen-Cahn u_t = eps Lap u + u - u^3 on [0, 2 pi)^2 (integer k).

    u0 ~ U[-0.05, 0.05] i.i.d. per point, then every mode with |k| > kcut is zeroed.
    L = -eps (kx^2 + ky^2) flattened over the ny x (nx//2+1) rfft2 modes.
    The linear ``+u`` term is kept inside the nonlinearity ``w - w^3``.
    """
    rng = np.random.default_rng(seed)
    ky = np.fft.fftfreq(ny, d=1.0 / ny)
    kx = np.fft.rfftfreq(nx, d=1.0 / nx)
    KY, KX = np.meshgrid(ky, kx, indexing="ij")
    mask = np.sqrt(KX**2 + KY**2) > kcut
    L = (-eps * (KX**2 + KY**2)).reshape(-1)
    u0s = []
    for _ in range(batch):
        w = rng.uniform(-0.05, 0.05, (ny, nx))
        uh = np.fft.rfft2(w)
        uh[mask] = 0.0
        u0s.append(uh.reshape(-1))
    return dict(kind="rfft2", ny=ny, nx=nx, L=L, nl=ac2d_nl(ny, nx), u0=u0s,
                nl_coef=1.0)


def cheb_allen_cahn_nl(x):
    """models.py:240-245: x - (w + x)^3 on the interior points, complex128."""
    xi = x[1:-1]

    def nl(u):
        val = xi - np.power(u + xi, 3)
        return np.asarray(val, dtype=np.complex128).ravel()
    return nl
