"""Quasi-discrete Hankel transform of order zero (oracle; test infrastructure).

Restates the reference's ``HankelTransform`` (transforms.py:43-240) and its
grid (grids.py:222-264, ``construct_r_kr_hankel``), plus ``mirror_grid``
(grids.py:267-321) and the radial cubic NLS nonlinearity of demos/nls.ipynb
cell 8.  With zeros j_1 < ... < j_{N+1} of J0:

    r_i = j_i rmax / j_{N+1},   k_i = j_i / rmax,
    Y_ij = 2 J0(j_i j_j / j_{N+1}) / (j_{N+1} J1(j_j)^2),
    ht(f) = S Y f,  iht(g) = Y g / S,  S = rmax^2 / j_{N+1}.

Pinned by: tests/golden/hankel.npz (the reference's grids, matrices,
transforms and the radial NLS runs); the reference tests' closed forms (the
transform of a Gaussian, 1e-10); ht/iht round trips.
"""
import numpy as np
import scipy.special as sp


def hankel_grid(nr, rmax):
    """(r, kr, zeros, jN) as construct_r_kr_hankel (grids.py:255-264)."""
    if nr < 4:
        raise ValueError("nr must be >= 4.")
    if rmax <= 0:
        raise ValueError("rmax must be positive.")
    z = sp.jn_zeros(0, nr + 1)
    zeros, jN = z[:-1], z[-1]
    return zeros * rmax / jN, zeros / rmax, zeros, jN


def hankel_matrix(zeros, jN):
    """Y of transforms.py:113-117, in the same operation order."""
    j1vec = sp.j1(zeros)
    arg = np.outer(zeros, zeros) / jN
    return 2 * sp.j0(arg) / (jN * j1vec**2)


class Hankel:
    def __init__(self, nr, rmax=1.0):
        self.nr, self.rmax = nr, rmax
        self.r, self.kr, self.zeros, self.jN = hankel_grid(nr, rmax)
        self.Y = hankel_matrix(self.zeros, self.jN)

    def scale(self):
        return self.rmax**2 / self.jN

    def ht(self, f):
        return self.scale() * self.Y.dot(f)

    def iht(self, g):
        return self.Y.dot(g) / self.scale()


def radial_nls_nl(H, gamma):
    """nl_func of demos/nls.ipynb cell 8: 1j gamma ht(|f|^2 f), f = iht(g)."""
    def nl(g):
        f = H.iht(g)
        f2 = f.real**2 + f.imag**2
        return 1j * gamma * H.ht(f2 * f)
    return nl


def mirror_grid(r, u=None, axis=-1):
    """grids.py:308-321: [-r reversed, r] and the matching mirror of u."""
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
