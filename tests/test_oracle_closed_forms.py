"""Pins of the oracle that do not need the reference: closed forms, brute-force
DFTs, exact solutions, convergence orders and mock-driven controller semantics
(restating tests/test_solveras.py:107-227 of the reference)."""
import math

import mpmath as mp
import numpy as np
import pytest

from oracle import grids, problems, solvers

mp.mp.dps = 40


def exact_psi(r, z):
    z = mp.mpc(z)
    if abs(z) < 1:  # series r! sum z^k / (k + r)!: no cancellation near 0
        return mp.factorial(r) * mp.nsum(lambda k: z**k / mp.factorial(k + r), [0, mp.inf])
    if r == 1:
        return (mp.e**z - 1) / z
    if r == 2:
        return 2 * (mp.e**z - 1 - z) / z**2
    return 6 * (mp.e**z - 1 - z - z**2 / 2) / z**3


@pytest.mark.parametrize("z", [0.5, -3.0, 2.5j, -1 + 1j, 40.0, -200.0])
def test_psi_closed_forms_vs_mpmath(z):
    for r, f in ((1, solvers.psi1), (2, solvers.psi2), (3, solvers.psi3)):
        got = complex(f(np.array([z], dtype=np.complex128))[0])
        ref = complex(exact_psi(r, z))
        assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))


def test_psi3_cancellation_near_cutoff_is_the_references_own_error():
    """Direct psi3 at |z| = modecutoff loses ~eps/|z|^3: the reference accepts ~1e-10 there."""
    z = np.array([0.0101, 0.0101j, -0.0101], dtype=np.complex128)
    errs = [abs(complex(solvers.psi3(z[i:i + 1])[0]) - complex(exact_psi(3, z[i]))) for i in range(3)]
    assert max(errs) < 1e-8
    assert max(errs) > 1e-14  # the cancellation is real: not exact


@pytest.mark.parametrize("z", [0.0, 1e-9, 0.004 - 0.007j, 0.0099j])
def test_contour_mean_equals_psi(z):
    """Mean of psi over the contour = psi at the centre (mean value property), etd35.py:255-264."""
    etd = solvers.ETDConfig()
    Z = z + solvers.contour_nodes(etd)
    for r, f in ((1, solvers.psi1), (2, solvers.psi2), (3, solvers.psi3)):
        got = complex(np.mean(f(Z)))
        ref = complex(exact_psi(r, z))
        assert abs(got - ref) < 1e-14


def dft(x):
    n = len(x)
    k = np.arange(n)
    return np.exp(-2j * np.pi * np.outer(k, k) / n) @ x


def test_advective_nl_vs_brute_force_dft():
    n = 16
    x, kx = grids.x_kx_rfft(n, 0.0, 2 * np.pi)
    u = np.cos(x) + 0.3 * np.sin(2 * x) + 0.1
    uf = np.fft.rfft(u)
    full = dft(u)
    kfull = 2 * np.pi * np.fft.fftfreq(n, x[1] - x[0])
    kfull[n // 2] = 0.0  # irfft ignores the imaginary Nyquist part, so i k u_hat contributes nothing there
    ux = np.real(np.conj(dft(np.conj(1j * kfull * full))) / n)  # inverse DFT via conjugation
    ref = -dft(u * ux)[: n // 2 + 1]
    got = problems.rfft_advect_nl(kx, -1.0)(uf)
    assert np.max(np.abs(got - ref)) < 1e-12


def test_nls_nl_vs_brute_force_dft():
    n = 16
    rng = np.random.default_rng(1)
    v = rng.normal(size=n) + 1j * rng.normal(size=n)
    ref = 1j * dft(np.abs(v) ** 2 * v)
    got = problems.nls_nl(1.0)(dft(v))
    assert np.max(np.abs(got - ref)) < 1e-12


def test_allen_cahn_2d_nl_single_mode_closed_form():
    """w = a cos(3x): w - w^3 = (a - 3a^3/4) cos(3x) - (a^3/4) cos(9x)."""
    ny = nx = 32
    a = 0.4
    x = np.arange(nx) * 2 * np.pi / nx
    w = np.tile(a * np.cos(3 * x), (ny, 1))
    nl = problems.ac2d_nl(ny, nx)
    out = nl(np.fft.rfft2(w).reshape(-1)).reshape(ny, nx // 2 + 1)
    expect = np.zeros_like(out)
    expect[0, 3] = (a - 0.75 * a**3) * ny * nx / 2
    expect[0, 9] = -(a**3 / 4) * ny * nx / 2
    assert np.max(np.abs(out - expect)) < 1e-10


def test_nls_single_soliton_exact():
    n = 512
    x, kx = grids.x_kx_fft(n, -32.0, 32.0)
    eta, v = 1.5, 0.7
    u0 = np.fft.fft(problems.nls_soliton_exact(x, 0.0, eta, v, -3.0))
    ue = np.fft.fft(problems.nls_soliton_exact(x, 2.0, eta, v, -3.0))
    drv = solvers.make_solver("ETD35", -0.5j * kx**2, problems.nls_nl(1.0), solvers.SolverConfig(epsilon=1e-9))
    uf = drv.evolve(u0, t0=0.0, tf=2.0, store_data=False)
    assert solvers.rel_l2(uf, ue) < 1e-7


@pytest.mark.parametrize("method,bar", [("ETD4", 1e-6), ("ETD5", 1e-6), ("IF4", 1e-5)])
def test_kdv_soliton_constant_step_bars(method, bar):
    """The reference's analytic bars (test_etd4.py:19, test_etd5.py:21, test_if4.py:35)."""
    p = problems.kdv_fixture()
    drv = solvers.make_solver(method, p["L"], p["nl"])
    uf = drv.evolve(p["u0"], t0=0.0, tf=p["tf"], h=p["h"], store_data=False)
    assert solvers.rel_l2(uf, p["exact"]) < bar


@pytest.mark.parametrize("method", ["ETD4", "ETD5", "IF4"])
def test_linear_problem_is_exact_exponential(method):
    """nl = 0: every step is u <- exp(hL) u (test_if4.py:108-130)."""
    L = np.array([-1.0, -0.2, 0.3j, -2.0 + 1j, 0.0])
    u0 = np.array([1.0, 2.0, 0.5j, 1 - 1j, 3.0], dtype=np.complex128)
    drv = solvers.make_solver(method, L, lambda u: np.zeros_like(u))
    uf = drv.evolve(u0, t0=0.0, tf=1.0, h=0.1, store_data=False)
    np.testing.assert_allclose(uf, np.exp(L * drv.tc) * u0, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("method,order", [("ETD4", 3.5), ("IF4", 3.5), ("ETD5", 4.5)])
def test_convergence_order_kdv(method, order):
    """Self-convergence slope (test_order_convergence.py:221-306): h = 1/32, 1/64, 1/128 vs 1/1024
    (binary fractions, so every run lands on tf exactly)."""
    p = problems.kdv_fixture()
    tf = 1.0

    def run(h):
        drv = solvers.make_solver(method, p["L"], p["nl"])
        return drv.evolve(p["u0"], t0=0.0, tf=tf, h=h, store_data=False)
    hs = [1 / 32, 1 / 64, 1 / 128]
    ref = run(1 / 1024)
    errs = [solvers.rel_l2(run(h), ref) for h in hs]
    slope = np.polyfit(np.log(hs), np.log(errs), 1)[0]
    assert slope > order, (errs, slope)


# -- controller semantics with canned (unew, err), restating test_solveras.py ----
class Canned:
    adaptive = True

    def __init__(self, unew, err, q=3):
        self.unew, self.err, self.q = unew, err, q

    def reset(self):
        pass

    def attempt(self, u, h, accept):
        return self.unew, self.err


def test_step_accepts_and_grows():
    d = solvers.AdaptiveDriver(Canned(np.array([2.0]), np.array([1e-12])))
    unew, h, hn = d.step(np.array([1.0]), 0.1)
    assert h == 0.1 and hn > 0.1 and d.accept


def test_step_rejects_until_max_loops():
    d = solvers.AdaptiveDriver(Canned(np.array([1.0]), np.array([1.0])), solvers.SolverConfig(epsilon=1e-4))
    d.MAX_LOOPS = 1
    with pytest.raises(solvers.MaxLoopsExceeded):
        d.step(np.array([1.0]), 0.1)


def test_step_minimum_step_reached():
    d = solvers.AdaptiveDriver(Canned(np.array([1.0]), np.array([1.0])), solvers.SolverConfig(minh=1e-6))
    d.MAX_LOOPS = 100
    with pytest.raises(solvers.MinimumStepReached):
        d.step(np.array([1.0]), 1e-5)


def test_reject_nan_inf_and_clamp():
    d = solvers.AdaptiveDriver(Canned(None, None))
    assert math.isclose(d.reject_h(np.nan, 0.1), d.MIN_S * 0.1)
    assert math.isclose(d.reject_h(np.inf, 0.1), d.MIN_S * 0.1)
    d.config.decr_f = 0.5
    d.MIN_S = 0.2
    assert math.isclose(d.reject_h(0.9, 1.0), 0.5)


def test_accept_growth_threshold():
    d = solvers.AdaptiveDriver(Canned(None, None))
    d.MAX_S = 10.0
    assert d.accept_h(2.0, 0.1) > 0.1
    d.config.incr_f = 5.0
    assert d.accept_h(2.0, 0.1) == 0.1


def test_pure_linear_adaptive_never_converges():
    """err == 0 => s = inf => reject forever (solveras.py:454,391): h shrinks by MIN_S per
    attempt and crosses minh before MAX_LOOPS attempts."""
    L = np.array([-1.0, -2.0])
    drv = solvers.make_solver("IF34", L, lambda u: np.zeros_like(u))
    with pytest.raises(solvers.MinimumStepReached):
        drv.evolve(np.array([1.0, 1.0], dtype=np.complex128), t0=0.0, tf=1.0, store_data=False)
    drv = solvers.make_solver("IF34", L, lambda u: np.zeros_like(u), solvers.SolverConfig(minh=1e-300))
    with pytest.raises(solvers.MaxLoopsExceeded):
        drv.evolve(np.array([1.0, 1.0], dtype=np.complex128), t0=0.0, tf=1.0, store_data=False)


def test_evolve_zero_or_negative_interval_returns_u0():
    drv = solvers.make_solver("ETD35", np.array([1.0]), lambda u: u**2)
    u0 = np.array([0.5 + 0j])
    assert np.allclose(drv.evolve(u0, 1.0, 0.0, 0.1, False), u0)
    assert np.allclose(drv.evolve(u0, 0.0, 0.0, 0.1, False), u0)


def test_constant_step_evolve_rejects_oversized_h():
    drv = solvers.make_solver("ETD4", np.array([-1.0]), lambda u: 0 * u)
    with pytest.raises(ValueError):
        drv.evolve(np.array([1.0 + 0j]), 0.0, 0.1, 0.2)
