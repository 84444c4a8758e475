"""The oracle reproduces the reference bitwise on every golden fixture.

Fixtures were written by tests/golden/make_golden.py from the reference.
Bitwise equality is expected because the oracle keeps NumPy's operation order.
"""
import numpy as np
import pytest

from oracle import grids, problems, solvers

ADAPTIVE = ("ETD35", "ETD34", "IF34", "IF45DP")
CONSTANT = ("ETD4", "ETD5", "IF4")


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    assert np.array_equal(a, b, equal_nan=True), float(np.max(np.abs(a - b)))


def run_logged(name, L, nl, u0, eps, **kw):
    drv = solvers.make_solver(name, L, nl, solvers.SolverConfig(epsilon=eps))
    uf = drv.evolve(u0, **kw)
    h = np.array([e[0] for e in drv.log])
    s = np.array([e[1] for e in drv.log])
    return drv, uf, h, s


def test_ks1024_if34_bitwise(golden):
    g = golden("ks1024_if34")
    p = problems.ks_readme()
    same(p["L"], g["L"])
    same(p["u0"], g["u0"])
    drv, uf, h, s = run_logged("IF34", p["L"], p["nl"], p["u0"], 1e-4, t0=0.0, tf=50.0,
                               store_freq=20)
    same(h, g["log_h"])
    same(s, g["log_s"])
    same(uf, g["u_final"])
    same(drv.t, g["t"])
    same([np.linalg.norm(v) for v in drv.u], g["snap_norm"])


def test_nls8192_etd35_bitwise(golden):
    g = golden("nls8192_etd35")
    p = problems.nls_two_soliton()
    same(p["L"], g["L"])
    same(p["u0"], g["u0"])
    _, uf, h, s = run_logged("ETD35", p["L"], p["nl"], p["u0"], 1e-6, t0=0.0, tf=20.0,
                             store_data=False)
    same(h, g["log_h"])
    same(s, g["log_s"])
    same(uf, g["u_final"])


BUILDERS = {
    "ETD35": solvers.etd35_coeffs, "ETD34": solvers.etd34_coeffs,
    "ETD4": solvers.etd4_coeffs, "ETD5": solvers.etd5_coeffs,
}


@pytest.mark.parametrize("method", ["ETD35", "ETD34", "ETD4", "ETD5", "IF34", "IF4", "IF45DP"])
@pytest.mark.parametrize("lname", ["real", "imag", "cplx"])
def test_coefficients_bitwise(golden, method, lname):
    g = golden("coeffs")
    L = g[f"L/{lname}"]
    etd = solvers.ETDConfig()
    for hi, h in enumerate(g["h"]):
        h = float(h)
        if method in BUILDERS:
            c = BUILDERS[method](L, h, etd)
        elif method == "IF45DP":
            c = solvers.if45dp_coeffs(L, h)
        else:
            c = solvers.if_exp_coeffs(L, h)
        prefix = f"{method}/{lname}/{hi}/"
        keys = [k for k in g if k.startswith(prefix)]
        assert keys
        for k in keys:
            name = k[len(prefix):]
            if method in ("ETD35", "ETD5") and name == "a64":
                # allocated, never assigned (etd35.py:144, etd5.py:107)
                assert not np.any(g[k])
                continue
            same(c[name], g[k])


@pytest.mark.parametrize("pname", ["ks64", "nls64"])
@pytest.mark.parametrize("method", ADAPTIVE + CONSTANT)
def test_three_steps_bitwise(golden, pname, method):
    g = golden("attempts")
    L, u = g[f"{pname}/L"], g[f"{pname}/u0"]
    if pname == "ks64":
        _, kx = grids.x_kx_rfft(64, 0.0, 32.0 * np.pi)
        nl = problems.rfft_advect_nl(kx, -1.0)
        h0 = 0.05
    else:
        nl = problems.nls_nl(1.0)
        h0 = 0.02
    drv = solvers.make_solver(method, L, nl, solvers.SolverConfig(epsilon=1e-3))
    v, h = u.copy(), h0
    for i in range(3):
        if method in ADAPTIVE:
            v, hu, h = drv.step(v, h)
            assert hu == g[f"{pname}/{method}/hused{i + 1}"]
            assert h == g[f"{pname}/{method}/hnext{i + 1}"]
        else:
            v = drv.step(v, h0)
        same(v, g[f"{pname}/{method}/u{i + 1}"])
    if method in ADAPTIVE:
        same([e[0] for e in drv.log], g[f"{pname}/{method}/log_h"])
        same([e[1] for e in drv.log], g[f"{pname}/{method}/log_s"])


@pytest.mark.parametrize("method", ADAPTIVE + CONSTANT)
def test_kdv_all_methods_bitwise(golden, method):
    g = golden("kdv_burgers")
    p = problems.kdv_fixture()
    same(p["u0"], g["u0"])
    same(p["L"], g["L"])
    if method in ADAPTIVE:
        drv, uf, h, s = run_logged(method, p["L"], p["nl"], p["u0"], 1e-4, t0=0.0, tf=p["tf"],
                                   h_init=p["h"], store_data=True, store_freq=10)
        same(h, g[f"{method}/log_h"])
        same(s, g[f"{method}/log_s"])
    else:
        drv = solvers.make_solver(method, p["L"], p["nl"])
        uf = drv.evolve(p["u0"], t0=0.0, tf=p["tf"], h=p["h"], store_data=True, store_freq=10)
    same(uf, g[f"{method}/u_final"])
    same(drv.t, g[f"{method}/t"])


@pytest.mark.parametrize("method,eps", [("IF34", 1e-4), ("IF45DP", 1e-3)])
def test_burgers_bitwise(golden, method, eps):
    g = golden("kdv_burgers")
    p = problems.burgers_fixture()
    same(p["u0"], g["burgers/u0"])
    _, uf, h, s = run_logged(method, p["L"], p["nl"], p["u0"], eps, t0=0, tf=0.85,
                             store_data=False)
    same(h, g[f"burgers/{method}/log_h"])
    same(s, g[f"burgers/{method}/log_s"])
    same(uf, g[f"burgers/{method}/u_final"])


@pytest.mark.parametrize("method", ADAPTIVE + CONSTANT)
def test_nls256_bitwise(golden, method):
    g = golden("nls256")
    L, u0 = g["L"], g["u0"]
    if method in ADAPTIVE:
        _, uf, _, _ = run_logged(method, L, problems.nls_nl(1.0), u0, 1e-5, t0=0.0, tf=2.0,
                                 store_data=False)
    else:
        drv = solvers.make_solver(method, L, problems.nls_nl(1.0))
        uf = drv.evolve(u0, t0=0.0, tf=2.0, h=0.01, store_data=False)
    same(uf, g[f"{method}/u_final"])


# -- diagonalized strategies (etd35.py:348-495, etd34.py:194-304, if34.py:137-196) --
def _cubic(a):
    return lambda u: a * u - u * u * u


@pytest.mark.parametrize("pname", ["real", "cplx"])
@pytest.mark.parametrize("method", ["ETD35", "ETD34", "IF34"])
def test_diagonalized_evolve_bitwise(golden, pname, method):
    g = golden("diagonalized")
    L, u0, a = g[f"{pname}/L"], g[f"{pname}/u0"], float(g[f"{pname}/a"])
    drv = solvers.make_solver(method, L, _cubic(a), solvers.SolverConfig(epsilon=1e-6), diagonalize=True)
    uf = drv.evolve(u0, t0=0.0, tf=3.0, store_data=True, store_freq=5)
    np.testing.assert_array_equal(uf, g[f"{pname}/{method}/u_final"])
    np.testing.assert_array_equal([x[0] for x in drv.log], g[f"{pname}/{method}/log_h"])
    np.testing.assert_array_equal([x[1] for x in drv.log], g[f"{pname}/{method}/log_s"])
    np.testing.assert_array_equal(drv.t, g[f"{pname}/{method}/t"])


@pytest.mark.parametrize("pname", ["real", "cplx"])
@pytest.mark.parametrize("method", ["ETD35", "ETD34", "IF34"])
def test_diagonalized_three_steps_bitwise(golden, pname, method):
    g = golden("diagonalized")
    L, u0, a = g[f"{pname}/L"], g[f"{pname}/u0"], float(g[f"{pname}/a"])
    drv = solvers.make_solver(method, L, _cubic(a), solvers.SolverConfig(epsilon=1e-4), diagonalize=True)
    v, h = u0, 0.2
    for i in range(3):
        v, hu, h = drv.step(v, h)
        np.testing.assert_array_equal(v, g[f"{pname}/{method}/u{i + 1}"])
        assert hu == g[f"{pname}/{method}/hused{i + 1}"] and h == g[f"{pname}/{method}/hnext{i + 1}"]


def test_diagonalized_rejects_bad_operators():
    with pytest.raises(ValueError):
        solvers.make_solver("ETD35", np.ones(4), _cubic(0.1), diagonalize=True)
    with pytest.raises(ValueError):
        solvers.make_solver("ETD35", np.zeros((3, 3)), _cubic(0.1), diagonalize=True)
    with pytest.raises(ValueError):
        solvers.make_solver("IF45DP", np.eye(3), _cubic(0.1), diagonalize=True)


# -- dense strategies (etd35.py:498-720, etd34.py:307-436, etd4.py:178-308, etd5.py:264-482,
#    if34.py:202-254, if4.py:125-199) --
@pytest.mark.parametrize("pname", ["real", "cplx"])
@pytest.mark.parametrize("method", ["ETD35", "ETD34", "IF34"])
def test_dense_adaptive_bitwise(golden, pname, method):
    from oracle import dense
    g = golden("dense")
    L, u0, a = g[f"{pname}/L"], g[f"{pname}/u0"], float(g[f"{pname}/a"])
    drv = dense.make_solver(method, L, _cubic(a), solvers.SolverConfig(epsilon=1e-6))
    uf = drv.evolve(u0, t0=0.0, tf=2.0, store_data=True, store_freq=4)
    np.testing.assert_array_equal(uf, g[f"{pname}/{method}/u_final"])
    np.testing.assert_array_equal([x[0] for x in drv.log], g[f"{pname}/{method}/log_h"])
    np.testing.assert_array_equal(drv.t, g[f"{pname}/{method}/t"])


@pytest.mark.parametrize("pname", ["real", "cplx"])
@pytest.mark.parametrize("method", ["ETD4", "ETD5", "IF4"])
def test_dense_constant_bitwise(golden, pname, method):
    from oracle import dense
    g = golden("dense")
    L, u0, a = g[f"{pname}/L"], g[f"{pname}/u0"], float(g[f"{pname}/a"])
    drv = dense.make_solver(method, L, _cubic(a))
    uf = drv.evolve(u0, t0=0.0, tf=1.0, h=0.025, store_data=False)
    np.testing.assert_array_equal(uf, g[f"{pname}/{method}/u_final"])


def test_dense_matches_diagonalized_on_a_diagonalizable_operator(golden):
    """Same operator, two strategies: the dense contour psi and the eigenbasis scheme agree."""
    from oracle import dense
    g = golden("dense")
    L, u0, a = g["real/L"], g["real/u0"], float(g["real/a"])
    cfg = solvers.SolverConfig(epsilon=1e-10)
    d1 = dense.make_solver("ETD35", L, _cubic(a), cfg)
    d2 = solvers.make_solver("ETD35", L, _cubic(a), cfg, diagonalize=True)
    u1 = d1.evolve(u0, t0=0.0, tf=1.0, store_data=False)
    u2 = d2.evolve(u0, t0=0.0, tf=1.0, store_data=False)
    assert solvers.rel_l2(u1, u2) < 1e-8


# -- demos/allen_cahn.ipynb: Chebyshev grid, dense and diagonalized strategies --
@pytest.mark.parametrize("strategy", ["dense", "diag"])
@pytest.mark.parametrize("method", ["ETD35", "ETD34", "IF34"])
def test_cheb_allen_cahn_bitwise(golden, strategy, method):
    from oracle import dense
    g = golden("cheb_allen_cahn")
    nl = problems.cheb_allen_cahn_nl(g["x"])
    cfg = solvers.SolverConfig(epsilon=1e-4)
    if strategy == "dense":
        etd = solvers.ETDConfig(contour_points=64, contour_radius=20)
        drv = dense.make_solver(method, g["L"], nl, cfg, etd)
    else:
        etd = solvers.ETDConfig(contour_points=32, contour_radius=1)
        drv = solvers.make_solver(method, g["L"], nl, cfg, etd, diagonalize=True)
    uf = drv.evolve(g["w0"], t0=0.0, tf=10.0, store_data=True, store_freq=10)
    np.testing.assert_array_equal(uf, g[f"{strategy}/{method}/u_final"])
    np.testing.assert_array_equal([x[0] for x in drv.log], g[f"{strategy}/{method}/log_h"])
    np.testing.assert_array_equal(drv.t, g[f"{strategy}/{method}/t"])
