"""Hankel transform oracle (oracle/transforms.py) against the reference's golden
vectors (tests/golden/hankel.npz) and the closed forms of the reference's own
tests (tests/test_transforms.py, tests/test_grids.py); the product's grid helpers
against the oracle (no device needed)."""
import numpy as np
import pytest
import torch

from oracle import solvers
from oracle import transforms as otr
from rkstiff_amd import grids


@pytest.mark.parametrize("key", ["12_2", "50_2", "100_5"])
def test_grid_matrix_and_transforms_match_reference(golden, key):
    g = golden("hankel")
    nr, rmax = int(key.split("_")[0]), float(key.split("_")[1])
    H = otr.Hankel(nr, rmax)
    np.testing.assert_array_equal(H.r, g[f"{key}/r"])
    np.testing.assert_array_equal(H.kr, g[f"{key}/kr"])
    np.testing.assert_array_equal(H.zeros, g[f"{key}/zeros"])
    np.testing.assert_array_equal(H.Y, g[f"{key}/Y"])
    f = g[f"{key}/f"]
    np.testing.assert_allclose(H.ht(f), g[f"{key}/ht_f"], rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(H.iht(f), g[f"{key}/iht_f"], rtol=1e-14, atol=1e-300)


def test_closed_forms_of_reference_tests():
    # Gaussian: G(k) = exp(-k^2 / (4 a^2)) / (2 a^2)
    H = otr.Hankel(50, 2.0)
    a = 4
    ex = np.exp(-H.kr**2 / (4 * a**2)) / (2 * a**2)
    assert np.linalg.norm(H.ht(np.exp(-(a**2) * H.r**2)) - ex) / np.linalg.norm(ex) < 1e-10
    # modulated Gaussian
    H = otr.Hankel(50, 4.0)
    sigma, w = 2, 0.5
    om = 1.0 / (4 * (sigma**2 + w**2))
    ex = -2 * om * np.exp(-sigma * om * H.kr**2) * (-w * np.cos(w * om * H.kr**2) + sigma * np.sin(w * om * H.kr**2))
    f = np.exp(-sigma * H.r**2) * np.sin(w * H.r**2)
    assert np.linalg.norm(H.ht(f) - ex) / np.linalg.norm(ex) < 1e-10
    # exponential decay: a / (a^2 + k^2)^(3/2), coarse grid
    H = otr.Hankel(25, 3.0)
    ex = 4 * np.power(16 + H.kr**2, -1.5)
    assert np.linalg.norm(H.ht(np.exp(-4 * H.r)) - ex) / np.linalg.norm(ex) < 1e-2
    # round trips
    H = otr.Hankel(40, 4.0)
    for f in (np.exp(-H.r**2), np.exp(-2 * H.r), np.exp(-H.r**2) * np.cos(H.r)):
        assert np.linalg.norm(H.iht(H.ht(f)) - f) / np.linalg.norm(f) < 1e-9
    H = otr.Hankel(30, 3.0)
    G = np.exp(-0.5 * H.kr**2)
    assert np.linalg.norm(H.ht(H.iht(G)) - G) / np.linalg.norm(G) < 1e-10
    # first zero of J0
    assert abs(otr.Hankel(10, 2.0).zeros[0] - 2.4048) < 1e-3
    with pytest.raises(ValueError):
        otr.hankel_grid(3, 1.0)
    with pytest.raises(ValueError):
        otr.hankel_grid(8, 0.0)


@pytest.mark.parametrize("method", ["ETD35", "ETD34", "IF34"])
def test_radial_nls_demo_matches_reference(golden, method):
    g = golden("hankel")
    H = otr.Hankel(100, 5.0)
    nl = otr.radial_nls_nl(H, 2)
    np.testing.assert_array_equal(nl(g["nls/u0"]), g[f"nls/{method}/nl_u0"])
    drv = solvers.make_solver(method, g["nls/L"], nl, solvers.SolverConfig(epsilon=1e-5))
    uf = drv.evolve(g["nls/u0"], t0=0.0, tf=0.8, store_data=True, store_freq=10)
    np.testing.assert_array_equal(uf, g[f"nls/{method}/u_final"])
    np.testing.assert_array_equal(np.array(drv.t), g[f"nls/{method}/t"])
    log = np.array([(h, s) for h, s, _ in drv.log])
    np.testing.assert_array_equal(log[:, 0], g[f"nls/{method}/log_h"])
    np.testing.assert_array_equal(log[:, 1], g[f"nls/{method}/log_s"])


def test_product_grids_match_oracle():
    r, kr, z, jN = grids.construct_r_kr_hankel(50, 2.0)
    ro, kro, zo, jNo = otr.hankel_grid(50, 2.0)
    np.testing.assert_array_equal(r.numpy(), ro)
    np.testing.assert_array_equal(kr.numpy(), kro)
    np.testing.assert_array_equal(z.numpy(), zo)
    assert jN == jNo
    with pytest.raises(ValueError):
        grids.construct_r_kr_hankel(3, 1.0)
    with pytest.raises(ValueError):
        grids.construct_r_kr_hankel(8, -1.0)


def test_mirror_grid():
    # tests/test_grids.py:67-105 restated for NumPy and torch inputs
    r = np.array([0, 1, 2, 3])
    u = np.array([10, 20, 30, 40])
    for rr, uu in ((r, u), (torch.as_tensor(r), torch.as_tensor(u))):
        rnew, unew = grids.mirror_grid(rr, uu)
        rnew, unew = np.asarray(rnew), np.asarray(unew)
        np.testing.assert_array_equal(rnew, [-3, -2, -1, 0, 0, 1, 2, 3])
        np.testing.assert_array_equal(unew, [40, 30, 20, 10, 10, 20, 30, 40])
        o_r, o_u = otr.mirror_grid(r, u)
        np.testing.assert_array_equal(rnew, o_r)
        np.testing.assert_array_equal(unew, o_u)
    rn, un = grids.mirror_grid(np.array([0, 1, 2]))
    assert un is None and rn.shape == (6,)
    m = np.array([[1, 2, 3], [4, 5, 6]])
    assert grids.mirror_grid(np.array([0, 1, 2]), m, axis=0)[1].shape == (4, 3)
    assert grids.mirror_grid(np.array([0, 1, 2]), m, axis=1)[1].shape == (2, 6)
    assert grids.mirror_grid(torch.arange(3), torch.as_tensor(m), axis=1)[1].shape == (2, 6)
    with pytest.raises(ValueError):
        grids.mirror_grid(np.array([0, 1, 2]), np.array([1, 2, 3]), axis=99)


def test_hankel_transform_needs_the_device():
    from rkstiff_amd.transforms import HankelTransform
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    with pytest.raises(RuntimeError):
        HankelTransform(16, 1.0)
