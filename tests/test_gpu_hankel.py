"""Hankel transform and the radial NLS on the device against the reference's golden
vectors (tests/golden/hankel.npz) and the closed forms of its tests."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tr():
    import rkstiff_amd.transforms as t
    return t


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("key", ["12_2", "50_2", "100_5"])
def test_transforms_match_reference(tr, golden, key):
    g = golden("hankel")
    nr, rmax = int(key.split("_")[0]), float(key.split("_")[1])
    ht = tr.HankelTransform(nr, rmax)
    np.testing.assert_array_equal(ht.hankel_matrix().numpy(), g[f"{key}/Y"])
    np.testing.assert_array_equal(ht.bessel_zeros().numpy(), g[f"{key}/zeros"])
    np.testing.assert_array_equal(ht.r.cpu().numpy(), g[f"{key}/r"])
    np.testing.assert_array_equal(ht.kr.cpu().numpy(), g[f"{key}/kr"])
    f = torch.as_tensor(g[f"{key}/f"]).cuda()
    a, b = ht.ht(f).cpu().numpy(), ht.iht(f).cpu().numpy()
    print(key, rel(a, g[f"{key}/ht_f"]), rel(b, g[f"{key}/iht_f"]))
    assert rel(a, g[f"{key}/ht_f"]) < 1e-14 and rel(b, g[f"{key}/iht_f"]) < 1e-14
    # real input, and a batch of rows in one launch
    fr = f.real.contiguous()
    np.testing.assert_allclose(ht.ht(fr).cpu().numpy(), g[f"{key}/ht_f"].real, rtol=1e-13,
                               atol=1e-15 * np.abs(g[f"{key}/ht_f"]).max())
    batch = torch.stack([f * (k + 1) for k in range(37)])
    out = ht.ht(batch).cpu().numpy()
    for k in range(37):
        assert rel(out[k], (k + 1) * g[f"{key}/ht_f"]) < 1e-14


def test_reference_test_cases_on_device(tr):
    ht = tr.HankelTransform(nr=50, rmax=2.0)
    a = 4
    ex = torch.exp(-ht.kr**2 / (4 * a**2)) / (2 * a**2)
    assert float(torch.linalg.norm(ht.ht(torch.exp(-(a**2) * ht.r**2)) - ex) / torch.linalg.norm(ex)) < 1e-10
    ht = tr.HankelTransform(nr=32, rmax=5.0)
    f = torch.exp(-ht.r**2)
    assert float(torch.linalg.norm(ht.iht(ht.ht(f)) - f) / torch.linalg.norm(f)) < 1e-10
    ht = tr.HankelTransform(nr=8, rmax=2.0)
    r0 = ht.r.clone()
    ht.nr = 16
    assert ht.nr == 16 and ht.r.shape[0] == 16 and ht.hankel_matrix().shape == (16, 16)
    with pytest.raises(ValueError, match="nr must be an integer"):
        ht.nr = 10.5
    with pytest.raises(ValueError, match="nr must be an integer ≥ 4"):
        ht.nr = 3
    assert ht.nr == 16
    ht = tr.HankelTransform(nr=8, rmax=2.0)
    ht.rmax = 4.0
    assert ht.rmax == 4.0
    np.testing.assert_allclose(ht.r.cpu().numpy(), 2.0 * r0.cpu().numpy(), rtol=1e-10)
    with pytest.raises(ValueError, match="rmax must be positive"):
        ht.rmax = 0.0
    assert ht.rmax == 4.0
    with pytest.raises(ValueError):
        ht.ht(torch.zeros(7, dtype=torch.float64, device="cuda"))
    with pytest.raises(TypeError):
        ht.ht(np.zeros(8))


@pytest.mark.parametrize("method", ["ETD35", "ETD34", "IF34"])
def test_radial_nls_demo(tr, golden, method):
    import rkstiff_amd as rk
    g = golden("hankel")
    ht = tr.HankelTransform(nr=100, rmax=5)
    nl = ht.cubic_nl(2.0)
    u0 = torch.as_tensor(g["nls/u0"]).cuda()
    assert rel(nl(u0).cpu().numpy(), g[f"nls/{method}/nl_u0"]) < 1e-13
    L = -1j * ht.kr**2
    np.testing.assert_array_equal(L.cpu().numpy(), g["nls/L"])
    solver = getattr(rk, method)(lin_op=L, nl_func=nl, config=rk.SolverConfig(epsilon=1e-5))
    uf = solver.evolve(u0, t0=0.0, tf=0.8, store_data=True, store_freq=10)
    err = rel(uf.cpu().numpy(), g[f"nls/{method}/u_final"])
    t = np.array([float(x) for x in solver.t])
    print(method, "radial NLS rel", err, "stored", len(t))
    assert err < 1e-10
    # same snapshots; the step sizes differ in their last digits (the GPU sums Y rows in
    # another order, and the embedded error estimate cancels those digits into s)
    assert len(t) == len(g[f"nls/{method}/t"])
    np.testing.assert_allclose(t, g[f"nls/{method}/t"], rtol=1e-8)
