"""GPU parity of the 2-D rfft2 engine (Allen-Cahn term) against the oracle.

The oracle's rfft2 nonlinearity is pinned by a closed form
(test_oracle_closed_forms.py) and its methods/controller by the reference's
golden vectors; here the engine runs the same methods on flattened N x N
rfft2 problems, as the reference requires of multi-D grids (solver.py:129-135).
"""
import numpy as np
import pytest

from oracle import problems, solvers

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ADAPTIVE = ("ETD35", "ETD34", "IF34", "IF45DP")
CONSTANT = ("ETD4", "ETD5", "IF4")


@pytest.fixture(scope="module")
def rk():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import rkstiff_amd
    from rkstiff_amd import _lib
    _lib.load()
    return rkstiff_amd


def dev(a, dtype=torch.complex128):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).cuda()


def host(t):
    return t.detach().cpu().numpy().astype(np.complex128)


def ac(n, seed=0, batch=1, kcut=None):
    return problems.allen_cahn_2d(n, n, eps=0.01, kcut=kcut if kcut is not None else n / 8, seed=seed,
                                  batch=batch)


@pytest.mark.parametrize("n", [32, 64, 128, 256, 512])
def test_ac2d_nl_matches_oracle(rk, n):
    print(f"ac2d nl n={n}", flush=True)
    p = ac(n, seed=n, kcut=n / 4)
    u = p["u0"][0] * 40.0  # amplitude ~1, so the cubic matters
    ref = p["nl"](u)
    nl = rk.AllenCahnNL(n)
    got = host(nl(dev(u)))
    assert solvers.rel_l2(got, ref) < 1e-13
    got32 = host(nl(dev(u, torch.complex64)))
    assert solvers.rel_l2(got32, ref) < 2e-6


@pytest.mark.parametrize("method", ADAPTIVE + CONSTANT)
def test_ac2d_fp64_matches_oracle(rk, method):
    """Whole runs on 32 x 32 (adaptive: same attempt sequence as the oracle)."""
    print(f"ac2d {method}", flush=True)
    p = ac(32, seed=1, kcut=8)
    u0 = p["u0"][0] * 20.0
    L, nl = rk.allen_cahn_ops(32)
    np.testing.assert_array_equal(L.numpy(), p["L"])
    cls = getattr(rk, method)
    # the reference's IF45DP error weight r4 = 17h/1920 (if45dp.py:234; Dormand-Prince has
    # 71/1920) leaves an O(h) error estimate, so it needs h ~ 1e-5 here: a shorter interval
    tf = 0.05 if method == "IF45DP" else 2.0
    if method in ADAPTIVE:
        solver = cls(L, nl, config=rk.SolverConfig(epsilon=1e-6))
        solver.log_capacity = 8192
        uf = solver.evolve(dev(u0), t0=0.0, tf=tf, store_data=False)
        drv = solvers.make_solver(method, p["L"], p["nl"], solvers.SolverConfig(epsilon=1e-6))
        ref = drv.evolve(u0, t0=0.0, tf=tf, store_data=False)
        log = solver.engine_log()
        assert len(log) == len(drv.log)
        np.testing.assert_allclose(log[:, 0], [x[0] for x in drv.log], rtol=1e-9)
    else:
        solver = cls(L, nl)
        uf = solver.evolve(dev(u0), t0=0.0, tf=2.0, h=0.05, store_data=False)
        drv = solvers.make_solver(method, p["L"], p["nl"])
        ref = drv.evolve(u0, t0=0.0, tf=2.0, h=0.05, store_data=False)
    err = solvers.rel_l2(host(uf), ref)
    print(method, "rel", err)
    assert err < 1e-11


def test_ac2d_step_and_snapshots(rk):
    p = ac(64, seed=2, kcut=12)
    u0 = p["u0"][0] * 20.0
    L, nl = rk.allen_cahn_ops(64)
    solver = rk.ETD35(L, nl, config=rk.SolverConfig(epsilon=1e-6))
    solver.SNAP_CHUNK = 4
    solver.evolve(dev(u0), t0=0.0, tf=1.0, store_freq=2)
    drv = solvers.make_solver("ETD35", p["L"], p["nl"], solvers.SolverConfig(epsilon=1e-6))
    drv.evolve(u0, t0=0.0, tf=1.0, store_freq=2)
    assert len(solver.t) == len(drv.t)
    np.testing.assert_allclose(solver.t, drv.t, rtol=1e-10)
    for a, b in zip(solver.u, drv.u):
        assert solvers.rel_l2(host(a), b) < 1e-11
    # step(): three steps with the returned suggestions
    s2 = rk.IF34(L, nl, config=rk.SolverConfig(epsilon=1e-6))
    d2 = solvers.make_solver("IF34", p["L"], p["nl"], solvers.SolverConfig(epsilon=1e-6))
    v, w, h, hw = dev(u0), u0, 0.1, 0.1
    for _ in range(3):
        v, hu, h = s2.step(v, h)
        w, hwu, hw = d2.step(w, hw)
        np.testing.assert_allclose([hu, h], [hwu, hw], rtol=1e-10)
        assert solvers.rel_l2(host(v), w) < 1e-11


@pytest.mark.parametrize("precision,bar", [("f64", 1e-11), ("mixed", 1e-6), ("f32", 1e-4)])
def test_ac2d_batch(rk, precision, bar):
    """A batch of trajectories in one launch set (per-trajectory dt) vs one-at-a-time oracle runs."""
    from rkstiff_amd.batch import BatchSolver
    n, B = 64, 5
    p = ac(n, seed=5, batch=B, kcut=12)
    L, nl = rk.allen_cahn_ops(n)
    bs = BatchSolver("ETD35", [(L, nl)] * B, precision=precision, config=rk.SolverConfig(epsilon=1e-5))
    u0 = [dev(u * 20.0) for u in p["u0"]]
    tfs = [1.0 + 0.5 * i for i in range(B)]
    uf = bs.evolve(u0, 0.0, tfs)
    assert not bs.failures()
    st = bs.stats()
    for i in range(B):
        drv = solvers.make_solver("ETD35", p["L"], p["nl"], solvers.SolverConfig(epsilon=1e-5))
        ref = drv.evolve(p["u0"][i] * 20.0, t0=0.0, tf=tfs[i], store_data=False)
        err = solvers.rel_l2(host(uf[i]), ref)
        print(precision, i, "rel", err, "attempts", st[i]["attempts"], len(drv.log))
        assert err < bar
        assert abs(st[i]["tc"] - tfs[i]) < 1e-12


def _advected(n, c=0.5):
    """Allen-Cahn lin_op plus advection along y, L = -eps |k|^2 + i c ky: L[r] != L[n - r]."""
    ky = np.fft.fftfreq(n, d=1.0 / n)[:, None] * np.ones((1, n // 2 + 1))
    return ky.reshape(-1) * 1j * c


@pytest.mark.parametrize("method", ["ETD35", "ETD34", "IF34"])
def test_ac2d_asymmetric_lin_op_matches_oracle(rk, method):
    """Rows r and n - r share coefficients only for a lin_op symmetric in ky; with advection along
    y every row keeps its own.  Engine vs oracle on 32 x 32 (same attempt sequence)."""
    p = ac(32, seed=3, kcut=8)
    u0 = p["u0"][0] * 20.0
    L0, nl = rk.allen_cahn_ops(32)
    L = L0.numpy() + _advected(32)
    cls = getattr(rk, method)
    solver = cls(torch.as_tensor(L), nl, config=rk.SolverConfig(epsilon=1e-6))
    solver.log_capacity = 8192
    uf = solver.evolve(dev(u0), t0=0.0, tf=1.0, store_data=False)
    drv = solvers.make_solver(method, L, p["nl"], solvers.SolverConfig(epsilon=1e-6))
    ref = drv.evolve(u0, t0=0.0, tf=1.0, store_data=False)
    assert len(solver.engine_log()) == len(drv.log)
    err = solvers.rel_l2(host(uf), ref)
    print(method, "asymmetric 2-D rel", err)
    assert err < 1e-11


@pytest.mark.parametrize("precision,bar", [("f64", 1e-11), ("mixed", 1e-6), ("f32", 1e-4)])
def test_ac2d_batch_symmetric_and_asymmetric(rk, precision, bar):
    """One launch set with symmetric (shared row coefficients) and advected (full) lin_ops."""
    from rkstiff_amd.batch import BatchSolver
    n = 64
    p = ac(n, seed=6, batch=3, kcut=12)
    L0, nl = rk.allen_cahn_ops(n)
    Ls = [L0.numpy().astype(np.complex128), L0.numpy() + _advected(n), L0.numpy().astype(np.complex128)]
    bs = BatchSolver("ETD35", [(torch.as_tensor(L), nl) for L in Ls], precision=precision,
                     config=rk.SolverConfig(epsilon=1e-5))
    dt = torch.complex64 if precision == "f32" else torch.complex128
    uf = bs.evolve([dev(u * 20.0, dt) for u in p["u0"]], 0.0, 1.0)
    assert not bs.failures()
    for i, L in enumerate(Ls):
        drv = solvers.make_solver("ETD35", L, p["nl"], solvers.SolverConfig(epsilon=1e-5))
        ref = drv.evolve(p["u0"][i] * 20.0, t0=0.0, tf=1.0, store_data=False)
        err = solvers.rel_l2(host(uf[i]), ref)
        print(precision, i, "rel", err)
        assert err < bar


def test_ac2d_rejects_bad_grids(rk):
    L = torch.zeros(48 * 25, dtype=torch.float64)
    with pytest.raises(ValueError):
        rk.ETD35(L, rk.AllenCahnNL(48)).evolve(dev(np.zeros(48 * 25)), 0.0, 1.0)
    with pytest.raises(ValueError):
        rk.AllenCahnNL(64, 32)


def test_ac2d_graph_replay_matches_direct_launches(rk, monkeypatch):
    """The captured attempt graph (replayed many times, with host syncs between
    groups of replays, over two evolve calls) gives bit-identical results to the
    same kernels launched directly."""
    from rkstiff_amd.batch import BatchSolver
    n, B = 64, 3
    p = ac(n, seed=9, batch=B, kcut=12)
    L, nl = rk.allen_cahn_ops(n)
    u0 = [dev(u * 20.0) for u in p["u0"]]
    outs = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("RKS_NO_GRAPH", flag)
        bs = BatchSolver("ETD35", [(L, nl)] * B, precision="f64", config=rk.SolverConfig(epsilon=1e-5))
        first = bs.evolve(u0, 0.0, 0.5)
        second = bs.evolve(first, 0.5, 1.5)
        outs[flag] = [host(u) for u in first + second]
    for a, b in zip(outs["0"], outs["1"]):
        np.testing.assert_array_equal(a, b)
