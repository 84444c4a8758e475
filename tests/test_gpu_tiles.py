"""The second FFT tile (twice the threads per trajectory, n = 512 .. 4096) against
the reference's golden run and the oracle, like the default tile.

A tile changes the FFTs' pass structure and the reductions' summation order,
so results agree with the default tile to rounding, not bit for bit; the
launch shapes (bucket concurrency, resident workgroups) are bit-identical
(tests/test_gpu_proposer.py)."""
import numpy as np
import pytest

from oracle import problems

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def mods():
    from rkstiff_amd import _lib
    from rkstiff_amd.config import SolverConfig
    from rkstiff_amd.engine import Plan
    from rkstiff_amd.nonlinear import AdvectiveNL, CubicNL
    return _lib, SolverConfig, Plan, AdvectiveNL, CubicNL


def test_ks1024_if34_tile1_matches_reference(mods, golden):
    _lib, SolverConfig, Plan, AdvectiveNL, _ = mods
    g = golden("ks1024_if34")
    p = problems.ks_readme()
    nl = AdvectiveNL(torch.as_tensor(p["kx"]), -1.0)
    plan = Plan("IF34", nl, _lib.F64, SolverConfig(), None, [nl.descriptor(torch.as_tensor(g["L"]))])
    assert plan.buckets()[0][3] == 2
    plan.set_launch(True, None, [1])
    plan.set_storage(20, 64, 4096)
    u0 = torch.as_tensor(g["u0"]).cuda()
    uf = torch.empty_like(u0)
    plan.evolve([u0], [uf], [0.0], [50.0], [0.5])
    st = plan.stats()[0]
    log = plan.attempt_log(0, st["log_count"])
    times, _ = plan.snapshots(0, st["snap_count"])
    err = rel(uf.cpu().numpy(), g["u_final"])
    print("tile 1 attempts", len(log), "golden", len(g["log_h"]), "rel", err)
    assert st["status"] == _lib.RKS_OK
    assert len(log) == len(g["log_h"])
    np.testing.assert_allclose(log[:, 0], g["log_h"], rtol=1e-9, atol=1e-12)
    assert err < 1e-11
    np.testing.assert_allclose(times, g["t"], rtol=1e-12)


@pytest.mark.parametrize("kind", ["advect", "cubic"])
@pytest.mark.parametrize("n", [512, 1024, 2048, 4096])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_tile1_nl_matches_oracle(mods, kind, n, prec):
    _lib, SolverConfig, Plan, AdvectiveNL, CubicNL = mods
    rng = np.random.default_rng(n)
    if kind == "advect":
        kx = 2 * np.pi * np.fft.rfftfreq(n, d=1.0 / n) / 7.0
        nl = AdvectiveNL(torch.as_tensor(kx), -1.0)
        m = n // 2 + 1
        u = (rng.standard_normal(m) + 1j * rng.standard_normal(m)) * np.exp(-np.arange(m) / 30.0)
        u[0] = u[0].real
        u[-1] = u[-1].real
        ref = problems.rfft_advect_nl(kx, -1.0)(u)
    else:
        nl = CubicNL(n, 1.0)
        u = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * np.exp(-np.minimum(np.arange(n), n - np.arange(n)) / 30.0)
        ref = problems.nls_nl(1.0)(u)
    P = _lib.F64 if prec == "f64" else _lib.F32
    dt = torch.complex128 if prec == "f64" else torch.complex64
    plan = Plan("IF34", nl, P, None, None, [nl.descriptor(torch.zeros(len(u), dtype=torch.float64))])
    ut = torch.as_tensor(u).to(dt).cuda()
    outs = []
    for tile in (0, 1):
        plan.set_launch(True, None, [tile])
        o = torch.empty_like(ut)
        plan.nl_eval([ut], [o])
        outs.append(o.cpu().numpy().astype(np.complex128))
    tol = 1e-13 if prec == "f64" else 2e-6
    e0, e1 = rel(outs[0], ref), rel(outs[1], ref)
    print(kind, n, prec, "tile0", e0, "tile1", e1)
    assert e0 < tol and e1 < tol


def test_tile1_evolve_agrees_with_default_tile(mods):
    _lib, SolverConfig, Plan, _, CubicNL = mods
    from rkstiff_amd.workloads import nls_ensemble
    ens = nls_ensemble(6, seed=5, sizes=(1024, 2048, 4096))
    descs = [nl.descriptor(torch.as_tensor(L)) for (_, _, L, nl, _) in ens]
    for prec, tol in ((_lib.F64, 1e-10), (_lib.F32, 5e-4)):
        plan = Plan("ETD35", ens[0][3], prec, SolverConfig(epsilon=1e-6), None, descs)
        dt = plan.dtype
        u0 = [torch.as_tensor(e[4]).to(dt).cuda() for e in ens]
        res = []
        for tile in (0, 1):
            nb = len(plan.buckets())
            plan.set_launch(True, None, [tile] * nb)
            uf = [torch.empty_like(u) for u in u0]
            plan.evolve(u0, uf, [0.0] * len(ens), [2.0] * len(ens), [0.02] * len(ens))
            st = plan.stats()
            assert all(s["status"] == _lib.RKS_OK for s in st)
            res.append(([u.cpu().numpy().astype(np.complex128) for u in uf], [s["attempts"] for s in st]))
        errs = [rel(a, b) for a, b in zip(res[1][0], res[0][0])]
        print(prec, "tile1 vs tile0", max(errs), res[0][1], res[1][1])
        assert max(errs) < tol
