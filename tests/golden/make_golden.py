"""Write the golden fixtures of tests/golden/ from the reference implementation.

Run once, in the CPU container where the reference tree is mounted:

    python tests/golden/make_golden.py [--reference /root/reference]

The reference package's ``__init__`` imports a version module generated at
install time (absent from the tree), so the package object is created from a
ModuleSpec pointing at the reference directory and its submodules are imported
from there unchanged.  No bytecode is written next to the reference.

Only data is written: inputs, outputs and attempt logs (h, s) as .npz, and a
timing calibration as JSON.  The fixtures pin what closed forms cannot: the
exact adaptive trajectories and step sequences, the psi coefficients at the
rounding level near ``modecutoff``, and single attempts of every method.
"""
import argparse
import importlib.machinery
import importlib.util
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.dont_write_bytecode = True


def load_reference(root):
    spec = importlib.machinery.ModuleSpec("rkstiff", None, is_package=True)
    pkg = importlib.util.module_from_spec(spec)
    pkg.__path__ = [os.path.join(root, "rkstiff")]
    sys.modules["rkstiff"] = pkg
    import rkstiff.etd35, rkstiff.etd34, rkstiff.etd4, rkstiff.etd5  # noqa: E401
    import rkstiff.if34, rkstiff.if45dp, rkstiff.if4, rkstiff.grids, rkstiff.models  # noqa: E401
    import rkstiff.transforms  # noqa: F401
    return rkstiff


def logged(cls):
    """Subclass recording (h, s) of every attempt made by the adaptive controller."""
    class Logged(cls):
        def __init__(self, *a, **k):
            super().__init__(*a, **k)
            self.attempt_log = []
            self._cur_h = None

        def _update_stages(self, u, h):
            self._cur_h = h
            return super()._update_stages(u, h)

        def _compute_s(self, u, err):
            s = super()._compute_s(u, err)
            self.attempt_log.append((self._cur_h, float(s)))
            return s
    Logged.__name__ = cls.__name__
    return Logged


def ks_nl(kx):
    def nl_func(u_fft):
        u = np.fft.irfft(u_fft)
        ux = np.fft.irfft(1j * kx * u_fft)
        return -np.fft.rfft(u * ux)
    return nl_func


def nls_nl(gamma=1.0):
    def nl_func(g):
        F = np.fft.ifft(g)
        F2 = F.real**2 + F.imag**2
        return 1j * gamma * np.fft.fft(F2 * F)
    return nl_func


def log_arrays(solver):
    lg = np.array(solver.attempt_log, dtype=np.float64).reshape(-1, 2)
    return lg[:, 0], lg[:, 1]


def diag_problems():
    """Two small non-diagonal lin_op matrices (real and complex eigenvalues) with a cubic nl."""
    rng = np.random.default_rng(11)
    n = 12
    S0 = np.eye(n) + 0.25 * rng.standard_normal((n, n))
    L_real = S0 @ np.diag(-np.linspace(0.5, 6.0, n)) @ np.linalg.inv(S0)
    u_real = rng.standard_normal(n) + 0j
    m = 10
    M = rng.standard_normal((m, m))
    A = rng.standard_normal((m, m))
    L_cplx = -(0.1 * M @ M.T + 0.5 * np.eye(m)) + 0.8 * (A - A.T)
    u_cplx = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    return {"real": (L_real, u_real, 0.5), "cplx": (L_cplx, u_cplx, 0.2)}


def cubic_nl(a):
    def nl_func(u):
        return a * u - u * u * u
    return nl_func


def case_diagonalized(rk):
    """G: diagonalized ETD35 / ETD34 / IF34 (etd35.py:348-495, etd34.py:194-304, if34.py:137-196)."""
    from rkstiff.solveras import SolverConfig
    classes = {"ETD35": rk.etd35.ETD35, "ETD34": rk.etd34.ETD34, "IF34": rk.if34.IF34}
    out = {}
    for pname, (L, u0, a) in diag_problems().items():
        out[f"{pname}/L"], out[f"{pname}/u0"], out[f"{pname}/a"] = L, u0, a
        for mname, cls in classes.items():
            sv = logged(cls)(lin_op=L, nl_func=cubic_nl(a), config=SolverConfig(epsilon=1e-6),
                             diagonalize=True)
            uf = sv.evolve(u0, t0=0.0, tf=3.0, store_data=True, store_freq=5)
            lh, ls = log_arrays(sv)
            out[f"{pname}/{mname}/u_final"] = uf
            out[f"{pname}/{mname}/log_h"], out[f"{pname}/{mname}/log_s"] = lh, ls
            out[f"{pname}/{mname}/t"] = np.array(sv.t)
            sv = cls(lin_op=L, nl_func=cubic_nl(a), config=SolverConfig(epsilon=1e-4), diagonalize=True)
            v, h = u0, 0.2
            for i in range(3):
                v, hu, h = sv.step(v, h)
                out[f"{pname}/{mname}/u{i + 1}"] = v
                out[f"{pname}/{mname}/hused{i + 1}"] = hu
                out[f"{pname}/{mname}/hnext{i + 1}"] = h
    np.savez_compressed(os.path.join(HERE, "diagonalized.npz"), **out)
    print("G diagonalized done", flush=True)


def dense_problems():
    """Small non-diagonal lin_op matrices whose spectrum times the step sizes stays inside the
    unit contour of the dense psi evaluation (real and complex eigenvalues)."""
    rng = np.random.default_rng(23)
    n = 8
    S0 = np.eye(n) + 0.3 * rng.standard_normal((n, n))
    L_real = S0 @ np.diag(-np.linspace(0.05, 0.8, n)) @ np.linalg.inv(S0)
    u_real = rng.standard_normal(n) + 0j
    A = rng.standard_normal((n, n))
    L_cplx = -0.2 * np.eye(n) + 0.15 * (A - A.T) - 0.05 * (A @ A.T) / n
    u_cplx = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return {"real": (L_real, u_real, 0.5), "cplx": (L_cplx, u_cplx, 0.3)}


def case_dense(rk):
    """H: dense strategies (etd35.py:498-720, etd34.py:307-436, etd4.py:178-308, etd5.py:264-482,
    if34.py:202-254, if4.py:125-199)."""
    from rkstiff.solveras import SolverConfig
    adaptive = {"ETD35": rk.etd35.ETD35, "ETD34": rk.etd34.ETD34, "IF34": rk.if34.IF34}
    constant = {"ETD4": rk.etd4.ETD4, "ETD5": rk.etd5.ETD5, "IF4": rk.if4.IF4}
    out = {}
    for pname, (L, u0, a) in dense_problems().items():
        out[f"{pname}/L"], out[f"{pname}/u0"], out[f"{pname}/a"] = L, u0, a
        for mname, cls in adaptive.items():
            sv = logged(cls)(lin_op=L, nl_func=cubic_nl(a), config=SolverConfig(epsilon=1e-6))
            uf = sv.evolve(u0, t0=0.0, tf=2.0, store_data=True, store_freq=4)
            lh, ls = log_arrays(sv)
            out[f"{pname}/{mname}/u_final"] = uf
            out[f"{pname}/{mname}/log_h"], out[f"{pname}/{mname}/log_s"] = lh, ls
            out[f"{pname}/{mname}/t"] = np.array(sv.t)
        for mname, cls in constant.items():
            sv = cls(lin_op=L, nl_func=cubic_nl(a))
            out[f"{pname}/{mname}/u_final"] = sv.evolve(u0, t0=0.0, tf=1.0, h=0.025, store_data=False)
    np.savez_compressed(os.path.join(HERE, "dense.npz"), **out)
    print("H dense done", flush=True)


def case_cheb_allen_cahn(rk):
    """I: demos/allen_cahn.ipynb cells 4-10 -- Chebyshev N = 20 Allen-Cahn, ETD34/ETD35/IF34 dense
    (64 contour points, radius 20) and diagonalized (32 points, radius 1), epsilon 1e-4.  The demo
    runs to tf = 100, where a 1e-14 change of w0 already moves the FP64 result by ~3e-4 (the
    fronts' metastable dynamics); the fixture stops at tf = 10, where that floor is ~1e-13."""
    from rkstiff.solveras import SolverConfig
    from rkstiff.etd import ETDConfig
    x, D = rk.grids.construct_x_dx_cheb(20, -1, 1)
    L, nl = rk.models.allen_cahn_ops(x, D, 0.01)
    u0 = 0.53 * x + 0.47 * np.sin(-1.5 * np.pi * x)
    w0 = (u0 - x)[1:-1].astype(np.complex128)
    out = {"x": x, "D": D, "L": L, "w0": w0}
    classes = {"ETD35": rk.etd35.ETD35, "ETD34": rk.etd34.ETD34, "IF34": rk.if34.IF34}
    for strat, etd in (("dense", ETDConfig(contour_points=64, contour_radius=20)),
                       ("diag", ETDConfig(contour_points=32, contour_radius=1))):
        for mname, cls in classes.items():
            kw = {} if mname == "IF34" else {"etd_config": etd}
            sv = logged(cls)(lin_op=L, nl_func=nl, config=SolverConfig(epsilon=1e-4),
                             diagonalize=(strat == "diag"), **kw)
            uf = sv.evolve(w0, t0=0.0, tf=10.0, store_data=True, store_freq=10)
            lh, ls = log_arrays(sv)
            out[f"{strat}/{mname}/u_final"] = uf
            out[f"{strat}/{mname}/log_h"], out[f"{strat}/{mname}/log_s"] = lh, ls
            out[f"{strat}/{mname}/t"] = np.array(sv.t)
    np.savez_compressed(os.path.join(HERE, "cheb_allen_cahn.npz"), **out)
    print("I chebyshev allen-cahn done", flush=True)


def case_hankel(rk):
    """J: transforms.py HankelTransform (grid, zeros, matrix, ht of a Gaussian) and the radial
    NLS of demos/nls.ipynb cells 6-12 (nr = 100, rmax = 5, L = -i kr^2, gamma = 2,
    u0 = 2 exp(-r^2)) with ETD35 / ETD34 / IF34, epsilon 1e-5, tf = 0.8, store_freq = 10."""
    from rkstiff.solveras import SolverConfig
    out = {}
    for nr, rmax in ((12, 2.0), (50, 2.0), (100, 5.0)):
        ht = rk.transforms.HankelTransform(nr=nr, rmax=rmax)
        key = f"{nr}_{rmax:g}"
        out[f"{key}/r"], out[f"{key}/kr"] = ht.r, ht.kr
        out[f"{key}/zeros"], out[f"{key}/Y"] = ht.bessel_zeros(), ht.hankel_matrix()
        f = np.exp(-16.0 * ht.r**2) * (1.0 + 0.5j * ht.r)
        out[f"{key}/f"], out[f"{key}/ht_f"], out[f"{key}/iht_f"] = f, ht.ht(f), ht.iht(f)
    HT = rk.transforms.HankelTransform(nr=100, rmax=5)
    L = -1j * HT.kr**2
    gamma = 2

    def nl_func(g):
        f = HT.iht(g)
        f2 = f.real**2 + f.imag**2
        return 1j * gamma * HT.ht(f2 * f)

    u0HT = HT.ht(2.0 * np.exp(-HT.r**2))
    out["nls/L"], out["nls/u0"] = L, u0HT
    classes = {"ETD35": rk.etd35.ETD35, "ETD34": rk.etd34.ETD34, "IF34": rk.if34.IF34}
    for mname, cls in classes.items():
        sv = logged(cls)(lin_op=L, nl_func=nl_func, config=SolverConfig(epsilon=1e-5))
        uf = sv.evolve(u0HT, t0=0.0, tf=0.8, store_data=True, store_freq=10)
        lh, ls = log_arrays(sv)
        out[f"nls/{mname}/u_final"], out[f"nls/{mname}/t"] = uf, np.array(sv.t)
        out[f"nls/{mname}/log_h"], out[f"nls/{mname}/log_s"] = lh, ls
        out[f"nls/{mname}/nl_u0"] = nl_func(u0HT)
    np.savez_compressed(os.path.join(HERE, "hankel.npz"), **out)
    print("J hankel done", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--skip-nls8192", action="store_true")
    ap.add_argument("--only", default="", help="comma list of cases to write (default: all)")
    args = ap.parse_args()
    rk = load_reference(args.reference)
    only = set(filter(None, args.only.split(",")))
    if only:
        if "G" in only:
            case_diagonalized(rk)
        if "H" in only:
            case_dense(rk)
        if "I" in only:
            case_cheb_allen_cahn(rk)
        if "J" in only:
            case_hankel(rk)
        return
    from rkstiff.solveras import SolverConfig
    from rkstiff.etd import ETDConfig
    grids, models = rk.grids, rk.models
    classes = {
        "ETD35": rk.etd35.ETD35, "ETD34": rk.etd34.ETD34, "IF34": rk.if34.IF34,
        "IF45DP": rk.if45dp.IF45DP, "ETD4": rk.etd4.ETD4, "ETD5": rk.etd5.ETD5,
        "IF4": rk.if4.IF4,
    }
    adaptive = ("ETD35", "ETD34", "IF34", "IF45DP")
    calib = {}

    # -- A: KS quickstart, IF34, n=1024, tf=50, store_freq=20 ------------------
    n = 1024
    x, kx = grids.construct_x_kx_rfft(n, 0.0, 32.0 * np.pi)
    L = kx**2 * (1 - kx**2)
    u0 = np.fft.rfft(np.cos(x / 16) * (1.0 + np.sin(x / 16)))
    s = logged(classes["IF34"])(lin_op=L, nl_func=ks_nl(kx))
    t0 = time.perf_counter()
    uf = s.evolve(u0, t0=0.0, tf=50.0, store_freq=20)
    calib["ks1024_if34_reference_s"] = time.perf_counter() - t0
    lh, ls = log_arrays(s)
    np.savez_compressed(os.path.join(HERE, "ks1024_if34.npz"), L=L, kx=kx, u0=u0, u_final=uf,
                        t=np.array(s.t), snap_norm=np.array([np.linalg.norm(v) for v in s.u]),
                        log_h=lh, log_s=ls)
    calib["ks1024_if34_attempts"] = int(lh.size)
    print("A ks1024 if34: attempts", lh.size, "stored", len(s.t),
          "%.2fs" % calib["ks1024_if34_reference_s"], flush=True)

    # -- B: two-soliton NLS, ETD35, n=8192, tf=20, eps=1e-6 ------------------
    if not args.skip_nls8192:
        n = 8192
        x, kx = grids.construct_x_kx_fft(n, -64.0, 64.0)
        L = -0.5j * kx**2
        u0 = np.fft.fft(2.0 / np.cosh(2.0 * (x + 10.0)) * np.exp(1j * x)
                        + 1.5 / np.cosh(1.5 * (x - 10.0)) * np.exp(-0.5j * x))
        s = logged(classes["ETD35"])(lin_op=L, nl_func=nls_nl(1.0),
                                     config=SolverConfig(epsilon=1e-6))
        t0 = time.perf_counter()
        uf = s.evolve(u0, t0=0.0, tf=20.0, store_data=False)
        calib["nls8192_etd35_reference_s"] = time.perf_counter() - t0
        lh, ls = log_arrays(s)
        np.savez_compressed(os.path.join(HERE, "nls8192_etd35.npz"), L=L, u0=u0, u_final=uf,
                            log_h=lh, log_s=ls)
        calib["nls8192_etd35_attempts"] = int(lh.size)
        print("B nls8192 etd35: attempts", lh.size,
              "%.2fs" % calib["nls8192_etd35_reference_s"], flush=True)

    # -- C: coefficients near and away from modecutoff -----------------------
    base = np.concatenate([
        [0.0, 1e-9, -1e-9, 1e-4, -1e-4],
        np.linspace(-0.0102, 0.0102, 21),
        -np.logspace(-3, 3, 25), np.logspace(-3, 1, 9)])
    Ls = {"real": base, "imag": 1j * base, "cplx": base * (0.6 - 0.8j)}
    hs = (0.5, 0.9999, 0.013)
    coeffs = {}
    etdc = ETDConfig()
    strat = {
        "ETD35": lambda L: rk.etd35._Etd35Diagonal(L, None, etdc),
        "ETD34": lambda L: rk.etd34._Etd34Diagonal(L, None, etdc),
        "ETD4": lambda L: rk.etd4._Etd4Diagonal(L, None, etdc),
        "ETD5": lambda L: rk.etd5._Etd5Diagonal(L, None, etdc),
        "IF34": lambda L: rk.if34._If34Diagonal(L, None),
        "IF4": lambda L: rk.if4._IF4Diagonal(L, None),
    }
    for lname, L in Ls.items():
        coeffs[f"L/{lname}"] = L
        for hi, h in enumerate(hs):
            for mname, mk in strat.items():
                obj = mk(L)
                obj.update_coeffs(h)
                for attr, val in vars(obj).items():
                    if attr.startswith("_") and attr[1:2] in ("E", "a", "b") and \
                            isinstance(val, np.ndarray):
                        coeffs[f"{mname}/{lname}/{hi}/{attr[1:]}"] = val
            obj = classes["IF45DP"](lin_op=L, nl_func=lambda u: u)
            obj._update_coeffs(h)
            for attr, val in vars(obj).items():
                if attr.startswith("_") and attr[1:2] in ("E", "a", "r") and \
                        isinstance(val, (np.ndarray, float)):
                    coeffs[f"IF45DP/{lname}/{hi}/{attr[1:]}"] = np.asarray(val)
    coeffs["h"] = np.array(hs)
    np.savez_compressed(os.path.join(HERE, "coeffs.npz"), **coeffs)
    print("C coeffs:", len(coeffs), "arrays", flush=True)

    # -- D: three steps of every method on small KS (rfft) and NLS (fft) -----
    att = {}
    xs, kxs = grids.construct_x_kx_rfft(64, 0.0, 32.0 * np.pi)
    Lks = kxs**2 * (1 - kxs**2)
    uks = np.fft.rfft(np.cos(xs / 16) * (1.0 + np.sin(xs / 16)) + 0.1 * np.sin(3 * xs / 16))
    xn, kxn = grids.construct_x_kx_fft(64, -16.0, 16.0)
    Lnls = -0.5j * kxn**2
    unls = np.fft.fft(1.2 / np.cosh(1.2 * xn) * np.exp(0.5j * xn))
    probs = {"ks64": (Lks, ks_nl(kxs), uks, 0.05), "nls64": (Lnls, nls_nl(1.0), unls, 0.02)}
    for pname, (L, nl, u, h0) in probs.items():
        att[f"{pname}/L"] = L
        att[f"{pname}/u0"] = u
        for mname, cls in classes.items():
            if mname in adaptive:
                sv = logged(cls)(lin_op=L, nl_func=nl, config=SolverConfig(epsilon=1e-3))
                v, h = u.copy(), h0
                for i in range(3):
                    v, hu, h = sv.step(v, h)
                    att[f"{pname}/{mname}/u{i + 1}"] = v
                    att[f"{pname}/{mname}/hused{i + 1}"] = np.float64(hu)
                    att[f"{pname}/{mname}/hnext{i + 1}"] = np.float64(h)
                lh, ls = log_arrays(sv)
                att[f"{pname}/{mname}/log_h"] = lh
                att[f"{pname}/{mname}/log_s"] = ls
            else:
                sv = cls(lin_op=L, nl_func=nl)
                v = u.copy()
                for i in range(3):
                    v = sv.step(v, h0)
                    att[f"{pname}/{mname}/u{i + 1}"] = v
    np.savez_compressed(os.path.join(HERE, "attempts.npz"), **att)
    print("D attempts:", len(att), "arrays", flush=True)

    # -- E: KdV soliton fixture, every method; Burgers IF34/IF45DP ----------
    kd = {}
    x, kx = grids.construct_x_kx_rfft(256, -30, 30)
    h, steps = 0.025, 200
    tf = h * steps
    u0 = np.fft.rfft(models.kdv_soliton(x, ampl=1.0, x0=-5.0, t=0))
    L, nl = models.kdv_ops(kx)
    kd["u0"], kd["L"] = u0, L
    for mname, cls in classes.items():
        if mname in adaptive:
            sv = logged(cls)(lin_op=L, nl_func=nl, config=SolverConfig(epsilon=1e-4))
            uf = sv.evolve(u0, t0=0.0, tf=tf, h_init=h, store_data=True, store_freq=10)
            lh, ls = log_arrays(sv)
            kd[f"{mname}/log_h"], kd[f"{mname}/log_s"] = lh, ls
        else:
            sv = cls(lin_op=L, nl_func=nl)
            uf = sv.evolve(u0, t0=0.0, tf=tf, h=h, store_data=True, store_freq=10)
        kd[f"{mname}/u_final"] = uf
        kd[f"{mname}/t"] = np.array(sv.t)
    xb, kxb = grids.construct_x_kx_rfft(1024, -np.pi, np.pi)
    Lb, nlb = models.burgers_ops(kxb, 0.0005)
    ub0 = np.fft.rfft(np.exp(-10 * np.sin(xb / 2) ** 2))
    kd["burgers/u0"], kd["burgers/L"] = ub0, Lb
    for mname, eps in (("IF34", 1e-4), ("IF45DP", 1e-3)):
        sv = logged(classes[mname])(lin_op=Lb, nl_func=nlb, config=SolverConfig(epsilon=eps))
        uf = sv.evolve(ub0, t0=0, tf=0.85, store_data=False)
        lh, ls = log_arrays(sv)
        kd[f"burgers/{mname}/u_final"] = uf
        kd[f"burgers/{mname}/log_h"], kd[f"burgers/{mname}/log_s"] = lh, ls
    np.savez_compressed(os.path.join(HERE, "kdv_burgers.npz"), **kd)
    print("E kdv/burgers done", flush=True)

    # -- F: NLS n=256 (complex fft path), adaptive methods, tf=2 ------------
    nl256 = {}
    x, kx = grids.construct_x_kx_fft(256, -32.0, 32.0)
    L = -0.5j * kx**2
    u0 = np.fft.fft(1.5 / np.cosh(1.5 * (x + 4)) * np.exp(1j * x)
                    + 1.0 / np.cosh(x - 4) * np.exp(-1j * x))
    nl256["L"], nl256["u0"] = L, u0
    for mname in adaptive:
        sv = logged(classes[mname])(lin_op=L, nl_func=nls_nl(1.0),
                                    config=SolverConfig(epsilon=1e-5))
        uf = sv.evolve(u0, t0=0.0, tf=2.0, store_data=False)
        lh, ls = log_arrays(sv)
        nl256[f"{mname}/u_final"] = uf
        nl256[f"{mname}/log_h"], nl256[f"{mname}/log_s"] = lh, ls
    for mname in ("ETD4", "ETD5", "IF4"):
        sv = classes[mname](lin_op=L, nl_func=nls_nl(1.0))
        nl256[f"{mname}/u_final"] = sv.evolve(u0, t0=0.0, tf=2.0, h=0.01, store_data=False)
    np.savez_compressed(os.path.join(HERE, "nls256.npz"), **nl256)
    print("F nls256 done", flush=True)

    case_diagonalized(rk)
    case_dense(rk)
    case_cheb_allen_cahn(rk)
    case_hankel(rk)

    calib["cpu"] = os.uname().machine
    calib["numpy"] = np.__version__
    calib["threads"] = os.environ.get("OMP_NUM_THREADS", "unset")
    old = {}
    path = os.path.join(HERE, "calibration.json")
    if os.path.exists(path):
        old = json.load(open(path))
    old.update(calib)
    json.dump(old, open(path, "w"), indent=1, sort_keys=True)
    print(json.dumps(calib, indent=1))


if __name__ == "__main__":
    main()
