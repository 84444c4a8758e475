"""Benchmark: trajectories/s to tf on one GPU, batched stepping vs the plain FP64 drop-in.

Workload: the two-soliton cubic NLS, u_t = (i/2) u_xx + i |u|^2 u, n = 8192
complex fft grid on [-64, 64), L = -(i/2) kx^2, ETD35 adaptive with
epsilon = 1e-6, t0 = 0, tf = 20.  A batch of ``--batch`` trajectories
(instance 0 the unperturbed pair, the others seeded shifts of positions and
phases) is integrated by one launch of the persistent stepping kernel; a
"step" is that batch integrated t0 -> tf, inputs resident in HBM.

Precision modes (all with FP64 coefficient generation, error norm and dt
control on the device):
  f64    FP64 throughout (the parity path)
  mixed  FP64 state, propagators and FFTs; FP32 NL arrays and the other
         coefficients
  f32    FP32 state, stages and FFTs
The headline value is the fastest mode whose max rel L2 error at tf against
the FP64 engine is <= 1e-6 (the error is in the line).

Reported beside the value:
  * the plain per-instance FP64 drop-in (one ETD35.evolve per trajectory),
  * the HBM roofline of the stepping kernel: algorithmic bytes per launch
    (per-attempt byte model below) / launch time from HIP events recorded on
    the launch stream,
  * the CPU baseline: the NumPy FP64 oracle on a bounded sample of the same
    workload, one trajectory per host core.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
                       [--modes f64,mixed,f32] [--plain P] [--no-cpu]
                       [--copy-mib X]   (first run a device copy of X MiB:
                                         the reference point for HBM counters)
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "trajectories/s to tf at ≤1e-6 rel L2 err vs FP64 (batched vs plain); % HBM peak"
HBM_PEAK_GBS = 8000.0

N = 8192
TF = 20.0
EPS = 1e-6


def nls_initial(n, seed):
    """Two colliding solitons; instance 0 is the unperturbed configuration."""
    dx = 128.0 / n
    x = np.arange(-64.0, 64.0, dx)
    rng = np.random.default_rng(seed)
    s1, s2 = (0.0, 0.0) if seed == 0 else rng.uniform(-1.0, 1.0, 2)
    p1, p2 = (0.0, 0.0) if seed == 0 else rng.uniform(0.0, 2 * np.pi, 2)
    u0 = (2.0 / np.cosh(2.0 * (x + 10.0 + s1)) * np.exp(1j * (x + p1))
          + 1.5 / np.cosh(1.5 * (x - 10.0 + s2)) * np.exp(-0.5j * x + 1j * p2))
    return np.fft.fft(u0)


def nls_lin_op(n):
    kx = 2 * np.pi * np.fft.fftfreq(n, d=128.0 / n)
    return -0.5j * kx**2


def bytes_per_attempt(m, e_u, e_nl, e_L=8):
    """Algorithmic HBM bytes of one ETD35 attempt (etd35.py:320-344).

    Per mode: the six stage combinations read u six times (e_u each) and 19 NL
    arrays (e_nl), write five NL arrays and the candidate, and read lin_op once
    per stage (e_L: one FP64 real per mode for the pure-imaginary NLS lin_op).
    Uniform precision reduces to M [(R_s + W) e_c + S e_L] with R_s=25, W=6, S=6.
    """
    return m * (6 * e_u + 19 * e_nl + 5 * e_nl + 1 * e_u + 6 * e_L)


def bytes_per_accept(m, e_u, e_nl):
    """NL1 = nl(u) after an accepted attempt (etd35.py:317-318): read u, write NL1."""
    return m * (e_u + e_nl)


# HBM bytes per trajectory of this workload measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# (scaled by a copy of known size; profiles/rks_kernel_nls8192_etd35_b256_hbm_counters.txt)
TRAFFIC_PER_TRAJECTORY = {"f64": 1.445e10, "mixed": 7.58e9}

MODES = {  # name: (precision id, state element bytes, NL element bytes)
    "f32": ("F32", 8, 8),
    "mixed": ("MIXED", 16, 8),
    "f64": ("F64", 16, 16),
}


# ---------------------------------------------------------------------------
# CPU baseline: the NumPy FP64 oracle (test infrastructure) on host cores
# ---------------------------------------------------------------------------
def _cpu_one(seed):
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import problems, solvers
    L = nls_lin_op(N)
    drv = solvers.make_solver("ETD35", L, problems.nls_nl(1.0), solvers.SolverConfig(epsilon=EPS))
    t0 = time.perf_counter()
    drv.evolve(nls_initial(N, seed), t0=0.0, tf=TF, store_data=False)
    return time.perf_counter() - t0


def cpu_baseline(workers):
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[k] = "1"
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(workers) as pool:
        per = pool.map(_cpu_one, list(range(workers)))
    wall = time.perf_counter() - t0
    return {"value": workers / wall, "unit": "trajectories/s", "cores": workers, "kind": "port",
            "sample": f"{workers} trajectories of the workload (one per core, NumPy FP64 oracle), "
                      f"wall {wall:.1f} s, single-trajectory {np.median(per):.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--plain", type=int, default=2, help="trajectories for the plain FP64 figure")
    ap.add_argument("--cpu-workers", type=int, default=0, help="0: min(16, cpu_count)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="f64,mixed,f32", help="precision modes to run (f64 always runs)")
    ap.add_argument("--copy-mib", type=int, default=0, help="device copy of this size first (counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; trajectories are independent, "
                                   "so N GPUs would each run their own share (see DESIGN.md)"}))
        sys.exit(2)

    import torch
    from rkstiff_amd import _lib
    from rkstiff_amd.engine import Plan
    from rkstiff_amd.config import SolverConfig
    from rkstiff_amd.nonlinear import CubicNL
    from rkstiff_amd.solvers import ETD35

    torch.cuda.init()
    dev = torch.device("cuda:0")
    B = args.batch
    L = nls_lin_op(N)
    nl = CubicNL(N, 1.0)
    cfg = SolverConfig(epsilon=EPS)
    u0_host = [nls_initial(N, i) for i in range(B)]
    t0s, tfs, hs = [0.0] * B, [TF] * B, [TF / 100.0] * B
    stream = torch.cuda.current_stream()

    def run_mode(name):
        prec, e_u, e_nl = MODES[name]
        dt = torch.complex64 if name == "f32" else torch.complex128
        plan = Plan("ETD35", nl, getattr(_lib, prec), cfg, None, [nl.descriptor(L) for _ in range(B)])
        u0 = [torch.as_tensor(u, dtype=dt, device=dev) for u in u0_host]
        uf = [torch.empty_like(u) for u in u0]
        for _ in range(args.warmup):
            plan.evolve(u0, uf, t0s, tfs, hs)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        kernel_ms = []
        t_start = time.perf_counter()
        for _ in range(args.steps):
            ev0.record(stream)
            plan.evolve(u0, uf, t0s, tfs, hs)
            ev1.record(stream)
            ev1.synchronize()
            kernel_ms.append(ev0.elapsed_time(ev1))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_start
        st = plan.stats()
        attempts = np.array([x["attempts"] for x in st])
        steps = np.array([x["steps"] for x in st])
        alg = float(attempts.sum()) * bytes_per_attempt(N, e_u, e_nl) + \
            float(steps.sum()) * bytes_per_accept(N, e_u, e_nl)
        k_ms = float(np.mean(kernel_ms))
        out = {"value": B * args.steps / wall, "ms_per_step": wall / args.steps * 1e3,
               "kernel_ms": k_ms, "achieved_GBs": alg / (k_ms * 1e-3) / 1e9,
               "attempts": float(attempts.mean()), "ok": all(x["status"] == 0 for x in st)}
        del plan
        return out, uf

    if args.copy_mib > 0:
        src = torch.empty(args.copy_mib * 2**18, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4}),
              flush=True)
        del src, dst
    modes = ["f64"] + [m for m in args.modes.split(",") if m in MODES and m != "f64"]
    results = {}
    finals = {}
    for name in modes:
        results[name], finals[name] = run_mode(name)
    ref = finals["f64"]
    for name in modes[1:]:
        errs = np.array([float(torch.linalg.norm(finals[name][i].to(torch.complex128) - ref[i])
                               / torch.linalg.norm(ref[i])) for i in range(B)])
        results[name]["rel_l2_max"] = float(errs.max())
        results[name]["rel_l2_median"] = float(np.median(errs))
    results["f64"]["rel_l2_max"] = 0.0
    results["f64"]["rel_l2_median"] = 0.0
    eligible = [k for k in results if results[k]["rel_l2_max"] <= 1e-6 and results[k]["ok"]]
    head = max(eligible, key=lambda k: results[k]["value"])
    hr = results[head]

    plain_s = None
    if args.plain > 0:
        plain = ETD35(torch.as_tensor(L), nl, config=cfg)
        u064 = [torch.as_tensor(u, dtype=torch.complex128, device=dev) for u in u0_host[:args.plain]]
        plain.evolve(u064[0], 0.0, 1.0, store_data=False)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for i in range(args.plain):
            plain.evolve(u064[i], 0.0, TF, store_data=False)
        torch.cuda.synchronize()
        plain_s = (time.perf_counter() - tp) / args.plain

    line = {
        "metric": METRIC,
        "value": round(hr["value"], 3),
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(hr["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "f32" else "f64",
        "data": "synthetic",
        "config": {"workload": "nls2soliton_n8192_etd35_tf20_eps1e-6_batched", "method": "ETD35",
                   "n": N, "tf": TF, "epsilon": EPS, "batch": B, "precision_mode": head,
                   "precision": {"f64": "f64 throughout", "mixed": "f64 state, propagators and FFTs; "
                                 "f32 NL arrays and other coefficients; f64 norm and dt control",
                                 "f32": "f32 stages and FFTs; f64 coefficients, norm, dt control"}[head],
                   "parallelism": "1 GPU, one workgroup per trajectory, persistent kernel"},
        "roofline": {"bound": "hbm", "achieved": round(hr["achieved_GBs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(hr["achieved_GBs"] / HBM_PEAK_GBS, 4),
                     "traffic": (None if head not in TRAFFIC_PER_TRAJECTORY
                                 else float(f"{TRAFFIC_PER_TRAJECTORY[head] * B:.4g}"))},
        "accuracy": {"rel_l2_vs_fp64_max": hr["rel_l2_max"], "rel_l2_vs_fp64_median": hr["rel_l2_median"],
                     "gate": 1e-6},
        "modes": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                  for k, v in results.items()},
        "kernel_ms_per_launch": round(hr["kernel_ms"], 3),
        "plain_fp64": None if plain_s is None else {"value": round(1.0 / plain_s, 4), "unit": "trajectories/s",
                                                    "s_per_trajectory": round(plain_s, 4)},
        "batched_over_plain": None if plain_s is None else round(hr["value"] * plain_s, 2),
    }
    if not args.no_cpu:
        workers = args.cpu_workers or min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(workers)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
