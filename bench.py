"""Benchmark: trajectories/s to tf for the batched FP32 stepping path on one GPU.

Workload (default ``nls``): the two-soliton cubic NLS of the reference's NLS
demo form, u_t = (i/2) u_xx + i |u|^2 u, n = 8192 complex fft grid on
[-64, 64), L = -(i/2) kx^2, ETD35 adaptive with epsilon = 1e-6, t0 = 0, tf = 20.
A batch of ``--batch`` trajectories (seeded perturbations of the soliton
positions and phases) is integrated per step by one launch of the persistent
stepping kernel: FP32 stages/FFTs, FP64 coefficients, error norm and dt
controller.  A "step" is one such batch integrated t0 -> tf.

Reported beside the value:
  * the plain per-instance FP64 drop-in (one ETD35.evolve per trajectory),
  * rel L2 error at tf of every FP32 trajectory against the FP64 engine,
  * the HBM roofline of the stepping kernel (algorithmic bytes per launch
    from the per-attempt byte model, divided by the launch time measured with
    HIP events on the launch stream),
  * the CPU baseline: the NumPy FP64 oracle on a bounded sample of the same
    workload on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
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


def bytes_per_attempt(m, e_c, e_L, S=6, R=25, W=6):
    """Algorithmic HBM bytes of one ETD35 attempt: M [(R+W) e_c + S e_L]."""
    return m * ((R + W) * e_c + S * e_L)


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
    u0 = [torch.as_tensor(u, dtype=torch.complex64, device=dev) for u in u0_host]
    uf = [torch.empty_like(u) for u in u0]
    plan = Plan("ETD35", nl, _lib.F32, cfg, None, [nl.descriptor(L) for _ in range(B)])
    t0s, tfs, hs = [0.0] * B, [TF] * B, [TF / 100.0] * B

    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        plan.evolve(u0, uf, t0s, tfs, hs)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    kernel_ms = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ev0.record(stream)
        plan.evolve(u0, uf, t0s, tfs, hs)
        ev1.record(stream)
        ev1.synchronize()
        kernel_ms.append(ev0.elapsed_time(ev1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    stats = plan.stats()
    status = [s["status"] for s in stats]
    attempts = np.array([s["attempts"] for s in stats])
    steps = np.array([s["steps"] for s in stats])

    # FP64 engine reference for every instance, and the plain per-instance FP64 drop-in
    plan64 = Plan("ETD35", nl, _lib.F64, cfg, None, [nl.descriptor(L) for _ in range(B)])
    u064 = [torch.as_tensor(u, dtype=torch.complex128, device=dev) for u in u0_host]
    uf64 = [torch.empty_like(u) for u in u064]
    plan64.evolve(u064, uf64, t0s, tfs, hs)
    torch.cuda.synchronize()
    errs = np.array([float(torch.linalg.norm(uf[i].to(torch.complex128) - uf64[i])
                           / torch.linalg.norm(uf64[i])) for i in range(B)])
    plain = ETD35(torch.as_tensor(L), nl, config=cfg)
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for i in range(args.plain):
        plain.evolve(u064[i], 0.0, TF, store_data=False)
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - tp) / args.plain

    ms_step = wall / args.steps * 1e3
    value = B * args.steps / wall
    b_att = bytes_per_attempt(N, 8, 4)
    alg_bytes = float(attempts.sum()) * b_att + float(steps.sum()) * N * 8
    k_ms = float(np.mean(kernel_ms))
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "nls2soliton_n8192_etd35_tf20_eps1e-6_batched", "method": "ETD35",
                   "n": N, "tf": TF, "epsilon": EPS, "batch": B,
                   "precision": "f32 stages and FFTs; f64 coefficients, error norm, dt control",
                   "parallelism": "1 GPU, one workgroup per trajectory, persistent kernel"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "accuracy": {"rel_l2_vs_fp64_max": float(errs.max()), "rel_l2_vs_fp64_median":
                     float(np.median(errs)), "all_ok": bool(all(s == 0 for s in status))},
        "attempts_per_trajectory": float(attempts.mean()),
        "kernel_ms_per_launch": round(k_ms, 3),
        "plain_fp64": {"value": round(1.0 / plain_s, 4), "unit": "trajectories/s",
                       "s_per_trajectory": round(plain_s, 4)},
        "batched_over_plain": round(value * plain_s, 2),
    }
    if not args.no_cpu:
        workers = args.cpu_workers or min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(workers)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
