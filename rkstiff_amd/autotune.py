"""Launch-shape autotuner for the persistent stepping kernels.

A plan groups its trajectories in size buckets, one persistent-kernel launch
per bucket.  Two launch choices change throughput and not results:

* whether the buckets run one after another on the caller's stream or
  overlap on private streams.  In a mixed-size ensemble a bucket's last
  long trajectories otherwise leave most CUs idle before the next bucket
  starts;
* how many workgroups of each bucket stay resident per CU (fewer resident
  trajectories share a CU's caches and LDS bandwidth with fewer others);
* the FFT tile of each bucket where a second one exists (n = 512 .. 4096):
  twice the threads per trajectory, half the modes per thread in the stage
  passes and transforms.

``tune_launch`` times short pilot evolves of the plan's own trajectories and
keeps the fastest shape.  It tries concurrent (the plan default) and
sequential buckets first,
then, bucket by bucket (largest n first), the second tile and then half and
a quarter of the resident workgroups.  A change is kept only when it is faster by more than
``min_gain``.  The shape is applied to the plan and returned with the timings
of every trial.
"""
import time

import numpy as np
import torch


def _timed(plan, u0, uf, t0s, tfs, hs, reps):
    plan.evolve(u0, uf, t0s, tfs, hs)  # warm
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        plan.evolve(u0, uf, t0s, tfs, hs)
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps


def tune_launch(plan, u0, t0, tf, h, reps=1, min_gain=0.02):
    """Tune ``plan``'s launch shape on pilot evolves over [t0, tf] from ``u0``.

    u0: list of device state tensors (the plan's dtype), one per trajectory;
    t0/tf/h: scalars or per-trajectory sequences (h: first step).
    Returns {"concurrent", "wg_per_cu", "seconds", "trials"}.
    """
    n = plan.ninst
    t0s = np.broadcast_to(np.asarray(t0, dtype=np.float64), (n,)).tolist()
    tfs = np.broadcast_to(np.asarray(tf, dtype=np.float64), (n,)).tolist()
    hs = np.broadcast_to(np.asarray(h, dtype=np.float64), (n,)).tolist()
    uf = [torch.empty_like(u) for u in u0]
    buckets = plan.buckets()
    trials = []
    if not buckets:
        # the 2-D engine: a fixed kernel sequence, nothing to choose
        sec = _timed(plan, u0, uf, t0s, tfs, hs, reps)
        return {"concurrent": None, "wg_per_cu": [], "tile": [], "seconds": sec,
                "trials": [{"concurrent": None, "wg_per_cu": [], "tile": [], "seconds": sec}], "buckets": []}

    def trial(conc, wg, ti):
        plan.set_launch(conc, wg, ti)
        sec = _timed(plan, u0, uf, t0s, tfs, hs, reps)
        trials.append({"concurrent": conc, "wg_per_cu": list(wg), "tile": list(ti), "seconds": sec})
        return sec

    nb = len(buckets)
    best_wg, best_ti = [0] * nb, [0] * nb
    best_conc = True  # the plan's default shape is the first trial
    best = trial(True, best_wg, best_ti)
    if nb > 1:
        sec = trial(False, best_wg, best_ti)
        if sec < best * (1.0 - min_gain):
            best, best_conc = sec, False
    for b, (_l2, _cnt, _mx, ntile) in enumerate(buckets):
        if ntile > 1:
            ti = list(best_ti)
            ti[b] = 1
            sec = trial(best_conc, best_wg, ti)
            if sec < best * (1.0 - min_gain):
                best, best_ti = sec, ti
    plan.set_launch(best_conc, best_wg, best_ti)
    for b in range(nb):
        # resident-workgroup limit of the chosen tile
        mx = plan.buckets()[b][2] if best_ti[b] == 0 else None
        if mx is None:
            continue
        for v in sorted({mx // 2, mx // 4} - {0, mx}, reverse=True):
            wg = list(best_wg)
            wg[b] = v
            sec = trial(best_conc, wg, best_ti)
            if sec < best * (1.0 - min_gain):
                best, best_wg = sec, wg
    plan.set_launch(best_conc, best_wg, best_ti)
    return {"concurrent": best_conc, "wg_per_cu": best_wg, "tile": best_ti, "seconds": best, "trials": trials,
            "buckets": [{"log2n": l2, "count": c, "max_wg_per_cu": m, "tiles": t} for l2, c, m, t in buckets]}
