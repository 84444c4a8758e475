"""Learned step-size proposer for batched evolves (opt-in).

The reference has no learned component.  After an accepted attempt its
controller keeps h unless s > incr_f, and then takes s * h
(solveras.py:543-554).  ``DtProposer`` replaces only that proposal.  A small
MLP reads the accepted state's 64-bin log-amplitude spectrum, ln h and ln s,
and predicts y ~ ln(h' s' / h): the log ratio between the largest step that
the next attempt would accept and the step just taken.  The step proposed is
h * exp(clamp(y - margin, ln min_s, ln max_s)).

These rules do not change:
* the embedded error estimate judges every attempt;
* a rejection uses the reference's rule;
* a non-finite output falls back to the reference's proposal.

The network is 66 (padded to 96) -> 256 -> 256 -> 256 -> 1 with exact GELU,
BF16 weights and FP32 accumulation.  Its forward, backward and Adam kernels
run on the device (rkstiff_amd/csrc/rks_mlp.hip).  Batch-1 inference runs
inside each trajectory's workgroup after an accepted step.

Workflow::

    solver = BatchSolver("ETD35", problems, precision="f32")
    solver.record(cap=400)                  # training records during evolve
    solver.evolve(u0, 0.0, tf)
    data = training_set(*solver.records(), solver.plan.cfg)
    net = DtProposer.for_data(data, seed=0)
    net.train(data, steps=3000)
    solver.set_proposer(net, margin=0.0)    # next evolves use the proposals
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib

NIN = _lib.PROPOSER_NIN
REC = _lib.PROPOSER_REC
NPARAM = _lib.PROPOSER_NPARAM
K0, HID = 96, 256
SHAPES = (("W1", (HID, K0)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("w4", (HID,)), ("b4", (1,)))
LOSSES = {"pinball": 0, "mse": 1}


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def init_params(seed=0):
    """He-uniform weights (the first layer over its 66 real inputs), zero biases; float32."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith("b"):
            parts.append(np.zeros(shape))
        elif name == "W1":
            w = np.zeros(shape)
            lim = math.sqrt(6.0 / NIN)
            w[:, :NIN] = rng.uniform(-lim, lim, size=(HID, NIN))
            parts.append(w)
        else:
            lim = math.sqrt(6.0 / shape[-1]) if name != "w4" else math.sqrt(3.0 / shape[-1])
            parts.append(rng.uniform(-lim, lim, size=shape))
    v = np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)
    assert v.size == NPARAM
    return v


def training_set(records, counts, cfg, max_rows=None):
    """Valid rows of recorded training data as a (N, 68) float32 device tensor.

    records: (ninst, cap, 68) device tensor; counts: records per instance;
    cfg: the plan's controller constants (min_s, max_s).  Targets are
    clipped to [ln min_s, ln max_s], the range of the proposals.
    """
    rows = [records[i, :c] for i, c in enumerate(counts) if c > 0]
    if not rows:
        return torch.empty((0, REC), dtype=torch.float32, device=records.device)
    data = torch.cat(rows)
    data = data[(data[:, NIN + 1] > 0.5) & torch.isfinite(data[:, :NIN + 1]).all(dim=1)]
    data = data.clone()
    data[:, NIN] = data[:, NIN].clamp(math.log(cfg.min_s), math.log(cfg.max_s))
    if max_rows is not None and data.shape[0] > max_rows:
        data = data[:max_rows]
    return data.contiguous()


class DtProposer:
    """A device-resident proposer network (one C-ABI ``rks_mlp`` handle)."""

    def __init__(self, params=None, mu=None, sigma=None, seed=0):
        if not torch.cuda.is_available():
            raise RuntimeError("rkstiff_amd needs a HIP device (torch.cuda.is_available() is False)")
        self.lib = _lib.load()
        p = init_params(seed) if params is None else np.ascontiguousarray(params, dtype=np.float32)
        if p.shape != (NPARAM,):
            raise ValueError(f"params must have {NPARAM} values")
        self.mu = np.zeros(NIN, np.float32) if mu is None else np.ascontiguousarray(mu, np.float32)
        self.sigma = np.ones(NIN, np.float32) if sigma is None else np.ascontiguousarray(sigma, np.float32)
        if self.mu.shape != (NIN,) or self.sigma.shape != (NIN,):
            raise ValueError(f"mu and sigma must have {NIN} values")
        h = ctypes.c_void_p()
        _lib.check(self.lib.rks_mlp_create(_fptr(p), _fptr(self.mu), _fptr(self.sigma), ctypes.byref(h)))
        self.h = h
        self.device = torch.cuda.current_device()
        self.history = []

    @classmethod
    def for_data(cls, data, seed=0, params=None):
        """A proposer whose input scaling is the mean / std of ``data``'s inputs."""
        x = data[:, :NIN].double()
        mu = x.mean(dim=0).cpu().numpy()
        sd = x.std(dim=0).cpu().numpy()
        return cls(params=params, mu=mu, sigma=np.maximum(sd, 1e-3), seed=seed)

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                self.lib.rks_mlp_destroy(h)
            except Exception:
                pass
            self.h = None

    def params(self):
        out = np.empty(NPARAM, np.float32)
        _lib.check(self.lib.rks_mlp_get_params(self.h, _fptr(out)))
        return out

    @staticmethod
    def _check_data(data):
        if data.dtype != torch.float32 or not data.is_cuda or data.dim() != 2 or data.shape[1] != REC:
            raise ValueError(f"records must be a (N, {REC}) float32 device tensor")
        return data.contiguous()

    def train(self, data, steps, batch=1024, lr=1e-3, loss="pinball", tau=0.1, seed=0):
        """``steps`` Adam steps on minibatches drawn from ``data``; returns the loss per step."""
        data = self._check_data(data)
        if loss not in LOSSES:
            raise ValueError(f"loss must be one of {sorted(LOSSES)}")
        n = data.shape[0]
        if n == 0:
            raise ValueError("no training records")
        g = torch.Generator(device="cpu").manual_seed(seed)
        per_epoch = max(1, n // batch)
        epochs = (steps + per_epoch - 1) // per_epoch
        idx = torch.cat([torch.randperm(n, generator=g)[: per_epoch * batch] if n >= batch
                         else torch.randint(0, n, (batch,), generator=g) for _ in range(epochs)])
        idx = idx[: steps * batch].to(torch.int32).to(data.device)
        losses = torch.empty(steps, dtype=torch.float32, device=data.device)
        _lib.check(self.lib.rks_mlp_train(self.h, ctypes.c_void_p(data.data_ptr()), n,
                                          ctypes.c_void_p(idx.data_ptr()), batch, steps, LOSSES[loss],
                                          float(tau), float(lr), ctypes.c_void_p(losses.data_ptr()),
                                          _stream()))
        out = losses.cpu().numpy()
        self.history.extend(out.tolist())
        return out

    def predict(self, data):
        """Network output y for each record (batched GEMM path)."""
        data = self._check_data(data)
        y = torch.empty(data.shape[0], dtype=torch.float32, device=data.device)
        _lib.check(self.lib.rks_mlp_forward(self.h, ctypes.c_void_p(data.data_ptr()), data.shape[0],
                                            ctypes.c_void_p(y.data_ptr()), _stream()))
        return y

    def predict_single(self, data, threads=256):
        """Network output through the batch-1 path that the stepping kernels run."""
        data = self._check_data(data)
        y = torch.empty(data.shape[0], dtype=torch.float32, device=data.device)
        _lib.check(self.lib.rks_mlp_infer_single(self.h, ctypes.c_void_p(data.data_ptr()), data.shape[0],
                                                 int(threads), ctypes.c_void_p(y.data_ptr()), _stream()))
        return y

    def gradient(self, data, loss="pinball", tau=0.1):
        """(gradient of the mean loss, outputs, mean loss) without an update; len(data) % 64 == 0."""
        data = self._check_data(data)
        if loss not in LOSSES:
            raise ValueError(f"loss must be one of {sorted(LOSSES)}")
        grad = torch.empty(NPARAM, dtype=torch.float32, device=data.device)
        y = torch.empty(data.shape[0], dtype=torch.float32, device=data.device)
        lv = torch.empty(1, dtype=torch.float32, device=data.device)
        _lib.check(self.lib.rks_mlp_gradient(self.h, ctypes.c_void_p(data.data_ptr()), data.shape[0],
                                             LOSSES[loss], float(tau), ctypes.c_void_p(grad.data_ptr()),
                                             ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(lv.data_ptr()),
                                             _stream()))
        return grad, y, float(lv.item())

    def state_dict(self):
        return {"params": self.params(), "mu": self.mu.copy(), "sigma": self.sigma.copy()}

    @classmethod
    def from_state_dict(cls, d):
        return cls(params=d["params"], mu=d["mu"], sigma=d["sigma"])
