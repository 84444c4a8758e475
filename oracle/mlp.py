"""FP64 restatement of the learned step-size proposer (oracle; test infrastructure).

The reference has no learned component: its next step after an accepted
attempt is ``s * h`` when ``s > incr_f`` and ``h`` otherwise
(solveras.py:543-554).  The proposer is an opt-in addition of this project.
This module is the checker for its device kernels, and it holds four parts:

* ``features``: the 66 inputs computed from an accepted state.  Inputs 0-63
  are a 64-bin log-amplitude spectrum.  Input 64 is ln h of the accepted
  step.  Input 65 is ln s of the accepted attempt.
* ``targets``: the training target taken from an attempt log.
* ``forward`` / ``backward``: the network, 66 (padded to 96) -> 256 -> 256
  -> 256 -> 1 with exact GELU.  It is evaluated with the device's BF16
  rounding points, so the device result can be compared element by element.
* ``adam``: one Adam update of the FP32 master parameters.

Pinned by: central finite differences of ``loss`` against ``backward`` in
FP64 without rounding (tests/test_oracle_mlp.py); a brute-force loop
over modes for ``features``; the closed form of GELU and its derivative
(erf); Adam's first step, whose update is exactly ``lr * sign(g)``.
"""
import math

import numpy as np
from scipy.special import erf

NBIN = 64
NIN = 66          # 64 bins + ln h + ln s
K0 = 96           # first-layer input width, padded for 16x16x32 MFMA tiles
HID = 256
S_FEAT_CAP = 1.0e4  # ln s is taken of min(s, S_FEAT_CAP) (s is DBL_MAX on a zero error)


# ----------------------------------------------------------------------------
# BF16
# ----------------------------------------------------------------------------
def bf16(x):
    """Round to bfloat16 (round to nearest even from float32), returned as float64."""
    f = np.asarray(x, dtype=np.float64).astype(np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


# ----------------------------------------------------------------------------
# parameters
# ----------------------------------------------------------------------------
SHAPES = (("W1", (HID, K0)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("w4", (HID,)), ("b4", (1,)))
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def init_params(seed):
    """He-uniform weights (first layer only over the 66 real inputs), zero biases.

    The flat FP32 vector follows SHAPES; the device trains this same vector.
    """
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in SHAPES:
        if name.startswith("b"):
            p[name] = np.zeros(shape)
        elif name == "W1":
            w = np.zeros(shape)
            lim = math.sqrt(6.0 / NIN)
            w[:, :NIN] = rng.uniform(-lim, lim, size=(HID, NIN))
            p[name] = w
        else:
            fan_in = shape[-1]
            lim = math.sqrt(6.0 / fan_in) if name != "w4" else math.sqrt(3.0 / fan_in)
            p[name] = rng.uniform(-lim, lim, size=shape)
    return {k: v.astype(np.float32).astype(np.float64) for k, v in p.items()}


def flatten(p):
    return np.concatenate([np.asarray(p[k], dtype=np.float64).reshape(-1) for k, _ in SHAPES])


def unflatten(v):
    p, o = {}, 0
    for name, shape in SHAPES:
        c = int(np.prod(shape))
        p[name] = np.asarray(v[o:o + c], dtype=np.float64).reshape(shape)
        o += c
    return p


# ----------------------------------------------------------------------------
# features and targets
# ----------------------------------------------------------------------------
def wavenumber_index(m, n, real):
    """|k| index of mode m: m for an rfft layout, min(m, n - m) for an fft layout."""
    m = np.asarray(m)
    return m if real else np.minimum(m, n - m)


def features(u_hat, n, real, h, s):
    """Raw proposer inputs of an accepted state u_hat (m modes of an n-point grid).

    Bin b holds the modes whose index w satisfies (w * 64) // (n/2 + 1) == b.
    Input b is 0.5 * ln(mean |u_hat|^2 over the bin / n^2 + 1e-30).
    Input 64 is ln h, and input 65 is ln(min(s, 1e4)).
    """
    u_hat = np.asarray(u_hat)
    m = np.arange(u_hat.shape[0])
    w = wavenumber_index(m, n, real)
    b = (w * NBIN) // (n // 2 + 1)
    p = np.abs(u_hat) ** 2
    out = np.empty(NIN)
    for j in range(NBIN):
        sel = b == j
        cnt = int(sel.sum())
        out[j] = 0.5 * math.log(p[sel].sum() / cnt / (float(n) * n) + 1e-30)
    out[64] = math.log(h)
    out[65] = math.log(min(s, S_FEAT_CAP))
    return out


def targets(log_h, log_s, accepted_idx, min_s=0.25, max_s=4.0):
    """Training target of the accepted attempt at index a: ln(h' s' / h).

    Here (h, s) is attempt a and (h', s') is attempt a + 1.  Because s scales
    as 1/h when err ~ h^q, h' s' is the largest step that the next attempt
    would have accepted (s = 1).  The target is clipped to [ln min_s, ln max_s].
    """
    a = np.asarray(accepted_idx)
    y = np.log(log_h[a + 1] * log_s[a + 1] / log_h[a])
    return np.clip(y, math.log(min_s), math.log(max_s))


def standardize(X, mu, sig):
    """(x - mu) / sig on the 66 inputs, padded with zeros to K0 columns."""
    X = np.atleast_2d(X)
    out = np.zeros((X.shape[0], K0))
    out[:, :NIN] = (X[:, :NIN] - mu) / sig
    return out


# ----------------------------------------------------------------------------
# network
# ----------------------------------------------------------------------------
def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(p, X, quant=True):
    """Network output for standardized inputs X (B x K0).

    With ``quant`` the BF16 rounding points of the device are applied: the
    inputs, every weight matrix/vector and every hidden activation are
    rounded to BF16; biases, pre-activations and the output stay at full
    precision.  Returns (yhat, cache).
    """
    q = bf16 if quant else (lambda v: np.asarray(v, dtype=np.float64))
    a = q(X)
    cache = {"a0": a}
    for i, (W, b) in enumerate((("W1", "b1"), ("W2", "b2"), ("W3", "b3")), start=1):
        z = a @ q(p[W]).T + p[b]
        a = q(gelu(z))
        cache[f"z{i}"] = z
        cache[f"a{i}"] = a
    yhat = a @ q(p["w4"]) + p["b4"][0]
    return yhat, cache


def loss(yhat, y, kind="pinball", tau=0.1):
    """Batch mean of the pinball loss rho_tau(y - yhat), or 0.5 (yhat - y)^2."""
    r = y - yhat
    if kind == "mse":
        return float(np.mean(0.5 * r * r))
    return float(np.mean(r * (tau - (r < 0))))


def loss_grad(yhat, y, kind="pinball", tau=0.1):
    """d loss / d yhat per sample (the batch mean included)."""
    B = yhat.shape[0]
    if kind == "mse":
        return (yhat - y) / B
    return ((y < yhat).astype(np.float64) - tau) / B


def backward(p, cache, g, quant=True):
    """Gradients of sum_b g_b * yhat_b with respect to every parameter.

    Device rounding points (``quant``): the GEMM operands dZ of the hidden
    layers are rounded to BF16 before they multiply W (for dA) and the
    activations (for dW), and the bias gradients sum the rounded dZ.  The
    head gradients use the unrounded g.
    """
    q = bf16 if quant else (lambda v: np.asarray(v, dtype=np.float64))
    gr = {}
    a3 = cache["a3"]
    gr["w4"] = a3.T @ g
    gr["b4"] = np.array([g.sum()])
    dA = np.outer(g, q(p["w4"]))
    for i in (3, 2, 1):
        dZ = q(dA * gelu_grad(cache[f"z{i}"]))
        gr[f"W{i}"] = dZ.T @ cache[f"a{i - 1}"]
        gr[f"b{i}"] = dZ.sum(axis=0)
        if i > 1:
            dA = dZ @ q(p[f"W{i}"])
    return gr


def adam(v, g, m, s, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step on flat vectors (t = 1 on the first step); returns (v, m, s)."""
    m = b1 * m + (1.0 - b1) * g
    s = b2 * s + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    sh = s / (1.0 - b2 ** t)
    return v - lr * mh / (np.sqrt(sh) + eps), m, s


def propose(p, feat, mu, sig, h, min_s=0.25, max_s=4.0, margin=0.0):
    """Next step after an accepted step of size h: h * exp(clip(yhat - margin))."""
    yhat, _ = forward(p, standardize(feat, mu, sig))
    y = float(yhat[0]) - margin
    return h * math.exp(min(max(y, math.log(min_s)), math.log(max_s)))
