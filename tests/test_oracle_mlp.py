"""Pins of the proposer oracle (oracle/mlp.py) on the CPU.

* backward: central finite differences of the loss in FP64 without rounding,
  on every parameter block (a dropped term, a transposed operand or a wrong
  activation derivative fails one of them);
* BF16 rounding: torch's float32 -> bfloat16 conversion (round to nearest
  even) and hand-checked ties;
* GELU: torch's exact (erf) GELU and the derivative by differences;
* features: a per-mode loop written from the definition, single-mode states
  and the amplitude-shift identity;
* targets: a synthetic log with err ~ h^q, where h' s' is known in closed form;
* Adam: the first step moves every coordinate by lr * sign(g).
"""
import math

import numpy as np
import pytest
import torch

from oracle import mlp


def small_params(seed, scale=1.0):
    p = mlp.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    for k in ("b1", "b2", "b3", "b4"):
        p[k] = 0.1 * rng.standard_normal(p[k].shape)
    return {k: v * (scale if k.startswith("W") else 1.0) for k, v in p.items()}


def random_inputs(B, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((B, mlp.K0))
    X[:, :mlp.NIN] = rng.standard_normal((B, mlp.NIN))
    y = 0.3 * rng.standard_normal(B)
    return X, y


@pytest.mark.parametrize("kind", ["mse", "pinball"])
def test_backward_matches_finite_differences(kind):
    p = small_params(1)
    X, y = random_inputs(8, 2)
    yhat, cache = mlp.forward(p, X, quant=False)
    if kind == "pinball":
        # keep every residual away from the kink so the loss is smooth there
        y = yhat + np.where(np.arange(8) % 2 == 0, 0.5, -0.5)
    g = mlp.loss_grad(yhat, y, kind, tau=0.2)
    gr = mlp.backward(p, cache, g, quant=False)
    rng = np.random.default_rng(3)
    eps = 1e-6
    for name, shape in mlp.SHAPES:
        flat = p[name].reshape(-1)
        # input-padding columns of W1 have exactly zero gradient
        cand = np.arange(flat.size)
        if name == "W1":
            cand = cand[(cand % mlp.K0) < mlp.NIN]
            pad = gr["W1"][:, mlp.NIN:]
            assert np.all(pad == 0.0)
        for i in rng.choice(cand, size=min(6, cand.size), replace=False):
            old = flat[i]
            flat[i] = old + eps
            lp = mlp.loss(mlp.forward(p, X, quant=False)[0], y, kind, tau=0.2)
            flat[i] = old - eps
            lm = mlp.loss(mlp.forward(p, X, quant=False)[0], y, kind, tau=0.2)
            flat[i] = old
            fd = (lp - lm) / (2 * eps)
            an = gr[name].reshape(-1)[i]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (name, i, fd, an)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-8, 8, 10000),
                        [0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3.0 * 2.0 ** -8, 65504.0, 3.0e38]])
    x32 = x.astype(np.float32)
    ref = torch.from_numpy(x32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mlp.bf16(x32), ref)
    # ties go to the even mantissa
    assert mlp.bf16(1.0 + 2.0 ** -8) == 1.0
    assert mlp.bf16(1.0 + 3.0 * 2.0 ** -8) == 1.0 + 2.0 ** -6


def test_gelu_and_derivative():
    z = np.linspace(-8.0, 8.0, 2001)
    ref = torch.nn.functional.gelu(torch.from_numpy(z), approximate="none").numpy()
    np.testing.assert_allclose(mlp.gelu(z), ref, rtol=1e-14, atol=1e-15)
    d = (mlp.gelu(z + 1e-6) - mlp.gelu(z - 1e-6)) / 2e-6
    np.testing.assert_allclose(mlp.gelu_grad(z), d, rtol=1e-7, atol=1e-8)
    assert mlp.gelu(0.0) == 0.0 and mlp.gelu_grad(0.0) == 0.5


def features_loop(u_hat, n, real, h, s):
    """The definition, one mode at a time."""
    sums = [0.0] * 64
    cnt = [0] * 64
    for m, v in enumerate(u_hat):
        w = m if real else min(m, n - m)
        b = (w * 64) // (n // 2 + 1)
        sums[b] += abs(v) ** 2
        cnt[b] += 1
    out = [0.5 * math.log(sums[b] / cnt[b] / (n * n) + 1e-30) for b in range(64)]
    return np.array(out + [math.log(h), math.log(min(s, 1e4))])


@pytest.mark.parametrize("n,real", [(256, True), (1024, True), (128, False), (512, False)])
def test_features_definition(n, real):
    rng = np.random.default_rng(n)
    m = n // 2 + 1 if real else n
    u = (rng.standard_normal(m) + 1j * rng.standard_normal(m)) * np.exp(-np.arange(m) / 40.0)
    f = mlp.features(u, n, real, 0.01, 1.3)
    np.testing.assert_allclose(f, features_loop(u, n, real, 0.01, 1.3), rtol=1e-12, atol=1e-12)
    # every bin holds at least one mode at these sizes
    w = mlp.wavenumber_index(np.arange(m), n, real)
    assert len(np.unique((w * 64) // (n // 2 + 1))) == 64
    # scaling the state by c shifts every spectral input by ln c
    f2 = mlp.features(3.0 * u, n, real, 0.01, 1.3)
    np.testing.assert_allclose(f2[:64] - f[:64], math.log(3.0), atol=1e-9)
    # zero-error cap
    assert mlp.features(u, n, real, 0.01, 1e300)[65] == pytest.approx(math.log(1e4))


def test_features_single_mode():
    n = 256
    u = np.zeros(n, complex)
    u[5] = n * 2.0        # physical amplitude 2 at |k| index 5 (fft layout)
    u[n - 5] = n * 2.0
    f = mlp.features(u, n, False, 0.1, 2.0)
    b = (5 * 64) // (n // 2 + 1)
    cnt = sum(1 for m in range(n) if (min(m, n - m) * 64) // (n // 2 + 1) == b)
    assert f[b] == pytest.approx(0.5 * math.log(2 * 4.0 / cnt + 1e-30), rel=1e-12)
    others = np.delete(f[:64], b)
    assert np.all(others == pytest.approx(0.5 * math.log(1e-30)))


def test_targets_closed_form():
    # err = C h^q  =>  s = safety (tol / (C h^q))^(1/q) = K / h with K = safety (tol/C)^(1/q);
    # then h' s' = K for every attempt, and the target is ln(K / h).
    q, safety, tol, C = 4, 0.8, 1e-6, 3.0
    K = safety * (tol / C) ** (1.0 / q)
    h = np.array([0.01, 0.012, 0.02, 0.015, 0.018])
    s = safety * (tol / (C * h ** q)) ** (1.0 / q)
    np.testing.assert_allclose(s * h, K, rtol=1e-13)
    y = mlp.targets(h, s, np.array([0, 1, 2, 3]), min_s=0.01, max_s=100.0)
    np.testing.assert_allclose(y, np.log(K / h[:4]), rtol=1e-12)
    # clipping to the proposal range
    y = mlp.targets(h, s, np.array([0]), min_s=0.25, max_s=1.5)
    assert y[0] == pytest.approx(min(math.log(K / h[0]), math.log(1.5)))


def test_adam_first_step_is_sign():
    rng = np.random.default_rng(9)
    v = rng.standard_normal(50)
    g = rng.standard_normal(50)
    v1, m, s = mlp.adam(v, g, np.zeros(50), np.zeros(50), 1, lr=1e-3, eps=0.0)
    np.testing.assert_allclose(v - v1, 1e-3 * np.sign(g), rtol=1e-12)


def test_quantized_forward_rounds_where_stated():
    p = small_params(4)
    X, _ = random_inputs(4, 6)
    yq, cq = mlp.forward(p, X, quant=True)
    # activations and inputs are exactly representable in BF16
    for k in ("a0", "a1", "a2", "a3"):
        np.testing.assert_array_equal(cq[k], mlp.bf16(cq[k]))
    yf, _ = mlp.forward(p, X, quant=False)
    # the rounding moves the output by a few BF16 ulps of its terms, not more
    assert np.max(np.abs(yq - yf)) < 0.05 * (1.0 + np.max(np.abs(yf)))
