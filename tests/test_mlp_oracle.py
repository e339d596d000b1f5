"""Pins the CPU oracle of the beam-surrogate network (oracle/mlp.py).

- BF16 rounding: known values, including ties to even.
- Encoding: closed forms at the beam direction and of the smooth terms.
- Gradient: the analytic back-propagation of the FP64 network against
  central finite differences of its loss, on parameters of every layer
  (a dropped term, a wrong sign, a transposed weight or a wrong SiLU
  derivative fails it).
- BF16 numerics: the emulated-rounding gradient stays close to the FP64
  one (the rounding points are where the kernels store BF16 values).
"""
import numpy as np

from oracle import mlp


def test_bf16_rounding_known_values():
    cases = [
        (1.0, 1.0),
        (1.0 + 2.0 ** -9, 1.0),                   # below half: down
        (1.0 + 2.0 ** -8, 1.0),                   # tie, 1.0 is even
        (1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -6),   # tie, odd -> up to even
        (1.0 + 1.5 * 2.0 ** -8, 1.0 + 2.0 ** -7),  # above half: up
        (-2.5, -2.5),
        # 2^100 (1 + 2.5/128 + 1/4096): above the half-way point -> 3/128.
        (2.0 ** 100 * (1 + 5 * 2.0 ** -8 + 2.0 ** -12),
         2.0 ** 100 * (1 + 3 * 2.0 ** -7)),
        (0.0, 0.0),
    ]
    for x, want in cases:
        got = float(mlp.bf16(np.array([x]))[0])
        assert got == want, (x, got, want)
    assert np.isnan(mlp.bf16(np.array([np.nan]))[0])


def test_encoding_closed_forms():
    rng = np.random.default_rng(0)
    fourier = rng.normal(size=(mlp.NF, 2)) * 5
    l0, m0, fmin, fmax = 0.3, -0.2, 50e6, 350e6
    x = mlp.encode(np.array([l0]), np.array([m0]), np.array([fmax]),
                   fourier, l0, m0, fmin, fmax)[0]
    assert x.shape == (mlp.D_IN,)
    assert np.allclose(x[:mlp.NF], 0.0)
    assert np.allclose(x[mlp.NF:2 * mlp.NF], 1.0)
    n = np.sqrt(1 - l0 ** 2 - m0 ** 2)
    fh = 1.0
    smooth = [l0, m0, n, fh, fh * fh, l0 * fh, m0 * fh, n * fh, l0 * l0,
              m0 * m0, l0 * m0, n * n, np.sin(np.pi * fh), np.cos(np.pi * fh),
              l0 * n, m0 * n]
    assert np.allclose(x[2 * mlp.NF:], smooth)
    # A shifted direction at half the top frequency: phase of frequency j
    # is 2 pi (F_j0 u + F_j1 v) with u = 0.5 (l - l0), v = 0.5 (m - m0).
    l, m, f = 0.1, 0.25, 0.5 * fmax
    y = mlp.encode(np.array([l]), np.array([m]), np.array([f]), fourier, l0,
                   m0, fmin, fmax)[0]
    ph = 2 * np.pi * (fourier[:, 0] * 0.5 * (l - l0) +
                      fourier[:, 1] * 0.5 * (m - m0))
    assert np.allclose(y[:mlp.NF], np.sin(ph))
    assert np.allclose(y[mlp.NF:2 * mlp.NF], np.cos(ph))
    assert np.isclose(y[2 * mlp.NF + 3], (2 * f - fmin - fmax) / (fmax - fmin))


def _case(seed, B=6):
    rng = np.random.default_rng(seed)
    theta = mlp.init_params(seed)
    # Non-zero biases so that every layer's bias gradient is exercised.
    layers = mlp.unpack(theta.copy())
    layers = [(W, rng.normal(scale=0.1, size=b.shape) * (
        1 if i < len(layers) - 1 else 0)) for i, (W, b) in enumerate(layers)]
    for i, (W, b) in enumerate(layers[-1:]):
        b[:mlp.N_OUT] = rng.normal(scale=0.1, size=mlp.N_OUT)
    theta = mlp.pack(layers)
    x = rng.normal(size=(B, mlp.D_IN))
    t = rng.normal(size=(B, mlp.N_OUT))
    return theta, x, t


def test_oracle_gradient_matches_finite_differences():
    theta, x, t = _case(1)
    loss, g = mlp.loss_and_grad(x, t, theta, bf16_numerics=False)
    assert g.shape == (mlp.NUM_PARAMS,)
    rng = np.random.default_rng(2)
    # Parameters of every layer: weights and biases (used rows only).
    picks = []
    o = 0
    for i, (n, k) in enumerate(mlp.LAYERS):
        rows = mlp.N_OUT if i == len(mlp.LAYERS) - 1 else n
        for _ in range(8):
            picks.append(o + rng.integers(0, rows) * k + rng.integers(0, k))
        for _ in range(3):
            picks.append(o + n * k + rng.integers(0, rows))
        o += n * k + n
    h = 1e-5
    for p in picks:
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        lp, _ = mlp.loss_and_grad(x, t, tp, bf16_numerics=False)
        lm, _ = mlp.loss_and_grad(x, t, tm, bf16_numerics=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9, (p, fd,
                                                                    g[p])
    # Unused output rows get no gradient.
    W4_off = sum(n * k + n for n, k in mlp.LAYERS[:-1])
    n4, k4 = mlp.LAYERS[-1]
    assert not np.any(g[W4_off + mlp.N_OUT * k4:W4_off + n4 * k4])


def test_oracle_bf16_numerics_stay_close_to_fp64():
    theta, x, t = _case(3, B=64)
    l64, g64 = mlp.loss_and_grad(x, t, theta, bf16_numerics=False)
    l16, g16 = mlp.loss_and_grad(x, t, theta, bf16_numerics=True)
    assert abs(l16 - l64) < 1e-2 * l64
    cos = np.dot(g16, g64) / np.linalg.norm(g16) / np.linalg.norm(g64)
    assert cos > 0.999
