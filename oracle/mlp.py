"""CPU oracle of the station-beam surrogate network (TEST INFRASTRUCTURE ONLY).

The surrogate has no counterpart in the reference: it approximates the
station beam E(l, m, f) that the reference evaluates exactly
(oskar/telescope/station/src/oskar_evaluate_station_beam_aperture_array.c:31-288:
element pattern x array-factor DFT, normalised), as a 2x2 complex matrix
per direction (8 reals, the layout of oskar_Mem OSKAR_*_COMPLEX_MATRIX).

Network (fixed architecture):
  x  = encode(l, m, f)                          D_IN = 128 features
  a0 = silu(W0 x + b0)                          256
  a_i = silu(W_i a_{i-1} + b_i), i = 1..3       256
  y  = W4 a3 + b4                               N_OUT_PAD = 32 rows, the
                                                first 8 are E, the rest zero
Loss: mean over samples and the 8 outputs of (y - E)^2.

Numerics of the GPU kernels, restated here so that the oracle computes what
they compute: matrix products take BF16 operands (round to nearest even)
and accumulate in FP32 (here: exactly, in FP64); the encoding, each
activation a_i, the loss gradient dy and each dz_i are rounded to BF16
where the kernels store them for the next product (bf16=True); weights
take part as their BF16 copies.  With bf16=False the same functions are
the plain FP64 network, whose analytic gradient the tests check against
finite differences.
"""
import numpy as np

D_IN = 128
NF = 56                  # Fourier frequencies: 2 * NF features
WIDTH = 256
DEPTH = 4
N_OUT = 8
N_OUT_PAD = 32
LAYERS = [(WIDTH, D_IN)] + [(WIDTH, WIDTH)] * (DEPTH - 1) + [
    (N_OUT_PAD, WIDTH)]
NUM_PARAMS = sum(n * k + n for n, k in LAYERS)


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as FP64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(a)
    out[nan] = np.nan
    return out


def unpack(theta):
    """Flat parameter vector -> [(W_i [n x k], b_i [n])]."""
    out, o = [], 0
    for n, k in LAYERS:
        W = theta[o:o + n * k].reshape(n, k)
        o += n * k
        b = theta[o:o + n]
        o += n
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def encode(l, m, f, fourier, l0, m0, f_min, f_max):
    """The 128 input features of directions (l, m) at frequencies f (Hz).

    fourier: (NF, 2) spatial frequencies (cycles per unit of the scaled
    offsets u = (f / f_max)(l - l0), v = (f / f_max)(m - m0)).  Features:
    sin and cos of 2 pi (F_j0 u + F_j1 v) for each j, then 16 smooth terms
    of (l, m, n, fh) with fh = (2 f - f_min - f_max) / (f_max - f_min)."""
    l = np.asarray(l, np.float64)
    m = np.asarray(m, np.float64)
    f = np.broadcast_to(np.asarray(f, np.float64), l.shape)
    n = np.sqrt(np.maximum(0.0, 1.0 - l * l - m * m))
    s = f / f_max
    u, v = s * (l - l0), s * (m - m0)
    ph = 2 * np.pi * (np.outer(u, fourier[:, 0]) + np.outer(v, fourier[:, 1]))
    fh = (2 * f - f_min - f_max) / (f_max - f_min)
    smooth = np.stack([l, m, n, fh, fh * fh, l * fh, m * fh, n * fh, l * l,
                       m * m, l * m, n * n, np.sin(np.pi * fh),
                       np.cos(np.pi * fh), l * n, m * n], axis=1)
    return np.concatenate([np.sin(ph), np.cos(ph), smooth], axis=1)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(x, theta, bf16_numerics=True):
    """Returns (y [B x 32], cache)."""
    r = bf16 if bf16_numerics else (lambda a: np.asarray(a, np.float64))
    layers = unpack(np.asarray(theta, np.float64))
    a = r(x)
    acts, zs = [a], []
    for i, (W, b) in enumerate(layers):
        z = a @ r(W).T + b
        if i == len(layers) - 1:
            y = z
            break
        if bf16_numerics:
            # The kernels keep z in FP32 for the backward pass.
            z = np.asarray(z, np.float32).astype(np.float64)
        zs.append(z)
        a = r(silu(z))
        acts.append(a)
    return y, dict(acts=acts, zs=zs, layers=layers)


def loss_and_grad(x, target, theta, bf16_numerics=True):
    """Mean-square loss over the 8 outputs and its gradient with respect to
    the flat parameter vector (rows 8..31 of the last layer get zero)."""
    r = bf16 if bf16_numerics else (lambda a: np.asarray(a, np.float64))
    y, c = forward(x, theta, bf16_numerics)
    B = x.shape[0]
    d = np.zeros_like(y)
    d[:, :N_OUT] = y[:, :N_OUT] - target
    loss = float(np.sum(d[:, :N_OUT] ** 2) / (B * N_OUT))
    dy = r(2.0 * d / (B * N_OUT))
    grads = []
    layers = c["layers"]
    dz = dy
    for i in range(len(layers) - 1, -1, -1):
        W, b = layers[i]
        a_in = c["acts"][i]
        gW = dz.T @ a_in
        gb = dz.sum(axis=0)
        if i == len(layers) - 1:
            gW[N_OUT:] = 0.0
            gb[N_OUT:] = 0.0
        grads.append((gW, gb))
        if i == 0:
            break
        da = dz @ r(W)
        dz = r(da * dsilu(c["zs"][i - 1]))
    return loss, pack(grads[::-1])


def init_params(seed):
    """Uniform(-1/sqrt(k), 1/sqrt(k)) weights, zero biases; rows 8..31 of
    the output layer zero."""
    rng = np.random.default_rng(seed)
    layers = []
    for i, (n, k) in enumerate(LAYERS):
        lim = 1.0 / np.sqrt(k)
        W = rng.uniform(-lim, lim, (n, k))
        b = np.zeros(n)
        if i == len(LAYERS) - 1:
            W[N_OUT:] = 0.0
        layers.append((W, b))
    return pack(layers)
