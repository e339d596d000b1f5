"""Pins the oracle of oskar_auto_correlate (oracle_auto_correlate_{d,f}).

  - against golden vectors from the reference's own compiled CPU bodies
    (define_auto_correlate.h, tests/golden/acorr.npz) at the reference
    test's shapes and value ranges;
  - against numpy complex matrix algebra (sum_i J B J^H);
  - against a closed form: a diagonal Jones diag(gx, gy) gives
    a = |gx|^2 sum(I+Q), d = |gy|^2 sum(I-Q), b = gx conj(gy) sum(U+iV),
    c = gy conj(gx) sum(U-iV);
  - against the golden-pinned cross-correlation oracle: two stations at one
    position carrying the same Jones give V_10 = J B J^H, the auto term.
A dropped conjugate, a swapped b/c or a missing Q, U or V term fails one of
these.
"""
import numpy as np
import pytest

import oracle
from helpers import load_golden, max_rel


def fixture_case(g, matrix, prec):
    nr = 8 if matrix else 2
    vis = oracle.auto_correlate(
        "double" if prec == "d" else "single", matrix,
        int(g["num_stations"]), g["jones_%d" % matrix], g["I"], g["Q"],
        g["U"], g["V"], vis=g["vis0_%d" % matrix].copy(),
        offset_out=int(g["offset_out"]))
    return vis, nr


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("matrix", [0, 1])
def test_acorr_oracle_matches_reference_kernels(matrix, prec):
    g = load_golden("acorr.npz")
    got, nr = fixture_case(g, matrix, prec)
    ref = g["vis_%d_%s" % (matrix, prec)]
    # Same sums in the same order as the reference's CPU bodies.
    tol = 1e-14 if prec == "d" else 1e-6
    assert max_rel(got, ref) < tol
    # Entries outside [offset, offset + stations) and the imaginary parts
    # of the diagonal terms (matrix) or of the value (scalar) are untouched.
    off, ns = int(g["offset_out"]), int(g["num_stations"])
    vis0 = g["vis0_%d" % matrix].astype(got.dtype)
    v, v0 = got.reshape(-1, nr), vis0.reshape(-1, nr)
    assert np.array_equal(v[:off], v0[:off])
    assert np.array_equal(v[off + ns:], v0[off + ns:])
    keep = [1, 7] if matrix else [1]
    assert np.array_equal(v[off:off + ns, keep], v0[off:off + ns, keep])


def complex_jones(j, ns, nsrc):
    j = j.reshape(ns, nsrc, 4, 2)
    return (j[..., 0] + 1j * j[..., 1]).reshape(ns, nsrc, 2, 2)


def test_acorr_oracle_matrix_algebra_brute_force():
    g = load_golden("acorr.npz")
    ns, nsrc = int(g["num_stations"]), int(g["num_sources"])
    J = complex_jones(g["jones_1"].astype(np.float64), ns, nsrc)
    I_, Q, U, V = (g[k].astype(np.float64) for k in "IQUV")
    B = np.zeros((nsrc, 2, 2), complex)
    B[:, 0, 0], B[:, 1, 1] = I_ + Q, I_ - Q
    B[:, 0, 1], B[:, 1, 0] = U + 1j * V, U - 1j * V
    want = np.einsum("sixy,iyz,siwz->sxw", J, B, np.conj(J))
    got = oracle.auto_correlate("double", 1, ns, g["jones_1"], I_, Q, U, V)
    got = got.reshape(ns, 4, 2)
    w = want.reshape(ns, 4)
    assert np.allclose(got[:, :, 0], w.real, rtol=1e-13, atol=0)
    assert np.allclose(got[:, [1, 2], 1], w.imag[:, [1, 2]], rtol=1e-12,
                       atol=1e-9)
    assert np.all(got[:, [0, 3], 1] == 0.0)
    # Scalar: sum |J|^2 I.
    js = g["jones_0"].astype(np.float64).reshape(ns, nsrc, 2)
    want_s = ((js ** 2).sum(-1) * I_).sum(1)
    got_s = oracle.auto_correlate("double", 0, ns, g["jones_0"], I_)
    assert np.allclose(got_s.reshape(ns, 2)[:, 0], want_s, rtol=1e-13)
    assert np.all(got_s.reshape(ns, 2)[:, 1] == 0.0)


def test_acorr_oracle_diagonal_jones_closed_form():
    rng = np.random.default_rng(11)
    ns, nsrc = 5, 64
    gx = rng.normal(size=ns) + 1j * rng.normal(size=ns)
    gy = rng.normal(size=ns) + 1j * rng.normal(size=ns)
    J = np.zeros((ns, nsrc, 8))
    J[:, :, 0], J[:, :, 1] = gx.real[:, None], gx.imag[:, None]
    J[:, :, 6], J[:, :, 7] = gy.real[:, None], gy.imag[:, None]
    I_ = rng.uniform(1, 2, nsrc)
    Q, U, V = (rng.uniform(-0.3, 0.3, nsrc) for _ in range(3))
    got = oracle.auto_correlate("double", 1, ns, J, I_, Q, U, V)
    got = got.reshape(ns, 4, 2)
    got = got[..., 0] + 1j * got[..., 1]
    sb = np.sum(U + 1j * V)
    assert np.allclose(got[:, 0], np.abs(gx) ** 2 * np.sum(I_ + Q), rtol=1e-13)
    assert np.allclose(got[:, 3], np.abs(gy) ** 2 * np.sum(I_ - Q), rtol=1e-13)
    assert np.allclose(got[:, 1], gx * np.conj(gy) * sb, rtol=1e-13)
    assert np.allclose(got[:, 2], gy * np.conj(gx) * np.conj(sb), rtol=1e-13)


@pytest.mark.parametrize("matrix", [0, 1])
def test_acorr_oracle_equals_cross_term_of_coincident_stations(matrix):
    g = load_golden("acorr.npz")
    ns, nsrc = 6, int(g["num_sources"])
    nr = 8 if matrix else 2
    jones = g["jones_%d" % matrix].astype(np.float64)[:ns * nsrc * nr]
    auto = oracle.auto_correlate("double", matrix, ns, jones, g["I"], g["Q"],
                                 g["U"], g["V"]).reshape(ns, nr)
    z = np.zeros(nsrc)
    for st in range(ns):
        j = jones.reshape(ns, nsrc * nr)[st]
        pair = np.concatenate([j, j])
        x = oracle.cross_correlate(
            "double", 0, 0, matrix, pair, g["I"], g["Q"], g["U"], g["V"],
            z, z, z + 1.0, z, z, z, np.zeros(2), np.zeros(2), np.zeros(2),
            np.zeros(2), np.zeros(2))
        want = x.reshape(nr).copy()
        if matrix:
            want[[1, 7]] = 0.0
        else:
            want[1] = 0.0
        assert max_rel(auto[st], want) < 1e-14
