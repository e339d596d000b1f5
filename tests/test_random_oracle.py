"""The oracle's random draws for time-variable element errors.

  - Philox4x32-10 against the published known-answer vectors of the
    Random123 distribution (counter and key all zero, all ones, and the
    hexadecimal digits of pi);
  - oskar_mem_random_gaussian's layout and Box-Muller: value j from Philox
    counter j / 4 (checked against a numpy restatement of the Box-Muller
    formula on the Philox outputs), mean 0 and variance 1 over 1M draws
    (five standard errors), no correlation between neighbouring draws;
  - element error factors with zero error scales are gain (cos, sin)(phase)
    to within a rounding.
"""
import math

import numpy as np

import oracle


def test_philox_known_answers():
    cases = [
        ([0, 0, 0, 0], [0, 0],
         [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2,
         [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
         [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, want in cases:
        assert list(oracle.philox4x32_10(ctr, key)) == want


def test_gaussian_layout_and_box_muller():
    seed, c1, c2, c3 = 42, 3, 7, 0x12345678
    g = oracle.random_gaussian(10, seed, c1, c2, c3)
    fa, fb = 1.0 / 2 ** 31, 1.0 / 2 ** 32
    for j in range(10):
        u = oracle.philox4x32_10([j // 4, c1, c2, c3], [seed, 0xCAFEF00D])
        a, b = (u[0], u[1]) if j % 4 < 2 else (u[2], u[3])
        t = math.pi * (int(a) * fa + 0.5 * fa)
        r = math.sqrt(-2.0 * math.log(int(b) * fb + 0.5 * fb))
        want = (math.sin(t) if j % 2 == 0 else math.cos(t)) * r
        assert abs(g[j] - want) <= 1e-15 * max(1.0, abs(want))


def test_gaussian_statistics():
    g = oracle.random_gaussian(1_000_000, 1, 0, 0, 0x12345678)
    n = g.size
    assert abs(g.mean()) < 5 / math.sqrt(n)
    assert abs(g.var() - 1) < 5 * math.sqrt(2 / n)
    assert abs(np.corrcoef(g[:-1], g[1:])[0, 1]) < 5 / math.sqrt(n)


def test_element_errors_without_error_scales():
    rng = np.random.default_rng(5)
    n = 33
    gain = rng.uniform(0.5, 1.5, n)
    phase = rng.uniform(-1, 1, n)
    z = np.zeros(n)
    err = oracle.element_errors(9, 2, 4, gain, z, phase, z).reshape(-1, 2)
    # Within one rounding of libm's separate cos and sin (the compiler may
    # fuse the oracle's pair into one sincos call).
    np.testing.assert_allclose(
        err[:, 0], [math.cos(p) * g for p, g in zip(phase, gain)],
        rtol=3e-16, atol=0)
    np.testing.assert_allclose(
        err[:, 1], [math.sin(p) * g for p, g in zip(phase, gain)],
        rtol=3e-16, atol=0)
