"""GPU parity of the station-factorised MFMA correlator against the FP64
CPU oracle of the reference's direct algorithm, through the C-ABI.

The oracle forms J_p(s) = K_p(s) E(s) (oskar_evaluate_jones_K semantics)
and runs the reference's double-precision cross-correlation on it.
Tolerance (relative RMS of all visibility reals vs FP64 on the same
rounded inputs): 1e-5 for FP16x3 and BF16x3.  Measured on an MI355X:
FP16x3 1.5e-7 on 5 sources, 3e-6 .. 6.2e-6 on 3,000 - 20,000 sources (the
error grows with the sources summed in FP32 per part and with the sky's
dynamic range); BF16x3 3.9e-6 .. 8.6e-6.
"""
import numpy as np
import pytest

from fact_cases import make_case, oracle_direct, rounded
from gpu_util import gpu_factorised
from helpers import max_rel, rel_rms

pytestmark = pytest.mark.gpu
TOL = {0: 1e-5, 1: 1e-5}


def run(c, prec, mode, freq, casa=0, **kw):
    return gpu_factorised(prec, mode, casa,
                          None if c["E"] is None else c["E"].ravel(),
                          c["I"], c["Q"], c["U"], c["V"], c["l"], c["m"],
                          c["n"], c["u"], c["v"], c["w"], freq, **kw)


@pytest.mark.parametrize("mode", [0, 1], ids=["fp16x3", "bf16x3"])
@pytest.mark.parametrize("shape", [(2, 5), (64, 3000), (130, 9001),
                                   (200, 20000)])
def test_factorised_matches_fp64_oracle(hip, mode, shape):
    ns, nsrc = shape
    print("factorised", mode, shape, flush=True)
    c = make_case(ns + nsrc, ns, nsrc)
    freq = 150e6
    for prec in ("f", "d"):
        want = oracle_direct(rounded(c, prec), freq)
        got = run(c, prec, mode, freq)
        err = rel_rms(got, want)
        print("  prec %s rel RMS %.3e  max rel %.3e" % (
            prec, err, max_rel(got, want)), flush=True)
        assert err < TOL[mode]


def test_factorised_casa_offset_uvfilter_and_identity_beam(hip):
    c = make_case(9, 70, 4001, pol=True, beam=False)
    freq = 80e6
    ns = c["u"].size
    nb = ns * (ns - 1) // 2
    rng = np.random.default_rng(1)
    init = rng.normal(size=3 * nb * 8)
    got = run(c, "d", 0, freq, casa=1, vis_init=init, offset_out=nb,
              uv_filter=(2000.0, 15000.0))
    want = oracle_direct(c, freq, casa=1, vis=init.copy(), offset_out=nb,
                         uv_min=2000.0, uv_max=15000.0)
    assert np.array_equal(got[:8 * nb], init[:8 * nb])
    assert np.array_equal(got[16 * nb:], init[16 * nb:])
    assert rel_rms(got[8 * nb:16 * nb] - init[8 * nb:16 * nb],
                   want[8 * nb:16 * nb] - init[8 * nb:16 * nb]) < 1e-5


def test_factorised_part_size_does_not_change_result_beyond_tolerance(hip):
    from oskar_amd import lib
    c = make_case(4, 96, 12000)
    freq = 200e6
    want = oracle_direct(rounded(c, "f"), freq)
    old = lib.oskar_cross_correlate_factorised_set_part_size(1024)
    try:
        a = run(c, "f", 0, freq)
    finally:
        lib.oskar_cross_correlate_factorised_set_part_size(old)
    b = run(c, "f", 0, freq)
    assert rel_rms(a, want) < 1e-5 and rel_rms(b, want) < 1e-5


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("mode", [0, 1], ids=["fp16x3", "bf16x3"])
def test_factorised_every_tile_shape(hip, shape, mode):
    """The one-call path with each tile shape (ragged station count, source
    count off the k-step and part sizes, FP32 and FP64)."""
    from oskar_amd import lib
    print("factorised shape", shape, "mode", mode, flush=True)
    c = make_case(77 + shape, 150, 9001)
    freq = 180e6
    old = lib.oskar_cross_correlate_factorised_set_tile_shape(shape)
    try:
        for prec in ("f", "d"):
            want = oracle_direct(rounded(c, prec), freq)
            got = run(c, prec, mode, freq)
            err = rel_rms(got, want)
            print("  prec %s rel RMS %.3e" % (prec, err), flush=True)
            assert err < TOL[mode]
    finally:
        lib.oskar_cross_correlate_factorised_set_tile_shape(old)
    assert lib.oskar_cross_correlate_factorised_set_tile_shape(9) == -1
