"""GPU parity of the per-station factorised correlator
(oskar_cross_correlate_factorised_jones: one station beam per station,
V = A D A^H with A_p(s) = K_p(s) E_p(s) L_s) against the FP64 oracle of the
reference's direct algorithm on J_p(s) = K_p(s) E_p(s)
(oskar/correlate/src/oskar_cross_correlate_omp.cpp:33-156), through the
C-ABI.  Tolerance: relative RMS of all visibility reals vs FP64 on the same
rounded inputs <= 1e-5.

Covered: random per-station beams with polarised sources (|pol| <= 0.2 I);
sources whose brightness matrix is not positive semi-definite (|pol| > I,
I = Q = 0 with U, V != 0, I = -Q), which take the LDL^H / UDU^H / eigen
branches with negative signs; CASA, offset, uv filter and untouched
outputs; a Jones with more sources than used; ragged station and source
counts; part sizes; FP32 and FP64; equality with the shared-beam path when
every station has the same beam; argument checks.
"""
import numpy as np
import pytest

from fact_cases import (make_case, oracle_direct, oracle_direct_stations,
                        rounded, station_beams)
from gpu_util import gpu_factorised, gpu_factorised_jones
from helpers import rel_rms

pytestmark = pytest.mark.gpu
TOL = 1e-5


def run(c, E, prec, mode, freq, **kw):
    return gpu_factorised_jones(prec, mode, kw.pop("casa", 0), E, c["I"],
                                c["Q"], c["U"], c["V"], c["l"], c["m"],
                                c["n"], c["u"], c["v"], c["w"], freq, **kw)


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


@pytest.mark.parametrize("mode", [0, 1], ids=["fp16x3", "bf16x3"])
@pytest.mark.parametrize("shape", [(2, 5), (65, 3001), (130, 9001),
                                   (200, 20000)])
def test_jones_matches_fp64_oracle(hip, mode, shape):
    ns, nsrc = shape
    print("factorised jones", mode, shape, flush=True)
    c = make_case(ns + 3 * nsrc, ns, nsrc)
    E = station_beams(ns + nsrc, ns, nsrc)
    freq = 150e6
    for prec in ("f", "d"):
        cr = rounded(c, prec)
        Er = f32(E) if prec == "f" else E
        want = oracle_direct_stations(cr, Er, freq)
        got = run(c, E, prec, mode, freq)
        err = rel_rms(got, want)
        print("  prec %s rel RMS %.3e" % (prec, err), flush=True)
        assert err < TOL


def test_jones_indefinite_brightness(hip):
    """Sources with |pol| > I, with I = Q = 0 (zero diagonal), and with
    I = -Q (a = 0): B is not positive semi-definite, so D carries -1
    signs; the result still equals the direct algorithm's."""
    ns, nsrc = 40, 2003
    c = make_case(5, ns, nsrc)
    rng = np.random.default_rng(6)
    k = nsrc // 4
    c["Q"] = c["Q"].copy(); c["U"] = c["U"].copy(); c["V"] = c["V"].copy()
    c["I"] = c["I"].copy()
    c["Q"][:k] = 1.5 * c["I"][:k] * rng.choice([-1, 1], k)       # |pol| > I
    c["I"][k:2 * k] = 0.0                                          # a = d = 0
    c["Q"][k:2 * k] = 0.0
    c["U"][k:2 * k] = rng.normal(size=k)
    c["V"][k:2 * k] = rng.normal(size=k)
    c["Q"][2 * k:2 * k + 20] = -c["I"][2 * k:2 * k + 20]          # a = 0
    c["U"][2 * k + 20:2 * k + 40] = 3.0 * c["I"][2 * k + 20:2 * k + 40]
    E = station_beams(7, ns, nsrc)
    freq = 120e6
    for mode in (0, 1):
        want = oracle_direct_stations(c, E, freq)
        got = run(c, E, "d", mode, freq)
        err = rel_rms(got, want)
        print("indefinite B, mode %d: rel RMS %.3e" % (mode, err), flush=True)
        assert err < TOL


def test_jones_casa_offset_uvfilter_ld(hip):
    """CASA conjugate transpose, offset_out, the uv filter (untouched rows),
    and a Jones holding more sources than num_sources."""
    ns, nsrc, extra = 9, 4001, 37
    c = make_case(9, ns, nsrc + extra, pol=True)
    E = station_beams(10, ns, nsrc + extra)
    freq = 80e6
    nb = ns * (ns - 1) // 2
    rng = np.random.default_rng(1)
    init = rng.normal(size=3 * nb * 8)
    got = run(c, E, "d", 0, freq, casa=1, vis_init=init, offset_out=nb,
              uv_filter=(2000.0, 15000.0), num_sources=nsrc)
    cs = {k: (v[:nsrc] if k in ("l", "m", "n", "I", "Q", "U", "V") else v)
          for k, v in c.items() if k != "E"}
    want = oracle_direct_stations(cs, E[:, :nsrc], freq, casa=1,
                                  vis=init.copy(), offset_out=nb,
                                  uv_min=2000.0, uv_max=15000.0)
    assert np.array_equal(got[:8 * nb], init[:8 * nb])
    assert np.array_equal(got[16 * nb:], init[16 * nb:])
    d_got = got[8 * nb:16 * nb] - init[8 * nb:16 * nb]
    d_want = want[8 * nb:16 * nb] - init[8 * nb:16 * nb]
    assert np.count_nonzero(d_want) < d_want.size      # some filtered
    assert np.array_equal(d_got == 0, d_want == 0)
    assert rel_rms(d_got, d_want) < TOL


def test_jones_part_sizes(hip):
    from oskar_amd import lib
    ns, nsrc = 70, 10001
    c = make_case(11, ns, nsrc)
    E = station_beams(12, ns, nsrc)
    freq = 200e6
    want = oracle_direct_stations(rounded(c, "f"), f32(E), freq)
    for part in (4, 512, 1000, 2048):
        old = lib.oskar_cross_correlate_factorised_jones_set_part_size(part)
        try:
            got = run(c, E, "f", 0, freq)
        finally:
            lib.oskar_cross_correlate_factorised_jones_set_part_size(old)
        err = rel_rms(got, want)
        print("part %d rel RMS %.3e" % (part, err), flush=True)
        assert err < TOL


def test_jones_identical_beams_equal_shared_path(hip):
    """Every station with the same beam: both factorised forms compute the
    same visibilities (each within its tolerance of FP64)."""
    ns, nsrc = 96, 6000
    c = make_case(13, ns, nsrc)
    E = np.broadcast_to(c["E"], (ns, nsrc, 8)).copy()
    freq = 110e6
    a = run(c, E, "d", 0, freq)
    b = gpu_factorised("d", 0, 0, c["E"].ravel(), c["I"], c["Q"], c["U"],
                       c["V"], c["l"], c["m"], c["n"], c["u"], c["v"],
                       c["w"], freq)
    want = oracle_direct(c, freq)
    assert rel_rms(a, want) < TOL and rel_rms(b, want) < TOL
    assert rel_rms(a, b) < 2 * TOL


def test_jones_argument_checks(hip):
    ns, nsrc = 6, 100
    c = make_case(14, ns, nsrc)
    E = station_beams(15, ns + 1, nsrc)      # one station too many
    run(c, E, "d", 0, 1e8, expect_status=-11)   # DIMENSION_MISMATCH
    E = station_beams(15, ns, nsrc - 1)      # too few sources
    run(c, E, "d", 0, 1e8, expect_status=-11)
    E = station_beams(15, ns, nsrc)
    run(c, E, "d", 5, 1e8, expect_status=-3)    # bad mode
