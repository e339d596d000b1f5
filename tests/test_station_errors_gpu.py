"""Element gains, phase offsets and apodisation weights in the aperture-array
station beam (oskar_station_evaluate_element_weights.c:34-64), through the
C-ABI:

  - random gains, phase offsets and complex apodisation weights: the beam
    equals the oracle's (double max rel < 1e-10, single RMS < 2e-5);
  - closed forms that share no code with the oracle:
      * phase offsets phi_i = -k x_i . (s1 - s0) on a station pointed at s0
        give the beam of the same station pointed at s1;
      * a common gain g and a common apodisation weight a multiply the
        array pattern by g a;
  - oskar_station_analyse sets the apply flags only for non-default values;
    a feed other than 0 or 1, or an element index out of range, is refused;
  - time-variable random errors (gain and phase error scales, the draws of
    the reference's Philox generator keyed by the station's seed, time
    index and unique id): the beam equals the oracle's, is the same for the
    same key and differs for another time index.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from gpu_util import sync_check
from helpers import max_rel, rel_rms
from oskar_amd import lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_DOUBLE, OSKAR_GPU,
                           OSKAR_MATRIX, OSKAR_SINGLE, Mem, Station,
                           StationWork, check_status)

pytestmark = pytest.mark.gpu
C0 = 299792458.0
AZEL = 4
FREQ = 110e6


def make(prec, ne, rng):
    P = OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE
    st = Station(P, OSKAR_GPU, ne)
    r = 19.0 * np.sqrt(rng.uniform(size=ne))
    t = rng.uniform(0, 2 * np.pi, ne)
    xt = np.stack([r * np.cos(t), r * np.sin(t), np.zeros(ne)])
    xm = xt + np.concatenate([rng.normal(scale=0.01, size=(2, ne)),
                              np.zeros((1, ne))])
    st.set_coords(xt, xm)
    return st, xt, xm


def analyse(st):
    s, fin = C.c_int(0), C.c_int(0)
    lib.oskar_station_analyse(st.ptr, C.byref(fin), C.byref(s))
    check_status(s, "oskar_station_analyse")


def directions(rng, n=1500):
    th = np.arccos(rng.uniform(0.0, 1.0, n))
    phi = rng.uniform(0, 2 * np.pi, n)
    return np.sin(th) * np.sin(phi), np.sin(th) * np.cos(phi), np.cos(th)


def beam(prec, st, x, y, z, expect_status=0):
    P = OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE
    mx, my, mz = (Mem.from_numpy(a, P) for a in (x, y, z))
    out = Mem(P | OSKAR_COMPLEX | OSKAR_MATRIX, OSKAR_GPU, x.size)
    work = StationWork(P, OSKAR_GPU)
    s = C.c_int(0)
    lib.oskar_evaluate_station_beam_aperture_array(
        st.ptr, work.ptr, x.size, mx.ptr, my.ptr, mz.ptr, 0, 0.0, FREQ,
        out.ptr, C.byref(s))
    if expect_status:
        assert s.value == expect_status
        return None
    check_status(s, "station beam")
    sync_check("station beam")
    return out.numpy().astype(np.float64)


def azel_dir(az, el):
    return np.array([math.cos(el) * math.sin(az), math.cos(el) * math.cos(az),
                     math.sin(el)])


@pytest.mark.parametrize("prec", ["f", "d"])
def test_errors_and_apodisation_match_oracle(hip, prec):
    print("element errors + apodisation", prec, flush=True)
    rng = np.random.default_rng(41)
    ne = 64
    st, xt, xm = make(prec, ne, rng)
    gain = rng.uniform(0.5, 1.5, ne)
    ph_deg = rng.uniform(-30, 30, ne)
    apod = rng.uniform(0.2, 1.0, ne) * np.exp(1j * rng.uniform(-0.3, 0.3, ne))
    s = C.c_int(0)
    for i in range(ne):
        lib.oskar_station_set_element_errors(st.ptr, 0, i, gain[i], 0.0,
                                             ph_deg[i], 0.0, C.byref(s))
        lib.oskar_station_set_element_weight(st.ptr, 0, i, apod[i].real,
                                             apod[i].imag, C.byref(s))
    check_status(s, "set element errors and weights")
    analyse(st)
    assert lib.oskar_station_apply_element_errors(st.ptr) == 1
    assert lib.oskar_station_apply_element_weight(st.ptr) == 1
    az, el = math.radians(30), math.radians(60)
    lib.oskar_station_set_phase_centre(st.ptr, AZEL, az, el)
    x, y, z = directions(rng)
    got = beam(prec, st, x, y, z)
    f = (lambda a: np.asarray(a, np.float32).astype(np.float64)) \
        if prec == "f" else (lambda a: np.asarray(a, np.float64))
    ph = np.radians(ph_deg)
    ap = np.stack([apod.real, apod.imag], 1).ravel()
    z0 = np.zeros(ne)
    err = oracle.element_errors(1, 0, 0, f(gain), z0, f(ph), z0)
    want = oracle.station_beam_aa(
        1, *[f(v) for v in xt], *[f(v) for v in xm], np.zeros(ne),
        np.zeros(ne, np.int32), np.array([0], np.int32),
        np.array([0.5 * C0 / FREQ]), azel_dir(az, el), f(x), f(y), f(z),
        FREQ, err=err, apod=f(ap))
    if prec == "d":
        assert max_rel(got, want) < 1e-10
    else:
        assert rel_rms(got, want) < 2e-5


def test_phase_offsets_steer_the_beam(hip):
    rng = np.random.default_rng(42)
    ne = 64
    st, xt, xm = make("d", ne, rng)
    ref, _, _ = make("d", ne, np.random.default_rng(42))
    s0 = (math.radians(30), math.radians(60))
    s1 = (math.radians(75), math.radians(45))
    k = 2 * math.pi * FREQ / C0
    d = azel_dir(*s1) - azel_dir(*s0)
    phi = -k * (xm[0] * d[0] + xm[1] * d[1] + xm[2] * d[2])
    s = C.c_int(0)
    for i in range(ne):
        lib.oskar_station_set_element_errors(st.ptr, 0, i, 1.0, 0.0,
                                             math.degrees(phi[i]), 0.0,
                                             C.byref(s))
    check_status(s, "set element errors")
    analyse(st)
    analyse(ref)
    lib.oskar_station_set_phase_centre(st.ptr, AZEL, *s0)
    lib.oskar_station_set_phase_centre(ref.ptr, AZEL, *s1)
    x, y, z = directions(rng)
    got = beam("d", st, x, y, z)
    want = beam("d", ref, x, y, z)
    assert max_rel(got, want) < 1e-10


def test_common_gain_and_apodisation_scale_the_pattern(hip):
    rng = np.random.default_rng(43)
    ne = 64
    st, _, _ = make("d", ne, rng)
    plain, _, _ = make("d", ne, np.random.default_rng(43))
    g, a = 2.0, 0.5 + 0.25j
    s = C.c_int(0)
    for i in range(ne):
        lib.oskar_station_set_element_errors(st.ptr, 0, i, g, 0.0, 0.0, 0.0,
                                             C.byref(s))
        lib.oskar_station_set_element_weight(st.ptr, 0, i, a.real, a.imag,
                                             C.byref(s))
    check_status(s, "set element errors and weights")
    analyse(st)
    analyse(plain)
    for t in (st, plain):
        lib.oskar_station_set_phase_centre(t.ptr, AZEL, math.radians(10),
                                           math.radians(70))
    x, y, z = directions(rng)
    got = beam("d", st, x, y, z).reshape(-1, 2)
    base = beam("d", plain, x, y, z).reshape(-1, 2)
    want = (base[:, 0] + 1j * base[:, 1]) * (g * a)
    np.testing.assert_allclose(got[:, 0] + 1j * got[:, 1], want, rtol=0,
                               atol=1e-12 * np.abs(want).max())


def test_analyse_flags_and_argument_checks(hip):
    rng = np.random.default_rng(44)
    st, _, _ = make("d", 16, rng)
    s = C.c_int(0)
    analyse(st)
    assert lib.oskar_station_apply_element_errors(st.ptr) == 0
    assert lib.oskar_station_apply_element_weight(st.ptr) == 0
    # Default values set explicitly change nothing.
    lib.oskar_station_set_element_errors(st.ptr, 0, 3, 1.0, 0.0, 0.0, 0.0,
                                         C.byref(s))
    lib.oskar_station_set_element_weight(st.ptr, 0, 3, 1.0, 0.0, C.byref(s))
    analyse(st)
    assert lib.oskar_station_apply_element_errors(st.ptr) == 0
    assert lib.oskar_station_apply_element_weight(st.ptr) == 0
    # Feed 2 and an index out of range are rejected (OSKAR_ERR_OUT_OF_RANGE).
    lib.oskar_station_set_element_weight(st.ptr, 2, 0, 1.0, 0.0, C.byref(s))
    assert s.value == -5
    s = C.c_int(0)
    lib.oskar_station_set_element_weight(st.ptr, 0, 16, 1.0, 0.0, C.byref(s))
    assert s.value == -5


def beam_t(prec, st, x, y, z, time_index):
    P = OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE
    mx, my, mz = (Mem.from_numpy(a, P) for a in (x, y, z))
    out = Mem(P | OSKAR_COMPLEX | OSKAR_MATRIX, OSKAR_GPU, x.size)
    work = StationWork(P, OSKAR_GPU)
    s = C.c_int(0)
    lib.oskar_evaluate_station_beam_aperture_array(
        st.ptr, work.ptr, x.size, mx.ptr, my.ptr, mz.ptr, int(time_index),
        0.0, FREQ, out.ptr, C.byref(s))
    check_status(s, "station beam")
    sync_check("station beam")
    return out.numpy().astype(np.float64)


@pytest.mark.parametrize("prec", ["f", "d"])
def test_random_element_errors_match_oracle(hip, prec):
    print("random element errors", prec, flush=True)
    rng = np.random.default_rng(45)
    ne = 70                      # odd number of Philox pairs
    st, xt, xm = make(prec, ne, rng)
    gain = rng.uniform(0.8, 1.2, ne)
    gerr = rng.uniform(0.0, 0.1, ne)
    ph_deg = rng.uniform(-10, 10, ne)
    perr_deg = rng.uniform(0.0, 5.0, ne)
    s = C.c_int(0)
    for i in range(ne):
        lib.oskar_station_set_element_errors(st.ptr, 0, i, gain[i], gerr[i],
                                             ph_deg[i], perr_deg[i],
                                             C.byref(s))
    check_status(s, "set element errors")
    lib.oskar_station_set_seed_time_variable_errors(st.ptr, 42)
    counter = C.c_int(7)
    lib.oskar_station_set_unique_ids(st.ptr, C.byref(counter))
    assert counter.value == 8 and lib.oskar_station_unique_id(st.ptr) == 7
    analyse(st)
    assert lib.oskar_station_apply_element_errors(st.ptr) == 1
    az, el = math.radians(200), math.radians(50)
    lib.oskar_station_set_phase_centre(st.ptr, AZEL, az, el)
    x, y, z = directions(rng, 1200)
    f = (lambda a: np.asarray(a, np.float32).astype(np.float64)) \
        if prec == "f" else (lambda a: np.asarray(a, np.float64))
    got3 = beam_t(prec, st, x, y, z, 3)
    err = oracle.element_errors(42, 3, 7, f(gain), f(gerr),
                                f(np.radians(ph_deg)),
                                f(np.radians(perr_deg)))
    want = oracle.station_beam_aa(
        1, *[f(v) for v in xt], *[f(v) for v in xm], np.zeros(ne),
        np.zeros(ne, np.int32), np.array([0], np.int32),
        np.array([0.5 * C0 / FREQ]), azel_dir(az, el), f(x), f(y), f(z),
        FREQ, err=err)
    if prec == "d":
        assert max_rel(got3, want) < 1e-10
    else:
        assert rel_rms(got3, want) < 2e-5
    # The same key gives the same beam; another time index another one.
    assert np.array_equal(beam_t(prec, st, x, y, z, 3), got3)
    got4 = beam_t(prec, st, x, y, z, 4)
    assert rel_rms(got4, got3) > 1e-3
