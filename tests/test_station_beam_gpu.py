"""GPU tests of oskar_convert_any_to_enu_directions and oskar_station_beam.

Conversions, pinned by geometry rather than by the formulas:
- RA/Dec: unit vectors; the celestial pole maps to (0, cos lat, sin lat)
  at any sidereal time; a source on the meridian (RA = LST) has x = 0 and
  elevation 90 deg - |lat - dec|;
- az/el: az = atan2(x, y), el = asin(z) recover the input;
- relative directions (l, m, n) about a reference point: the origin
  (0, 0, 1) maps to the reference point's RA/Dec direction, and the map is a
  rotation (dot products of directions are preserved).
Station beam (FP64): an aperture-array station's beam of ENU directions,
divided by its value in the normalisation direction, against the CPU
oracle of the aperture-array beam (max relative error 1e-10); RA/Dec input
against the same beam of the converted directions; normalisation off
equals oskar_evaluate_station_beam_aperture_array; isotropic stations give
the identity; a beam direction below the horizon gives zeros.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from gpu_util import sync_check
from helpers import max_rel
from oskar_amd import capi, lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_DOUBLE, OSKAR_GPU,
                           OSKAR_MATRIX, Mem, Station, StationWork,
                           check_status)

pytestmark = pytest.mark.gpu
REL_DIR, ENU_DIR, RADEC, AZEL = 0, 1, 2, 4
C0 = 299792458.0


def to_enu(kind, a, b, c=None, ref=(0.0, 0.0), lst=0.0, lat=0.0):
    P = OSKAR_DOUBLE
    ins = [Mem.from_numpy(x, P) for x in (a, b, c if c is not None else a)]
    outs = [Mem(P, OSKAR_GPU, 1) for _ in range(3)]
    s = C.c_int(0)
    lib.oskar_convert_any_to_enu_directions(
        kind, a.size, capi.MEM3(*[m.ptr for m in ins]), float(ref[0]),
        float(ref[1]), float(lst), float(lat),
        capi.MEM3(*[m.ptr for m in outs]), C.byref(s))
    check_status(s, "convert")
    sync_check("convert")
    return [m.numpy()[:a.size] for m in outs]


def test_radec_to_enu_geometry(hip):
    rng = np.random.default_rng(1)
    lat, lst = np.radians(-26.7), 1.234
    ra = rng.uniform(0, 2 * np.pi, 500)
    dec = rng.uniform(-np.pi / 2, np.pi / 2, 500)
    x, y, z = to_enu(RADEC, ra, dec, lst=lst, lat=lat)
    assert np.allclose(x * x + y * y + z * z, 1.0, atol=1e-14)
    # The south celestial pole from a southern site, at any sidereal time.
    for t in (0.0, 2.0, 5.5):
        x, y, z = to_enu(RADEC, np.array([0.7, 3.1]),
                         np.array([-np.pi / 2, -np.pi / 2]), lst=t, lat=lat)
        assert np.allclose(x, 0.0, atol=1e-15)
        assert np.allclose(y, -np.cos(lat), atol=1e-15)
        assert np.allclose(z, -np.sin(lat), atol=1e-15)
    # On the meridian: no east component, elevation 90 - |lat - dec|.
    dec = np.radians(np.array([-80.0, -40.0, 0.0, 30.0]))
    x, y, z = to_enu(RADEC, np.full(4, lst), dec, lst=lst, lat=lat)
    assert np.allclose(x, 0.0, atol=1e-15)
    el = np.degrees(np.arcsin(z))
    assert np.allclose(el, 90.0 - np.abs(np.degrees(lat - dec)), atol=1e-9)


def test_azel_round_trip(hip):
    rng = np.random.default_rng(2)
    az = rng.uniform(-np.pi, np.pi, 300)
    el = rng.uniform(-np.pi / 2 + 1e-3, np.pi / 2 - 1e-3, 300)
    x, y, z = to_enu(AZEL, az, el)
    assert np.allclose(np.arctan2(x, y), az, atol=1e-12)
    assert np.allclose(np.arcsin(z), el, atol=1e-12)


def test_relative_directions_are_a_rotation(hip):
    rng = np.random.default_rng(3)
    ra0, dec0, lst, lat = 0.9, np.radians(-35.0), 2.2, np.radians(-30.7)
    th = np.arccos(rng.uniform(-1, 1, 200))
    ph = rng.uniform(0, 2 * np.pi, 200)
    l = np.sin(th) * np.cos(ph)
    m = np.sin(th) * np.sin(ph)
    n = np.cos(th)
    x, y, z = to_enu(REL_DIR, np.r_[0.0, l], np.r_[0.0, m], np.r_[1.0, n],
                     ref=(ra0, dec0), lst=lst, lat=lat)
    x0, y0, z0 = to_enu(RADEC, np.array([ra0]), np.array([dec0]), lst=lst,
                        lat=lat)
    assert np.allclose([x[0], y[0], z[0]], [x0[0], y0[0], z0[0]], atol=1e-14)
    a = np.stack([l, m, n], 1)
    b = np.stack([x[1:], y[1:], z[1:]], 1)
    assert np.allclose(a @ a.T, b @ b.T, atol=1e-13)


def make_aa_station(rng, ne=48):
    st = Station(OSKAR_DOUBLE, OSKAR_GPU, ne)
    r = 15.0 * np.sqrt(rng.uniform(size=ne))
    t = rng.uniform(0, 2 * np.pi, ne)
    xt = np.stack([r * np.cos(t), r * np.sin(t), np.zeros(ne)])
    st.set_coords(xt, xt)
    s = C.c_int(0)
    lib.oskar_station_resize_element_types(st.ptr, 1, C.byref(s))
    lib.oskar_element_set_element_type(lib.oskar_station_element(st.ptr, 0),
                                       b"Dipole", C.byref(s))
    check_status(s, "station")
    lib.oskar_station_set_position(st.ptr, np.radians(116.7),
                                   np.radians(-26.7), 0.0, 0.0, 0.0, 0.0)
    return st, xt


def station_beam(st, work, kind, n, coords, norm, mjd, freq, ref=(0, 0),
                 offset_out=0, beam=None):
    P = OSKAR_DOUBLE
    ins = [Mem.from_numpy(c, P) for c in coords]
    if beam is None:
        beam = Mem(P | OSKAR_COMPLEX | OSKAR_MATRIX, OSKAR_GPU,
                   n + offset_out)
    s = C.c_int(0)
    lib.oskar_station_beam(
        st.ptr, work.ptr, kind, n, capi.MEM3(*[m.ptr for m in ins]),
        float(ref[0]), float(ref[1]), norm[0], float(norm[1]),
        float(norm[2]), 0, float(mjd), float(mjd), float(freq),
        offset_out, beam.ptr, C.byref(s))
    check_status(s, "oskar_station_beam")
    sync_check("oskar_station_beam")
    return beam.numpy()


def oracle_aa(xt, bd, x, y, z, freq):
    ne = xt.shape[1]
    return oracle.station_beam_aa(
        1, xt[0], xt[1], xt[2], xt[0], xt[1], xt[2], np.zeros(ne),
        np.zeros(ne, np.int32), np.array([0], np.int32),
        np.array([0.5 * C0 / freq]), bd, x, y, z, freq)


def test_station_beam_normalised_matches_oracle(hip):
    print("station beam normalised", flush=True)
    rng = np.random.default_rng(4)
    st, xt = make_aa_station(rng)
    az, el = np.radians(40.0), np.radians(65.0)
    lib.oskar_station_set_phase_centre(st.ptr, AZEL, az, el)
    work = StationWork(OSKAR_DOUBLE, OSKAR_GPU)
    n, freq, mjd = 700, 120e6, 59000.3
    th = np.arccos(rng.uniform(0.05, 1.0, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    x, y, z = np.sin(th) * np.sin(ph), np.sin(th) * np.cos(ph), np.cos(th)
    norm = (AZEL, az, el)                   # normalise at the beam direction
    got = station_beam(st, work, ENU_DIR, n, (x, y, z), norm, mjd, freq,
                       offset_out=3)
    bd = np.array([np.cos(el) * np.sin(az), np.cos(el) * np.cos(az),
                   np.sin(el)])
    want = oracle_aa(xt, bd, x, y, z, freq).reshape(n, 8)
    ref = oracle_aa(xt, bd, bd[:1], bd[1:2], bd[2:], freq)
    scale = 1.0 / np.sqrt(np.sum(ref ** 2) / 2.0)
    assert max_rel(got.reshape(-1, 8)[3:], want * scale) < 1e-10
    # Normalisation off: the aperture-array beam itself.
    lib.oskar_station_set_normalise_final_beam(st.ptr, 0)
    got = station_beam(st, work, ENU_DIR, n, (x, y, z), norm, mjd, freq)
    assert max_rel(got.reshape(n, 8), want) < 1e-10


def test_station_beam_radec_sources(hip):
    print("station beam radec", flush=True)
    rng = np.random.default_rng(5)
    st, xt = make_aa_station(rng)
    ra0, dec0 = np.radians(201.4), np.radians(-43.0)
    lib.oskar_station_set_phase_centre(st.ptr, RADEC, ra0, dec0)
    work = StationWork(OSKAR_DOUBLE, OSKAR_GPU)
    n, freq, mjd = 400, 90e6, 60123.41
    ra = ra0 + rng.uniform(-0.3, 0.3, n)
    dec = dec0 + rng.uniform(-0.3, 0.3, n)
    norm = (RADEC, ra0, dec0)
    got = station_beam(st, work, RADEC, n, (ra, dec, dec), norm, mjd, freq)
    lst = lib.oskar_convert_mjd_to_gast_fast(mjd) + np.radians(116.7)
    x, y, z = to_enu(RADEC, ra, dec, lst=lst, lat=np.radians(-26.7))
    want = station_beam(st, work, ENU_DIR, n, (x, y, z), norm, mjd, freq)
    assert max_rel(got, want) < 1e-12


def test_station_beam_isotropic_and_below_horizon(hip):
    print("station beam isotropic / below horizon", flush=True)
    rng = np.random.default_rng(6)
    st, xt = make_aa_station(rng)
    work = StationWork(OSKAR_DOUBLE, OSKAR_GPU)
    n = 50
    x, y, z = rng.normal(size=(3, n))
    lib.oskar_station_set_phase_centre(st.ptr, AZEL, 0.3, 0.9)
    lib.oskar_station_set_station_type(st.ptr, 1)          # isotropic
    got = station_beam(st, work, ENU_DIR, n, (x, y, z), (AZEL, 0.3, 0.9),
                       58000.0, 1e8).reshape(n, 8)
    assert np.all(got == np.array([1, 0, 0, 0, 0, 0, 1, 0]))
    lib.oskar_station_set_station_type(st.ptr, 0)
    lib.oskar_station_set_phase_centre(st.ptr, AZEL, 0.3, -0.1)
    init = Mem.from_numpy(np.ones(n * 8), OSKAR_DOUBLE | OSKAR_COMPLEX |
                          OSKAR_MATRIX)
    got = station_beam(st, work, ENU_DIR, n, (x, y, z), (AZEL, 0.3, 0.9),
                       58000.0, 1e8, beam=init)
    assert not np.any(got)
