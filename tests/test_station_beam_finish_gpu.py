"""GPU tests of the last steps of oskar_station_beam: the horizon blank and
the normalisation.  For a flat station of one dipole type one kernel forms
the whole beam with them (array factor, element pattern, blank, division
by the beam in the normalisation direction, whose array factor it sums in
a different order); other stations go through the separate steps and one
kernel blanks and normalises while it copies the beam out.

The reference blanks the aperture-array beam below the horizon
(oskar_evaluate_station_beam_aperture_array.c:238-239, flat stations only)
and then divides it by its value at the normalisation direction
(oskar_station_beam.c:204-216, oskar_mem_normalise).  Expected values here
come from oskar_evaluate_station_beam_aperture_array itself, evaluated on
the same directions with the normalisation direction appended, blanked by
that function, then normalised in numpy:
  - flat stations, FP64 and FP32, matrix and scalar beams, normalisation on
    and off, about a third of the directions below the horizon (their rows
    must be exact zeros), with an output offset; elements in a plane and
    (FP32 and FP64 matrix) with heights, a 3D array;
  - element gains and phases with time-variable random errors and
    apodisation weights, at a time index other than 0;
  - a hierarchical station (alike children), whose top level is not
    blanked (its children are), normalisation on.
"""
import ctypes as C

import numpy as np
import pytest

from gpu_util import sync_check
from helpers import max_rel
from oskar_amd import capi, lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_DOUBLE, OSKAR_GPU,
                           OSKAR_MATRIX, OSKAR_SINGLE, Mem, Station,
                           StationView, StationWork, check_status)

pytestmark = pytest.mark.gpu
ENU_DIR, AZEL = 1, 4
FREQ, MJD = 110e6, 59500.25
AZ, EL = np.radians(35.0), np.radians(70.0)


def azel_dir(az, el):
    return np.array([np.cos(el) * np.sin(az), np.cos(el) * np.cos(az),
                     np.sin(el)])


def flat_station(P, rng, ne=40, height=0.0):
    st = Station(P, OSKAR_GPU, ne)
    r = 12.0 * np.sqrt(rng.uniform(size=ne))
    t = rng.uniform(0, 2 * np.pi, ne)
    xt = np.stack([r * np.cos(t), r * np.sin(t),
                   height * rng.uniform(-1.0, 1.0, ne)])
    st.set_coords(xt, xt)
    s = C.c_int(0)
    lib.oskar_station_resize_element_types(st.ptr, 1, C.byref(s))
    lib.oskar_element_set_element_type(lib.oskar_station_element(st.ptr, 0),
                                       b"Dipole", C.byref(s))
    check_status(s, "station")
    lib.oskar_station_set_position(st.ptr, np.radians(21.4),
                                   np.radians(-30.7), 0.0, 0.0, 0.0, 0.0)
    lib.oskar_station_set_phase_centre(st.ptr, AZEL, AZ, EL)
    return st


def directions(rng, n):
    v = rng.normal(size=(3, n))
    v /= np.linalg.norm(v, axis=0)
    v[2, : n // 8] = -np.abs(v[2, : n // 8])    # some surely below
    return v


def aa_beam(st, work, P, matrix, x, y, z, time_index=0):
    """The aperture-array beam of the given directions (blanked by the
    function itself)."""
    n = x.size
    ins = [Mem.from_numpy(v, P) for v in (x, y, z)]
    t = P | OSKAR_COMPLEX | (OSKAR_MATRIX if matrix else 0)
    out = Mem(t, OSKAR_GPU, n)
    s = C.c_int(0)
    lib.oskar_evaluate_station_beam_aperture_array(
        st.ptr, work.ptr, n, ins[0].ptr, ins[1].ptr, ins[2].ptr, time_index,
        lib.oskar_convert_mjd_to_gast_fast(MJD), FREQ, out.ptr, C.byref(s))
    check_status(s, "aa beam")
    sync_check("aa beam")
    return out.numpy().reshape(n, 8 if matrix else 2)


def station_beam(st, work, P, matrix, x, y, z, offset_out, time_index=0):
    n = x.size
    ins = [Mem.from_numpy(v, P) for v in (x, y, z)]
    t = P | OSKAR_COMPLEX | (OSKAR_MATRIX if matrix else 0)
    out = Mem(t, OSKAR_GPU, n + offset_out)
    s = C.c_int(0)
    lib.oskar_station_beam(
        st.ptr, work.ptr, ENU_DIR, n, capi.MEM3(*[m.ptr for m in ins]), 0.0,
        0.0, AZEL, float(AZ), float(EL), time_index, MJD, MJD, FREQ,
        offset_out, out.ptr, C.byref(s))
    check_status(s, "oskar_station_beam")
    sync_check("oskar_station_beam")
    return out.numpy().reshape(n + offset_out, 8 if matrix else 2)


def normalised(v, ref):
    """oskar_mem_normalise's scale (amplitude / sqrt 2 for a matrix)."""
    amp = np.sum(ref.astype(np.float64) ** 2)
    if ref.size == 8:
        amp /= 2.0
    return v / np.sqrt(amp)


CASES = [(p, m, nrm, 0.0) for p in ("d", "s") for m in (True, False)
         for nrm in (1, 0)] + [("d", True, 1, 0.4), ("s", True, 1, 0.4)]


@pytest.mark.parametrize("prec,matrix,normalise,height", CASES)
def test_blank_and_normalise_flat_station(hip, prec, matrix, normalise,
                                          height):
    print("station beam finish: %s matrix=%d normalise=%d height=%g"
          % (prec, matrix, normalise, height), flush=True)
    P = OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE
    f = np.float64 if prec == "d" else np.float32
    rng = np.random.default_rng(70 + 2 * matrix + normalise)
    st = flat_station(P, rng, height=height)
    lib.oskar_station_set_normalise_final_beam(st.ptr, normalise)
    work = StationWork(P, OSKAR_GPU)
    n, off = 3001, 5
    x, y, z = (v.astype(f) for v in directions(rng, n))
    bd = azel_dir(AZ, EL).astype(f)
    aa = aa_beam(st, work, P, matrix, *(np.append(v, b) for v, b in
                                        zip((x, y, z), bd)))
    got = station_beam(st, work, P, matrix, x, y, z, off)[off:]
    below = z < 0
    assert below.sum() > n // 4
    assert not np.any(got[below])
    want = normalised(aa[:n], aa[n]) if normalise else aa[:n]
    tol = 1e-12 if prec == "d" else 2e-6
    assert max_rel(got[~below], want[~below]) < tol


@pytest.mark.parametrize("prec", ["d", "s"])
def test_element_errors_and_weights(hip, prec):
    print("station beam finish: element errors, %s" % prec, flush=True)
    P = OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE
    f = np.float64 if prec == "d" else np.float32
    rng = np.random.default_rng(79)
    ne = 40
    st = flat_station(P, rng, ne=ne)
    s = C.c_int(0)
    for i in range(ne):
        lib.oskar_station_set_element_errors(
            st.ptr, 0, i, rng.uniform(0.8, 1.2), rng.uniform(0.0, 0.1),
            rng.uniform(-10, 10), rng.uniform(0.0, 5.0), C.byref(s))
        a = rng.uniform(0.3, 1.0) * np.exp(1j * rng.uniform(-0.3, 0.3))
        lib.oskar_station_set_element_weight(st.ptr, 0, i, a.real, a.imag,
                                             C.byref(s))
    lib.oskar_station_set_seed_time_variable_errors(st.ptr, 9)
    fin = C.c_int(0)
    lib.oskar_station_analyse(st.ptr, C.byref(fin), C.byref(s))
    check_status(s, "errors and weights")
    assert lib.oskar_station_apply_element_errors(st.ptr) == 1
    assert lib.oskar_station_apply_element_weight(st.ptr) == 1
    lib.oskar_station_set_normalise_final_beam(st.ptr, 1)
    work = StationWork(P, OSKAR_GPU)
    n = 2000
    x, y, z = (v.astype(f) for v in directions(rng, n))
    bd = azel_dir(AZ, EL).astype(f)
    aa = aa_beam(st, work, P, True, *(np.append(v, b) for v, b in
                                      zip((x, y, z), bd)), time_index=5)
    got = station_beam(st, work, P, True, x, y, z, 0, time_index=5)
    assert not np.any(got[z < 0])
    tol = 1e-12 if prec == "d" else 2e-6
    assert max_rel(got, normalised(aa[:n], aa[n])) < tol
    # Another time index draws other errors: the beam changes.
    other = station_beam(st, work, P, True, x, y, z, 0, time_index=6)
    assert max_rel(other, got) > 1e-4


def test_normalise_hierarchical_station(hip):
    print("station beam finish: hierarchical", flush=True)
    P = OSKAR_DOUBLE
    rng = np.random.default_rng(77)
    st = flat_station(P, rng, ne=6)
    s = C.c_int(0)
    lib.oskar_station_create_child_stations(st.ptr, C.byref(s))
    check_status(s, "children")
    child_xy = rng.uniform(-2.0, 2.0, size=(2, 9))
    xt = np.vstack([child_xy, np.zeros(9)])
    for i in range(6):
        ch = StationView(lib.oskar_station_child(st.ptr, i), P)
        lib.oskar_station_resize(ch.ptr, 9, C.byref(s))
        check_status(s, "child resize")
        ch.set_coords(xt, xt)
        lib.oskar_station_set_phase_centre(ch.ptr, AZEL, AZ, EL)
    lib.oskar_station_set_normalise_final_beam(st.ptr, 1)
    work = StationWork(P, OSKAR_GPU)
    n = 900
    x, y, z = directions(rng, n)
    bd = azel_dir(AZ, EL)
    aa = aa_beam(st, work, P, True, *(np.append(v, b) for v, b in
                                      zip((x, y, z), bd)))
    got = station_beam(st, work, P, True, x, y, z, 0)
    assert not np.any(got[z < 0])
    assert max_rel(got, normalised(aa[:n], aa[n])) < 1e-12
