"""oskar_evaluate_jones_E through the C-ABI
(oskar/interferometer/src/oskar_evaluate_jones_E.c:88-192): row i of E is
the beam of station i's model from oskar_station_beam (itself pinned
against the CPU oracle of the aperture-array beam), bit for bit, without
duplication (model i) and with it (model type_map[i], each model evaluated
once and copied); argument errors.
"""
import ctypes as C

import numpy as np
import pytest

from gpu_util import sync_check
from oskar_amd import capi, lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_DOUBLE, OSKAR_GPU,
                           OSKAR_MATRIX, Jones, Mem, StationView,
                           StationWork, Telescope, _Borrowed, check_status)

pytestmark = pytest.mark.gpu
AZEL = 4
CM64 = OSKAR_DOUBLE | OSKAR_COMPLEX | OSKAR_MATRIX
FREQ, MJD = 140e6, 60000.25


def telescope(rng, num_stations, num_models):
    tel = Telescope(OSKAR_DOUBLE, OSKAR_GPU, num_stations)
    s = C.c_int(0)
    lib.oskar_telescope_resize_station_array(tel.ptr, num_models, C.byref(s))
    check_status(s, "resize_station_array")
    lib.oskar_telescope_set_phase_centre(tel.ptr, AZEL, 0.3, 1.1)
    for k in range(num_models):
        ne = 20 + 7 * k
        st = StationView(lib.oskar_telescope_station(tel.ptr, k),
                         OSKAR_DOUBLE)
        lib.oskar_station_resize(st.ptr, ne, C.byref(s))
        lib.oskar_station_resize_element_types(st.ptr, 1, C.byref(s))
        lib.oskar_element_set_element_type(
            lib.oskar_station_element(st.ptr, 0), b"Dipole", C.byref(s))
        check_status(s, "station model")
        r = 12.0 * np.sqrt(rng.uniform(size=ne))
        t = rng.uniform(0, 2 * np.pi, ne)
        st.set_coords(np.stack([r * np.cos(t), r * np.sin(t),
                                np.zeros(ne)]))
        lib.oskar_station_set_phase_centre(st.ptr, AZEL, 0.2 * k, 1.0)
    return tel


def sources(rng, n):
    az = rng.uniform(-np.pi, np.pi, n)
    el = rng.uniform(0.05, np.pi / 2, n)
    return [Mem.from_numpy(a, OSKAR_DOUBLE) for a in (az, el, el)]


def jones_E(tel, num_stations, n, coords, work, extra=0, expect=0):
    E = Jones(CM64, OSKAR_GPU, num_stations, n + extra)
    s = C.c_int(0)
    lib.oskar_evaluate_jones_E(E.ptr, AZEL, n,
                               capi.MEM3(*[m.ptr for m in coords]), 0.0, 0.0,
                               tel.ptr, 0, MJD, MJD, FREQ, work.ptr,
                               C.byref(s))
    if expect:
        assert s.value == expect
        return None
    check_status(s, "oskar_evaluate_jones_E")
    sync_check("oskar_evaluate_jones_E")
    return E.mem.numpy().reshape(num_stations, n + extra, 8)


def model_beam(tel, k, n, coords, work):
    out = Mem(CM64, OSKAR_GPU, n)
    s = C.c_int(0)
    lib.oskar_station_beam(
        lib.oskar_telescope_station(tel.ptr, k), work.ptr, AZEL, n,
        capi.MEM3(*[m.ptr for m in coords]), 0.0, 0.0, AZEL, 0.3, 1.1, 0,
        MJD, MJD, FREQ, 0, out.ptr, C.byref(s))
    check_status(s, "oskar_station_beam")
    sync_check("oskar_station_beam")
    return out.numpy().reshape(n, 8)


def test_jones_E_one_model_per_station(hip):
    rng = np.random.default_rng(1)
    ns, n = 4, 3001
    tel = telescope(rng, ns, ns)
    coords = sources(rng, n)
    work = StationWork(OSKAR_DOUBLE, OSKAR_GPU)
    E = jones_E(tel, ns, n, coords, work)
    for i in range(ns):
        want = model_beam(tel, i, n, coords, work)
        assert np.array_equal(E[i], want), i
    assert not np.array_equal(E[0], E[1])


def test_jones_E_duplication_type_map(hip):
    rng = np.random.default_rng(2)
    ns, n, extra = 6, 2048, 5
    tel = telescope(rng, ns, 3)
    lib.oskar_telescope_set_allow_station_beam_duplication(tel.ptr, 1)
    assert lib.oskar_telescope_allow_station_beam_duplication(tel.ptr) == 1
    tmap = _Borrowed(lib.oskar_telescope_station_type_map(tel.ptr))
    assert np.array_equal(tmap.numpy(), np.zeros(ns, np.int32))
    types = np.array([2, 0, 2, 1, 0, 2], np.int32)
    arr = np.ctypeslib.as_array(
        C.cast(lib.oskar_mem_void(tmap.ptr), C.POINTER(C.c_int)), (ns,))
    arr[:] = types
    coords = sources(rng, n)
    work = StationWork(OSKAR_DOUBLE, OSKAR_GPU)
    E = jones_E(tel, ns, n, coords, work, extra=extra)
    beams = [model_beam(tel, k, n, coords, work) for k in range(3)]
    for i in range(ns):
        assert np.array_equal(E[i, :n], beams[types[i]]), i
    s = C.c_int(0)
    lib.oskar_telescope_set_unique_stations(tel.ptr, 1, C.byref(s))
    assert np.array_equal(tmap.numpy(), np.arange(ns))


def test_jones_E_errors(hip):
    rng = np.random.default_rng(3)
    tel = telescope(rng, 5, 2)
    coords = sources(rng, 10)
    work = StationWork(OSKAR_DOUBLE, OSKAR_GPU)
    jones_E(tel, 4, 10, coords, work, expect=-11)      # station count
    jones_E(tel, 5, 10, coords, work, expect=-5)       # models < stations
