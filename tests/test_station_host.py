"""Host logic of the station model, on the CPU (no device calls):
oskar_station_resize keeps every per-feed element array at the station's
length and gives new elements the reference's defaults
(oskar_station_resize.c:20-78: gain 1 and weight 1 + 0i for both feeds,
everything else 0), so a station grown after its gains, errors, phases or
weights were set is analysed and evaluated over arrays of the right length.
"""
import ctypes as C

import numpy as np

from oskar_amd import lib
from oskar_amd.mem import (OSKAR_CPU, OSKAR_DOUBLE, _Borrowed, Station,
                           check_status)


def _arr(ptr):
    assert ptr, "array not created"
    return _Borrowed(ptr).numpy()


def test_resize_after_errors_and_weights():
    ne, ne2 = 6, 11
    st = Station(OSKAR_DOUBLE, OSKAR_CPU, ne)
    s = C.c_int(0)
    lib.oskar_station_set_element_errors(st.ptr, 0, 2, 1.5, 0.1, 10.0, 2.0,
                                         C.byref(s))
    lib.oskar_station_set_element_weight(st.ptr, 0, 3, 0.5, -0.25, C.byref(s))
    # Feed 1 data created by setting one Y-feed value.
    lib.oskar_station_set_element_weight(st.ptr, 1, 4, 2.0, 0.0, C.byref(s))
    check_status(s, "set errors and weights")
    fin = C.c_int(0)
    lib.oskar_station_analyse(st.ptr, C.byref(fin), C.byref(s))
    check_status(s, "oskar_station_analyse")
    assert lib.oskar_station_apply_element_errors(st.ptr) == 1
    assert lib.oskar_station_apply_element_weight(st.ptr) == 1
    lib.oskar_station_resize(st.ptr, ne2, C.byref(s))
    check_status(s, "oskar_station_resize")
    lib.oskar_station_analyse(st.ptr, C.byref(fin), C.byref(s))
    check_status(s, "oskar_station_analyse after resize")
    for feed in (0, 1):
        gain = _arr(lib.oskar_station_element_gain_const(st.ptr, feed))
        gerr = _arr(lib.oskar_station_element_gain_error_const(st.ptr, feed))
        ph = _arr(lib.oskar_station_element_phase_offset_rad_const(st.ptr,
                                                                   feed))
        perr = _arr(lib.oskar_station_element_phase_error_rad_const(st.ptr,
                                                                    feed))
        w = _arr(lib.oskar_station_element_weight_const(st.ptr, feed))
        assert gain.size == gerr.size == ph.size == perr.size == ne2
        assert w.size == 2 * ne2
        # Old values kept, new elements at the defaults.
        assert gain[2] == 1.5 and gerr[2] == 0.1
        assert abs(ph[2] - np.radians(10.0)) < 1e-15
        assert abs(perr[2] - np.radians(2.0)) < 1e-15
        assert np.all(gain[ne:] == 1.0) and np.all(gerr[ne:] == 0.0)
        assert np.all(ph[ne:] == 0.0) and np.all(perr[ne:] == 0.0)
        wc = w[0::2] + 1j * w[1::2]
        assert np.all(wc[ne:] == 1.0)
        assert wc[3] == (0.5 - 0.25j if feed == 0 else 0.5 - 0.25j)
        assert wc[4] == (1.0 if feed == 0 else 2.0)
    # Shrinking keeps the first elements.
    lib.oskar_station_resize(st.ptr, 4, C.byref(s))
    check_status(s, "oskar_station_resize (shrink)")
    gain = _arr(lib.oskar_station_element_gain_const(st.ptr, 0))
    w1 = _arr(lib.oskar_station_element_weight_const(st.ptr, 1))
    assert gain.size == 4 and gain[2] == 1.5 and w1.size == 8


def test_resize_before_any_errors_creates_nothing():
    st = Station(OSKAR_DOUBLE, OSKAR_CPU, 5)
    s = C.c_int(0)
    lib.oskar_station_resize(st.ptr, 9, C.byref(s))
    check_status(s, "oskar_station_resize")
    assert not lib.oskar_station_element_gain_const(st.ptr, 0)
    assert not lib.oskar_station_element_weight_const(st.ptr, 1)
    lib.oskar_station_set_element_errors(st.ptr, 0, 8, 2.0, 0.0, 0.0, 0.0,
                                         C.byref(s))
    check_status(s, "set errors on a grown station")
    gain = _arr(lib.oskar_station_element_gain_const(st.ptr, 0))
    assert gain.size == 9 and gain[8] == 2.0 and np.all(gain[:8] == 1.0)
