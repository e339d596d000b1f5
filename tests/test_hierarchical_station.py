"""Child-station host logic (no GPU): oskar_station_create_child_stations
(oskar_station_create_child_stations.c:15-44), oskar_station_different
(oskar_station_different.c:15-200), the identical-children flag of
oskar_station_analyse (oskar_station_analyse.c:193-220), recursive copies
and unique ids (oskar_station_accessors.c:503-515)."""
import ctypes as C

import numpy as np

from oskar_amd import lib
from oskar_amd.mem import (OSKAR_CPU, OSKAR_DOUBLE, Station, StationView,
                           check_status)

OSKAR_ERR_MEMORY_ALLOC_FAILURE = -6  # oskar_global.h


def tree(n_children=3, n_leaf=4, seed=0):
    rng = np.random.default_rng(seed)
    parent = Station(OSKAR_DOUBLE, OSKAR_CPU, n_children)
    parent.set_coords(rng.normal(size=(3, n_children)) * [[10], [10], [0]])
    s = C.c_int(0)
    lib.oskar_station_create_child_stations(parent.ptr, C.byref(s))
    check_status(s, "oskar_station_create_child_stations")
    xyz = rng.normal(size=(3, n_leaf)) * [[2], [2], [0]]
    for i in range(n_children):
        ch = StationView(lib.oskar_station_child(parent.ptr, i), OSKAR_DOUBLE)
        assert lib.oskar_station_num_elements(ch.ptr) == 0
        lib.oskar_station_resize(ch.ptr, n_leaf, C.byref(s))
        ch.set_coords(xyz)
    return parent


def analyse(st, fin=0):
    f, s = C.c_int(fin), C.c_int(0)
    lib.oskar_station_analyse(st.ptr, C.byref(f), C.byref(s))
    check_status(s, "oskar_station_analyse")
    return f.value


def test_create_children_once():
    parent = tree()
    assert lib.oskar_station_has_child(parent.ptr) == 1
    s = C.c_int(0)
    lib.oskar_station_create_child_stations(parent.ptr, C.byref(s))
    assert s.value == OSKAR_ERR_MEMORY_ALLOC_FAILURE
    leaf = Station(OSKAR_DOUBLE, OSKAR_CPU, 2)
    assert lib.oskar_station_has_child(leaf.ptr) == 0


def test_identical_children_and_differences():
    parent = tree()
    analyse(parent)
    assert lib.oskar_station_identical_children(parent.ptr) == 1
    c0 = lib.oskar_station_child(parent.ptr, 0)
    c2 = lib.oskar_station_child(parent.ptr, 2)
    s = C.c_int(0)
    assert lib.oskar_station_different(c0, c2, C.byref(s)) == 0
    # The unique id is not compared.
    counter = C.c_int(5)
    lib.oskar_station_set_unique_ids(parent.ptr, C.byref(counter))
    assert counter.value == 9
    assert [lib.oskar_station_unique_id(lib.oskar_station_child(
        parent.ptr, i)) for i in range(3)] == [6, 7, 8]
    assert lib.oskar_station_different(c0, c2, C.byref(s)) == 0
    # A different element position, phase centre or apodisation weight.
    lib.oskar_station_set_element_weight(c2, 0, 1, 0.5, 0.0, C.byref(s))
    check_status(s, "oskar_station_set_element_weight")
    analyse(parent)
    assert lib.oskar_station_different(c0, c2, C.byref(s)) == 1
    assert lib.oskar_station_identical_children(parent.ptr) == 0
    parent = tree()
    c1 = lib.oskar_station_child(parent.ptr, 1)
    lib.oskar_station_set_phase_centre(c1, 4, 0.1, 1.0)
    analyse(parent)
    assert lib.oskar_station_identical_children(parent.ptr) == 0
    # Stations with and without children differ.
    a, b = tree(), Station(OSKAR_DOUBLE, OSKAR_CPU, 3)
    assert lib.oskar_station_different(a.ptr, b.ptr, C.byref(s)) == 1


def test_random_errors_end_the_identity_check():
    parent = tree()
    s = C.c_int(0)
    for i in range(3):
        c = lib.oskar_station_child(parent.ptr, i)
        # Gain 1, phase 0, gain error scale 0.1: random errors only.
        lib.oskar_station_set_element_errors(c, 0, 0, 1.0, 0.1, 0.0, 0.0,
                                             C.byref(s))
    check_status(s, "oskar_station_set_element_errors")
    assert analyse(parent) == 1
    assert lib.oskar_station_identical_children(parent.ptr) == 0
    # A finished check handed in also leaves the flag clear.
    parent = tree()
    analyse(parent, fin=1)
    assert lib.oskar_station_identical_children(parent.ptr) == 0


def test_copy_is_recursive():
    parent = tree(seed=3)
    analyse(parent)
    s = C.c_int(0)
    cp = lib.oskar_station_create_copy(parent.ptr, OSKAR_CPU, C.byref(s))
    check_status(s, "oskar_station_create_copy")
    try:
        assert lib.oskar_station_has_child(cp) == 1
        assert lib.oskar_station_identical_children(cp) == 1
        assert lib.oskar_station_different(parent.ptr, cp, C.byref(s)) == 0
        lib.oskar_station_set_phase_centre(
            lib.oskar_station_child(cp, 1), 4, 0.3, 0.9)
        assert lib.oskar_station_different(parent.ptr, cp, C.byref(s)) == 1
    finally:
        lib.oskar_station_free(cp, C.byref(C.c_int(0)))


def test_feed1_data_host_logic():
    """Feed-1 element data (oskar_station_set_element_coords.c,
    oskar_station_accessors.c:259-275, oskar_station_analyse.c:62-79):
    accessors fall back to feed 0's arrays until feed 1 has its own, which
    start as a copy of feed 0's; any feed-1 array clears common_pol_beams;
    copies and comparisons include them."""
    rng = np.random.default_rng(9)
    ne = 5
    xyz = rng.normal(size=(3, ne)) * [[3], [3], [0]]
    st = Station(OSKAR_DOUBLE, OSKAR_CPU, ne)
    st.set_coords(xyz)
    f0 = lib.oskar_station_element_true_enu_metres(st.ptr, 0, 0)
    assert lib.oskar_station_element_true_enu_metres(st.ptr, 1, 0) == f0
    assert lib.oskar_station_common_pol_beams(st.ptr) == 1
    twin = lib.oskar_station_create_copy(st.ptr, OSKAR_CPU, C.byref(C.c_int(0)))
    s = C.c_int(0)
    try:
        lib.oskar_station_set_element_coords(
            st.ptr, 1, 3, (C.c_double * 3)(1.5, -2.0, 0.0),
            (C.c_double * 3)(1.5, -2.0, 0.0), C.byref(s))
        check_status(s, "oskar_station_set_element_coords")
        f1 = lib.oskar_station_element_true_enu_metres(st.ptr, 1, 0)
        assert f1 and f1 != f0
        analyse(st)
        assert lib.oskar_station_common_pol_beams(st.ptr) == 0
        from oskar_amd.mem import _Borrowed
        x1 = _Borrowed(f1).numpy()
        want = xyz[0].copy()
        want[3] = 1.5
        np.testing.assert_array_equal(x1, want)
        np.testing.assert_array_equal(_Borrowed(f0).numpy(), xyz[0])
        assert lib.oskar_station_different(st.ptr, twin, C.byref(s)) == 1
        cp = lib.oskar_station_create_copy(st.ptr, OSKAR_CPU, C.byref(s))
        assert lib.oskar_station_different(st.ptr, cp, C.byref(s)) == 0
        lib.oskar_station_free(cp, C.byref(C.c_int(0)))
        lib.oskar_station_set_element_weight(st.ptr, 2, 0, 1.0, 0.0,
                                             C.byref(s))
        assert s.value == -5   # OSKAR_ERR_OUT_OF_RANGE
    finally:
        lib.oskar_station_free(twin, C.byref(C.c_int(0)))
