"""Separate X and Y beams of an aperture-array station
(oskar_evaluate_station_beam_aperture_array.c:89-90, 204-218): when a station
holds feed-1 (Y) element data, each feed's weights and DFT, over that feed's
coordinates, gains and apodisation, fill its own row of the Jones matrix.

Checked through the C-ABI against closed forms on the common-beam path
(pinned against the oracle elsewhere): the X row equals that of a station
built from feed 0's data alone and the Y row that of a station built from
feed 1's, bit for bit in both precisions; and against the oracle's beams of
those two stations (double max rel < 1e-10).  Also: feed-1 data clear
common_pol_beams; elements not set for feed 1 keep feed 0's values; a scalar
beam uses feed 0 only.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from helpers import max_rel
from oskar_amd import lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_DOUBLE, OSKAR_GPU,
                           OSKAR_SINGLE, Mem, Station, StationWork,
                           check_status)
from gpu_util import sync_check
from test_station_errors_gpu import FREQ, analyse, azel_dir, beam, directions

pytestmark = pytest.mark.gpu
C0 = 299792458.0
AZEL = 4


def set_feed(st, feed, xt, xm, gain, ph_deg, apod):
    s = C.c_int(0)
    for i in range(xt.shape[1]):
        lib.oskar_station_set_element_coords(
            st.ptr, feed, i, (C.c_double * 3)(*xm[:, i]),
            (C.c_double * 3)(*xt[:, i]), C.byref(s))
        lib.oskar_station_set_element_errors(st.ptr, feed, i, gain[i], 0.0,
                                             ph_deg[i], 0.0, C.byref(s))
        lib.oskar_station_set_element_weight(st.ptr, feed, i, apod[i].real,
                                             apod[i].imag, C.byref(s))
    check_status(s, "set feed %d" % feed)


@pytest.mark.parametrize("prec", ["f", "d"])
def test_feed_rows_equal_single_feed_stations(hip, prec):
    print("separate X/Y beams", prec, flush=True)
    P = OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE
    rng = np.random.default_rng(71)
    ne = 40
    feeds = []
    for _ in range(2):
        r = 12.0 * np.sqrt(rng.uniform(size=ne))
        t = rng.uniform(0, 2 * np.pi, ne)
        xt = np.stack([r * np.cos(t), r * np.sin(t), np.zeros(ne)])
        xm = xt + np.concatenate([rng.normal(scale=0.01, size=(2, ne)),
                                  np.zeros((1, ne))])
        feeds.append((xt, xm, rng.uniform(0.7, 1.3, ne),
                      rng.uniform(-20, 20, ne),
                      rng.uniform(0.4, 1, ne) * np.exp(1j * rng.uniform(
                          -0.3, 0.3, ne))))
    pointing = (math.radians(120), math.radians(58))
    both = Station(P, OSKAR_GPU, ne)
    analyse(both)
    assert lib.oskar_station_common_pol_beams(both.ptr) == 1
    set_feed(both, 0, *feeds[0])
    set_feed(both, 1, *feeds[1])
    analyse(both)
    assert lib.oskar_station_common_pol_beams(both.ptr) == 0
    singles = []
    for fd in feeds:
        st = Station(P, OSKAR_GPU, ne)
        set_feed(st, 0, *fd)
        analyse(st)
        assert lib.oskar_station_common_pol_beams(st.ptr) == 1
        singles.append(st)
    for st in [both] + singles:
        lib.oskar_station_set_phase_centre(st.ptr, AZEL, *pointing)
    x, y, z = directions(rng, 1500)
    z[:25] = -np.abs(z[:25])
    got = beam(prec, both, x, y, z).reshape(-1, 4, 2)
    bx = beam(prec, singles[0], x, y, z).reshape(-1, 4, 2)
    by = beam(prec, singles[1], x, y, z).reshape(-1, 4, 2)
    # Row X = Jones elements a, b; row Y = c, d.
    np.testing.assert_array_equal(got[:, :2], bx[:, :2])
    np.testing.assert_array_equal(got[:, 2:], by[:, 2:])
    assert np.all(got[:25] == 0)
    if prec == "d":
        for row, (xt, xm, gain, ph, apod) in zip((slice(0, 2), slice(2, 4)),
                                                 feeds):
            err = oracle.element_errors(1, 0, 0, gain, np.zeros(ne),
                                        np.radians(ph), np.zeros(ne))
            ap = np.stack([apod.real, apod.imag], 1).ravel()
            want = oracle.station_beam_aa(
                1, *xt, *xm, np.zeros(ne), np.zeros(ne, np.int32),
                np.array([0], np.int32), np.array([0.5 * C0 / FREQ]),
                azel_dir(*pointing), x, y, z, FREQ, err=err,
                apod=ap).reshape(-1, 4, 2)
            assert max_rel(got[:, row], want[:, row]) < 1e-10


def test_unset_feed1_elements_keep_feed0_values(hip):
    rng = np.random.default_rng(72)
    ne = 16
    xt = np.stack([rng.normal(size=ne) * 5, rng.normal(size=ne) * 5,
                   np.zeros(ne)])
    st = Station(OSKAR_DOUBLE, OSKAR_GPU, ne)
    st.set_coords(xt)
    plain = Station(OSKAR_DOUBLE, OSKAR_GPU, ne)
    plain.set_coords(xt)
    s = C.c_int(0)
    # One feed-1 value, equal to feed 0's: separate beams, same values.
    lib.oskar_station_set_element_cable_length_error(st.ptr, 1, 2, 0.0,
                                                     C.byref(s))
    check_status(s, "oskar_station_set_element_cable_length_error")
    analyse(st)
    assert lib.oskar_station_common_pol_beams(st.ptr) == 0
    for t in (st, plain):
        lib.oskar_station_set_phase_centre(t.ptr, AZEL, 0.3, 1.1)
    x, y, z = directions(rng, 500)
    np.testing.assert_array_equal(beam("d", st, x, y, z),
                                  beam("d", plain, x, y, z))
    # A scalar beam uses feed 0 only.
    lib.oskar_station_set_element_cable_length_error(st.ptr, 1, 2, 0.37,
                                                     C.byref(s))
    analyse(st)
    outs = []
    for t in (st, plain):
        mx, my, mz = (Mem.from_numpy(a, OSKAR_DOUBLE) for a in (x, y, z))
        out = Mem(OSKAR_DOUBLE | OSKAR_COMPLEX, OSKAR_GPU, x.size)
        work = StationWork(OSKAR_DOUBLE, OSKAR_GPU)
        lib.oskar_evaluate_station_beam_aperture_array(
            t.ptr, work.ptr, x.size, mx.ptr, my.ptr, mz.ptr, 0, 0.0, FREQ,
            out.ptr, C.byref(s))
        check_status(s, "scalar station beam")
        sync_check("scalar station beam")
        outs.append(out.numpy())
    np.testing.assert_array_equal(outs[0], outs[1])
