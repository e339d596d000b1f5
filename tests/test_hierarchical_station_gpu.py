"""Hierarchical aperture-array stations, whose elements are child stations
(oskar_evaluate_station_beam_aperture_array.c:242-287), through the C-ABI:

  - different children (own layouts and phase centres; the parent with
    normalisation and apodisation): the beam equals the oracle's
    composition, each child's oracle beam as the signal of the parent's
    oracle dftw (double max rel < 1e-10, single RMS < 2e-5);
  - three levels (children with children of their own): the same
    composition applied twice;
  - alike children (evaluated once and copied): the pattern multiplication
    theorem, beam = child beam x the parent's array factor computed in
    numpy, at 1.1 M directions, which takes two passes of the recursion;
  - a parent with the array pattern off is a settings error.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from gpu_util import sync_check
from helpers import max_rel, rel_rms
from oskar_amd import lib
from oskar_amd.mem import (OSKAR_DOUBLE, OSKAR_GPU, OSKAR_SINGLE, Station,
                           StationView, check_status)
from test_station_errors_gpu import FREQ, analyse, azel_dir, beam, directions

pytestmark = pytest.mark.gpu
C0 = 299792458.0
AZEL = 4
K = 2 * math.pi * FREQ / C0
OSKAR_ERR_SETTINGS_TELESCOPE = -23  # oskar_global.h


def layout(ne, radius, rng, jitter=0.0):
    r = radius * np.sqrt(rng.uniform(size=ne))
    t = rng.uniform(0, 2 * np.pi, ne)
    xt = np.stack([r * np.cos(t), r * np.sin(t), np.zeros(ne)])
    xm = xt + np.concatenate([rng.normal(scale=jitter, size=(2, ne)),
                              np.zeros((1, ne))])
    return xt, xm


def setup_children(parent, prec, ne_child, rng, pointing, alike=False):
    """Gives every element of parent a leaf child station of ne_child
    dipoles; returns each child's (xt, xm, pointing)."""
    P = OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE
    s = C.c_int(0)
    lib.oskar_station_create_child_stations(parent.ptr, C.byref(s))
    check_status(s, "oskar_station_create_child_stations")
    spec = []
    n = lib.oskar_station_num_elements(parent.ptr)
    common = layout(ne_child, 3.0, rng, 0.002)
    for i in range(n):
        ch = StationView(lib.oskar_station_child(parent.ptr, i), P)
        lib.oskar_station_resize(ch.ptr, ne_child, C.byref(s))
        check_status(s, "oskar_station_resize")
        xt, xm = common if alike else layout(ne_child, 3.0, rng, 0.002)
        ch.set_coords(xt, xm)
        az, el = pointing if alike else (
            pointing[0] + rng.normal(scale=0.02),
            pointing[1] + rng.normal(scale=0.02))
        lib.oskar_station_set_phase_centre(ch.ptr, AZEL, az, el)
        spec.append((xt, xm, (az, el)))
    return spec


def oracle_leaf(xt, xm, pointing, x, y, z, f):
    ne = xt.shape[1]
    return oracle.station_beam_aa(
        1, *[f(v) for v in xt], *[f(v) for v in xm], np.zeros(ne),
        np.zeros(ne, np.int32), np.array([0], np.int32),
        np.array([0.5 * C0 / FREQ]), azel_dir(*pointing), f(x), f(y), f(z),
        FREQ)


def oracle_parent(xt, xm, pointing, signals, x, y, z, f, norm, apod=None):
    ne = xt.shape[1]
    w = np.zeros(2 * ne)
    d = azel_dir(*pointing)
    oracle.lib().oracle_element_weights_dft_d(
        ne, *[oracle._p(np.ascontiguousarray(f(v))) for v in xm],
        oracle._p(np.zeros(ne)), K, float(d[0]), float(d[1]), float(d[2]),
        oracle._p(w))
    if apod is not None:
        wc = (w[0::2] + 1j * w[1::2]) * apod
        w = np.stack([wc.real, wc.imag], 1).ravel()
    return oracle.dftw("double", 1, norm, K, w, *[f(v) for v in xt],
                       f(x), f(y), f(z), None, np.concatenate(signals))


def prec_fn(prec):
    if prec == "f":
        return lambda a: np.asarray(a, np.float32).astype(np.float64)
    return lambda a: np.asarray(a, np.float64)


@pytest.mark.parametrize("prec", ["f", "d"])
def test_different_children_match_oracle(hip, prec):
    print("hierarchical station, different children", prec, flush=True)
    rng = np.random.default_rng(51)
    P = OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE
    ne = 12
    parent = Station(P, OSKAR_GPU, ne)
    pxt, pxm = layout(ne, 25.0, rng, 0.01)
    parent.set_coords(pxt, pxm)
    pointing = (math.radians(40), math.radians(65))
    spec = setup_children(parent, prec, 16, rng, pointing)
    apod = rng.uniform(0.3, 1.0, ne) * np.exp(1j * rng.uniform(-0.2, 0.2, ne))
    s = C.c_int(0)
    for i in range(ne):
        lib.oskar_station_set_element_weight(parent.ptr, 0, i, apod[i].real,
                                             apod[i].imag, C.byref(s))
    check_status(s, "oskar_station_set_element_weight")
    lib.oskar_station_set_normalise_array_pattern(parent.ptr, 1)
    lib.oskar_station_set_phase_centre(parent.ptr, AZEL, *pointing)
    analyse(parent)
    assert lib.oskar_station_has_child(parent.ptr) == 1
    assert lib.oskar_station_identical_children(parent.ptr) == 0
    x, y, z = directions(rng, 1200)
    z[:40] = -np.abs(z[:40])          # below the horizon: blanked
    got = beam(prec, parent, x, y, z)
    f = prec_fn(prec)
    sig = [oracle_leaf(xt, xm, pt, x, y, z, f) for xt, xm, pt in spec]
    want = oracle_parent(pxt, pxm, pointing, sig, x, y, z, f, 1,
                         f(apod.real) + 1j * f(apod.imag))
    assert np.all(got.reshape(-1, 8)[:40] == 0)
    if prec == "d":
        assert max_rel(got, want) < 1e-10
    else:
        assert rel_rms(got, want) < 2e-5


def test_three_levels_match_oracle(hip):
    print("hierarchical station, three levels", flush=True)
    rng = np.random.default_rng(52)
    f = prec_fn("d")
    top = Station(OSKAR_DOUBLE, OSKAR_GPU, 4)
    txt, txm = layout(4, 120.0, rng, 0.0)
    top.set_coords(txt, txm)
    pointing = (math.radians(200), math.radians(55))
    s = C.c_int(0)
    lib.oskar_station_create_child_stations(top.ptr, C.byref(s))
    check_status(s, "oskar_station_create_child_stations")
    mids = []
    for i in range(4):
        mid = StationView(lib.oskar_station_child(top.ptr, i), OSKAR_DOUBLE)
        lib.oskar_station_resize(mid.ptr, 5, C.byref(s))
        mxt, mxm = layout(5, 20.0, rng, 0.005)
        mid.set_coords(mxt, mxm)
        mpt = (pointing[0] + 0.01 * i, pointing[1] - 0.01 * i)
        lib.oskar_station_set_phase_centre(mid.ptr, AZEL, *mpt)
        leaves = setup_children(mid, "d", 8, rng, mpt)
        mids.append((mxt, mxm, mpt, leaves))
    lib.oskar_station_set_phase_centre(top.ptr, AZEL, *pointing)
    counter = C.c_int(0)
    lib.oskar_station_set_unique_ids(top.ptr, C.byref(counter))
    assert counter.value == 1 + 4 + 4 * 5
    assert lib.oskar_station_unique_id(
        lib.oskar_station_child(lib.oskar_station_child(top.ptr, 1), 0)) == 8
    analyse(top)
    x, y, z = directions(rng, 800)
    got = beam("d", top, x, y, z)
    sig = []
    for mxt, mxm, mpt, leaves in mids:
        leaf_sig = [oracle_leaf(a, b, c, x, y, z, f) for a, b, c in leaves]
        sig.append(oracle_parent(mxt, mxm, mpt, leaf_sig, x, y, z, f, 0))
    want = oracle_parent(txt, txm, pointing, sig, x, y, z, f, 0)
    assert max_rel(got, want) < 1e-10


def test_alike_children_pattern_multiplication(hip):
    print("hierarchical station, alike children, 1.1M directions", flush=True)
    rng = np.random.default_rng(53)
    ne = 6
    parent = Station(OSKAR_DOUBLE, OSKAR_GPU, ne)
    pxt, _ = layout(ne, 30.0, rng, 0.0)
    parent.set_coords(pxt)
    pointing = (math.radians(10), math.radians(70))
    spec = setup_children(parent, "d", 10, rng, pointing, alike=True)
    lib.oskar_station_set_phase_centre(parent.ptr, AZEL, *pointing)
    lib.oskar_station_set_normalise_array_pattern(parent.ptr, 1)
    analyse(parent)
    assert lib.oskar_station_identical_children(parent.ptr) == 1
    x, y, z = directions(rng, (1 << 20) + 50000)
    got = beam("d", parent, x, y, z).reshape(-1, 8)
    # The child as a station of its own.
    xt, xm, pt = spec[0]
    child = Station(OSKAR_DOUBLE, OSKAR_GPU, xt.shape[1])
    child.set_coords(xt, xm)
    lib.oskar_station_set_phase_centre(child.ptr, AZEL, *pt)
    cb = beam("d", child, x, y, z).reshape(-1, 8)
    s0 = azel_dir(*pointing)
    ph = K * (np.outer(x, pxt[0]) + np.outer(y, pxt[1]) - (s0 @ pxt)[None, :])
    af = np.exp(1j * ph).sum(axis=1) / ne
    cc = cb[:, 0::2] + 1j * cb[:, 1::2]
    want = cc * af[:, None]
    gc = got[:, 0::2] + 1j * got[:, 1::2]
    err = np.abs(gc - want)
    assert np.max(err) < 1e-11 * max(1.0, np.max(np.abs(want)))


def test_children_need_the_array_pattern(hip):
    rng = np.random.default_rng(54)
    parent = Station(OSKAR_DOUBLE, OSKAR_GPU, 3)
    pxt, _ = layout(3, 10.0, rng)
    parent.set_coords(pxt)
    setup_children(parent, "d", 4, rng, (0.0, 1.2))
    lib.oskar_station_set_enable_array_pattern(parent.ptr, 0)
    analyse(parent)
    x, y, z = directions(rng, 64)
    beam("d", parent, x, y, z, expect_status=OSKAR_ERR_SETTINGS_TELESCOPE)
