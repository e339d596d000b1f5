"""Elements with their own orientations in the aperture-array station beam
(oskar_evaluate_station_beam_aperture_array.c:178-201: one pattern per
element, turned by its own Euler angles, row j of the DFT's signal being
element j's), through the C-ABI.

The array pattern is linear in the elements, so a station whose elements
fall into two orientation groups has the beam of the two sub-stations, each
of one common orientation (the path pinned against the oracle elsewhere),
added: checked against those sub-stations on the GPU (a closed form that
shares no code with the per-element branch) and against the oracle's
sub-station beams (double max rel < 1e-10, single RMS < 2e-5).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from helpers import max_rel, rel_rms
from oskar_amd import lib
from oskar_amd.mem import (OSKAR_DOUBLE, OSKAR_GPU, OSKAR_SINGLE, Station,
                           check_status)
from test_station_errors_gpu import FREQ, analyse, azel_dir, beam, directions

pytestmark = pytest.mark.gpu
C0 = 299792458.0
AZEL = 4


def station(prec, xt, xm, alpha_deg, swap_xy=0, norm_element=0):
    P = OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE
    ne = xt.shape[1]
    st = Station(P, OSKAR_GPU, ne)
    st.set_coords(xt, xm)
    s = C.c_int(0)
    for i in range(ne):
        for feed in (0, 1):
            lib.oskar_station_set_element_feed_angle(
                st.ptr, feed, i, float(alpha_deg[i]), 0.0, 0.0, C.byref(s))
    check_status(s, "oskar_station_set_element_feed_angle")
    lib.oskar_station_set_swap_xy(st.ptr, swap_xy)
    lib.oskar_station_set_normalise_element_pattern(st.ptr, norm_element)
    analyse(st)
    return st


@pytest.mark.parametrize("prec,swap_xy,norm_element",
                         [("d", 0, 0), ("d", 1, 1), ("f", 0, 1)])
def test_two_orientation_groups(hip, prec, swap_xy, norm_element):
    print("per-element orientation", prec, swap_xy, norm_element, flush=True)
    rng = np.random.default_rng(61)
    ne = 48
    r = 15.0 * np.sqrt(rng.uniform(size=ne))
    t = rng.uniform(0, 2 * np.pi, ne)
    xt = np.stack([r * np.cos(t), r * np.sin(t), np.zeros(ne)])
    xm = xt + np.concatenate([rng.normal(scale=0.01, size=(2, ne)),
                              np.zeros((1, ne))])
    group = rng.uniform(size=ne) < 0.5
    angles = (0.0, 35.0)
    alpha = np.where(group, angles[1], angles[0])
    pointing = (math.radians(25), math.radians(62))
    full = station(prec, xt, xm, alpha, swap_xy, norm_element)
    assert lib.oskar_station_common_element_orientation(full.ptr) == 0
    subs = []
    for g, a in zip((False, True), angles):
        sel = group == g
        sub = station(prec, xt[:, sel], xm[:, sel], np.full(sel.sum(), a),
                      swap_xy, norm_element)
        assert lib.oskar_station_common_element_orientation(sub.ptr) == 1
        subs.append((sub, sel, a))
    x, y, z = directions(rng, 1500)
    z[:30] = -np.abs(z[:30])
    for st in [full] + [t[0] for t in subs]:
        lib.oskar_station_set_phase_centre(st.ptr, AZEL, *pointing)
    got = beam(prec, full, x, y, z)
    parts = [beam(prec, sub, x, y, z) for sub, _, _ in subs]
    # Closed form: the two sub-stations added.
    if prec == "d":
        assert max_rel(got, parts[0] + parts[1]) < 1e-12
    else:
        assert rel_rms(got, parts[0] + parts[1]) < 1e-6
    # Oracle: the same sum of its sub-station beams.
    f = (lambda a: np.asarray(a, np.float32).astype(np.float64)) \
        if prec == "f" else (lambda a: np.asarray(a, np.float64))
    want = 0.0
    for _, sel, a in subs:
        n = int(sel.sum())
        want = want + oracle.station_beam_aa(
            1, *[f(v) for v in xt[:, sel]], *[f(v) for v in xm[:, sel]],
            np.zeros(n), np.zeros(n, np.int32), np.array([0], np.int32),
            np.array([0.5 * C0 / FREQ]), azel_dir(*pointing), f(x), f(y),
            f(z), FREQ, euler_x0=math.radians(a), euler_y0=math.radians(a),
            swap_xy=swap_xy, norm_element=norm_element)
    assert np.all(got.reshape(-1, 8)[:30] == 0)
    if prec == "d":
        assert max_rel(got, want) < 1e-10
    else:
        assert rel_rms(got, want) < 2e-5


def test_element_type_out_of_range(hip):
    rng = np.random.default_rng(62)
    ne = 8
    xt = np.stack([rng.normal(size=ne), rng.normal(size=ne), np.zeros(ne)])
    st = station("d", xt, xt, np.arange(ne) * 10.0)
    s = C.c_int(0)
    lib.oskar_station_set_element_type(st.ptr, 3, 1, C.byref(s))
    check_status(s, "oskar_station_set_element_type")
    lib.oskar_station_set_phase_centre(st.ptr, AZEL, 0.0, 1.2)
    x, y, z = directions(rng, 64)
    beam("d", st, x, y, z, expect_status=-5)  # OSKAR_ERR_OUT_OF_RANGE


@pytest.mark.parametrize("prec,swap_xy", [("d", 0), ("d", 1), ("f", 0)])
def test_mixed_angles_closed_form(hip, prec, swap_xy):
    """Every element with its own X and Y feed angles: the beam equals
    sum_a exp(-i k xm_a.b) exp(i k xt_a.s) g_a(s), with g_a the textbook
    dipole field of that element's two feeds (X along (cos ax, sin ax, 0),
    Y along (sin ay, -cos ay, 0)); no oracle code is involved."""
    from dipole_closed_form import dipole_matrix
    print("per-element orientation closed form", prec, swap_xy, flush=True)
    rng = np.random.default_rng(63)
    ne = 40
    r = 15.0 * np.sqrt(rng.uniform(size=ne))
    t = rng.uniform(0, 2 * np.pi, ne)
    xt = np.stack([r * np.cos(t), r * np.sin(t), np.zeros(ne)])
    xm = xt + np.concatenate([rng.normal(scale=0.01, size=(2, ne)),
                              np.zeros((1, ne))])
    ax = rng.uniform(-180, 180, ne)
    ay = ax + rng.uniform(-20, 20, ne)
    P = OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE
    st = Station(P, OSKAR_GPU, ne)
    st.set_coords(xt, xm)
    s = C.c_int(0)
    for i in range(ne):
        lib.oskar_station_set_element_feed_angle(
            st.ptr, 0, i, float(ax[i]), 0.0, 0.0, C.byref(s))
        lib.oskar_station_set_element_feed_angle(
            st.ptr, 1, i, float(ay[i]), 0.0, 0.0, C.byref(s))
    check_status(s, "oskar_station_set_element_feed_angle")
    lib.oskar_station_set_swap_xy(st.ptr, swap_xy)
    analyse(st)
    assert lib.oskar_station_common_element_orientation(st.ptr) == 0
    pointing = (math.radians(-40), math.radians(55))
    lib.oskar_station_set_phase_centre(st.ptr, AZEL, *pointing)
    x, y, z = directions(rng, 800)
    got = beam(prec, st, x, y, z).reshape(-1, 4, 2)
    got = (got[..., 0] + 1j * got[..., 1]).reshape(-1, 2, 2)
    if prec == "f":
        f = lambda a: np.asarray(a, np.float32).astype(np.float64)  # noqa
        xt, xm, x, y, z = f(xt), f(xm), f(x), f(y), f(z)
    k = 2 * np.pi * FREQ / C0
    b = azel_dir(*pointing)
    w = np.exp(-1j * k * (xm.T @ b))
    want = np.zeros((x.size, 2, 2), complex)
    for a in range(ne):
        ph = np.exp(1j * k * (xt[0, a] * x + xt[1, a] * y + xt[2, a] * z))
        want += (w[a] * ph)[:, None, None] * dipole_matrix(
            math.radians(ax[a]), math.radians(ay[a]), x, y, z, np.pi / 2,
            swap_xy)
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    rms = rel_rms(np.stack([got.real, got.imag]),
                  np.stack([want.real, want.imag]))
    print("max err / peak %.3e, rel RMS %.3e" % (err, rms), flush=True)
    if prec == "d":
        assert err < 1e-10
    else:
        assert rms < 2e-5
