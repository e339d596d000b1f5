"""The oracle's relative directions and Gaussian source parameters, pinned:

  * against the reference's own function, compiled from its sources
    (oracle/_ref/liboskar_sky_ref.so, fixture tests/golden/sky_params.npz):
    bit for bit in float and double, fit failures and flux columns
    included, on the reference test's sky and on an all-sky random set;
  * against closed forms that do not share the oracle's arithmetic:
      - a u^2 + 2 b u v + c v^2 is the exponent of the Fourier transform of
        an elliptical Gaussian, 2 pi^2 (s_maj^2 u_maj^2 + s_min^2 u_min^2),
        with (u_maj, u_min) the baseline rotated to the source's axes
        (position angle from north through east);
      - a small circular source at angular distance d from the phase centre
        projects onto the tangent plane as an ellipse foreshortened by
        cos d along the offset (orthographic projection);
      - relative direction cosines are the dot products of the unit vector
        with the east, north and up axes at the phase centre.
"""
import math

import numpy as np
import pytest

import oracle
from helpers import load_golden

RA, DEC, I_, Q_, U_, V_, MAJ, MIN, PA = 1, 2, 3, 4, 5, 6, 10, 11, 12
EXT_A, EXT_B, EXT_C = 0x100, 0x101, 0x102
ARCSEC = math.pi / 180 / 3600
K = 0.5 * math.pi ** 2 / (2 * math.log(2))  # 2 pi^2 sigma^2 / FWHM^2


def dt(prec):
    return np.float32 if prec == "f" else np.float64


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("case", ["reftest", "rand"])
def test_gaussian_oracle_equals_reference(prec, case):
    g = load_golden("sky_params.npz")
    ra0, dec0 = g["%s_ra0_dec0" % case]
    inp = {k: g["%s_in_%d" % (case, k)].astype(dt(prec))
           for k in (RA, DEC, MAJ, MIN, PA)}
    n = inp[RA].size
    start = np.full(n, -7.0, dt(prec))
    a, b, c, failed = oracle.gaussian_parameters(
        prec, inp[RA], inp[DEC], inp[MAJ], inp[MIN], inp[PA], ra0, dec0,
        start, start, start)
    assert int(failed.sum()) == int(g["%s_failed_%s" % (case, prec)])
    for out, col in ((a, EXT_A), (b, EXT_B), (c, EXT_C)):
        np.testing.assert_array_equal(
            out, g["%s_out_%d_%s" % (case, col, prec)])
    if case == "rand":
        # Sources with both axes zero are skipped.  (A zero minor axis is
        # a line, which fails only where the projection keeps it straight.)
        assert not failed[100:200].any()
        assert np.all(a[100:200] == -7.0)
        assert failed[:100].any()
        # zero_failed = 1 clears nothing in the reference: its call to
        # oskar_sky_clear_source_flux (:155-158) still sees the fit's error
        # in *status, so it returns at once (oskar_sky_clear_source_flux.c
        # :17).  Every flux column comes back unchanged.
        for col in (I_, Q_, U_, V_):
            np.testing.assert_array_equal(
                g["rand_out_%d_%s" % (col, prec)],
                g["rand_in_%d" % col].astype(dt(prec)))


def ft_exponent(maj, mn, pa, u, v):
    """2 pi^2 (s_maj^2 u_maj^2 + s_min^2 u_min^2) of an elliptical Gaussian
    whose major axis points at position angle pa (from +m towards +l)."""
    s2 = lambda fwhm: (fwhm / (2 * math.sqrt(2 * math.log(2)))) ** 2
    u_maj = u * np.sin(pa) + v * np.cos(pa)
    u_min = u * np.cos(pa) - v * np.sin(pa)
    return 2 * math.pi ** 2 * (s2(maj) * u_maj ** 2 + s2(mn) * u_min ** 2)


def test_gaussian_oracle_at_phase_centre_is_fourier_exponent():
    rng = np.random.default_rng(3)
    n = 400
    ra0, dec0 = 1.1, -0.4
    maj = rng.uniform(1, 60, n) * ARCSEC
    mn = maj * rng.uniform(0.2, 0.95, n)
    pa = rng.uniform(0, math.pi, n)
    a, b, c, failed = oracle.gaussian_parameters(
        "d", np.full(n, ra0), np.full(n, dec0), maj, mn, pa, ra0, dec0)
    assert not failed.any()
    u, v = rng.normal(size=(2, n)) * 2e4
    q = a * u ** 2 + 2 * b * u * v + c * v ** 2
    np.testing.assert_allclose(q, ft_exponent(maj, mn, pa, u, v),
                               rtol=1e-7)


@pytest.mark.parametrize("axis", ["dec", "ra"])
@pytest.mark.parametrize("offset_deg", [5.0, 30.0, 60.0])
def test_gaussian_oracle_orthographic_foreshortening(axis, offset_deg):
    # A circular source at distance d from the phase centre along one
    # axis: foreshortened by cos d along that axis, full size across it.
    fwhm = 10 * ARCSEC
    d = math.radians(offset_deg)
    ra0, dec0 = 0.7, 0.0
    ra, dec = (ra0, dec0 + d) if axis == "dec" else (ra0 + d, dec0)
    a, b, c, failed = oracle.gaussian_parameters(
        "d", np.array([ra]), np.array([dec]), np.array([fwhm]),
        np.array([fwhm]), np.array([0.0]), ra0, dec0)
    assert not failed.any()
    big, small = K * fwhm ** 2, K * (fwhm * math.cos(d)) ** 2
    # Offset along Dec shortens m (the v coefficient c); along RA, l (a).
    want_a, want_c = (big, small) if axis == "dec" else (small, big)
    assert a[0] == pytest.approx(want_a, rel=1e-3)
    assert c[0] == pytest.approx(want_c, rel=1e-3)
    assert abs(b[0]) < 1e-3 * big


@pytest.mark.parametrize("prec", ["f", "d"])
def test_relative_directions_oracle_equals_reference(prec):
    g = load_golden("sky_params.npz")
    ra0, dec0 = g["rand_ra0_dec0"]
    l, m, n = oracle.relative_directions(
        prec, g["rand_in_%d" % RA].astype(dt(prec)),
        g["rand_in_%d" % DEC].astype(dt(prec)), ra0, dec0)
    np.testing.assert_array_equal(l, g["reldir_l_%s" % prec])
    np.testing.assert_array_equal(m, g["reldir_m_%s" % prec])
    np.testing.assert_array_equal(n, g["reldir_n_%s" % prec])


def test_relative_directions_oracle_are_enu_projections():
    rng = np.random.default_rng(4)
    ra = rng.uniform(0, 2 * math.pi, 1000)
    dec = np.arcsin(rng.uniform(-1, 1, 1000))
    ra0, dec0 = 4.0, 0.9
    l, m, n = oracle.relative_directions("d", ra, dec, ra0, dec0)
    s = np.stack([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra),
                  np.sin(dec)])
    east = np.array([-math.sin(ra0), math.cos(ra0), 0.0])
    north = np.array([-math.sin(dec0) * math.cos(ra0),
                      -math.sin(dec0) * math.sin(ra0), math.cos(dec0)])
    up = np.array([math.cos(dec0) * math.cos(ra0),
                   math.cos(dec0) * math.sin(ra0), math.sin(dec0)])
    np.testing.assert_allclose(l, east @ s, atol=1e-15)
    np.testing.assert_allclose(m, north @ s, atol=1e-15)
    np.testing.assert_allclose(n, up @ s, atol=1e-15)
