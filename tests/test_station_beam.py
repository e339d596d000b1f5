"""Sidereal time used by oskar_station_beam (host functions, no GPU).

Pinned by published values rather than by the formula:
- GMST at 2000-01-01 12h UT1 (MJD 51544.5) is 18.697374558 h (the IAU
  1982 expression's constant term, as tabulated by the USNO);
- one solar day advances GMST by 2 pi (1.00273790935 - 1) rad (the
  ratio of mean sidereal to solar time);
- the equation of the equinoxes is the nutation in longitude times
  cos(obliquity); its two leading terms (-17.20 arcsec sin(node of the
  Moon), -1.32 arcsec sin(2 x mean longitude of the Sun)) bound it by
  18.52 arcsec x cos(23.44 deg) = 1.13 s of time, reached within a few per
  cent over an 18.6-year node cycle; it is what separates GAST from
  GMST.
"""
import numpy as np

from oskar_amd import lib


def test_gmst_at_j2000():
    got = lib.oskar_convert_mjd_to_gmst(51544.5)
    want = 18.697374558 * np.pi / 12.0
    assert abs(got - want) < 1e-9


def test_gmst_advances_at_the_sidereal_rate():
    for mjd in (51544.5, 58000.25, 60500.9):
        d = lib.oskar_convert_mjd_to_gmst(mjd + 1.0) - \
            lib.oskar_convert_mjd_to_gmst(mjd)
        d = (d + np.pi) % (2 * np.pi) - np.pi
        assert abs(d - 2 * np.pi * 0.00273790935) < 2e-8
    # Range [0, 2 pi) for dates before and after J2000.
    for mjd in np.linspace(40000.0, 70000.0, 77):
        g = lib.oskar_convert_mjd_to_gmst(float(mjd))
        assert 0.0 <= g < 2 * np.pi


def test_equation_of_the_equinoxes():
    bound = (17.20 + 1.32) / 3600.0 * np.pi / 180.0 * np.cos(
        np.radians(23.44))
    vals = []
    for mjd in np.linspace(51544.5, 51544.5 + 6800.0, 400):
        e = lib.oskar_equation_of_equinoxes_fast(float(mjd))
        gast = lib.oskar_convert_mjd_to_gast_fast(float(mjd))
        gmst = lib.oskar_convert_mjd_to_gmst(float(mjd))
        assert abs(e) < 1.01 * bound
        assert abs(gast - gmst - e) < 1e-15
        vals.append(e)
    # Over one 18.6-year node cycle both extremes come close to the bound.
    assert min(vals) < -0.95 * bound and max(vals) > 0.95 * bound
