"""Textbook far field of a thin centre-fed dipole, in vector form, for
closed-form checks of the aperture-array station beam's dipole branch.

A dipole of half-length h along the unit vector d radiates, towards the
unit direction s, the field

    E(s) = -t(s.d) * (d - (d.s) s),   t(c) = (cos(k h c) - cos(k h)) / (1 - c^2)

(|E| = |cos(k h cos psi) - cos k h| / sin psi with psi the angle between s
and d: nulls along the dipole's axis, the maximum broadside).  The overall
sign is the reference's convention (define_evaluate_dipole_pattern.h:3-15
writes E_theta = -cos(phi) cos(theta) t, E_phi = sin(phi) t for a dipole
along phi = 0).  The field is resolved on the spherical unit vectors of s,
theta_hat = (cos th cos ph, cos th sin ph, -sin th) and
phi_hat = (-sin ph, cos ph, 0), with th = atan2(sqrt(x^2 + y^2), z) and
ph = atan2(y, x).

The axes follow from the reference's feed offsets
(oskar_evaluate_station_beam_aperture_array.c:170-177: orientation_x =
euler_x + pi/2, orientation_y = euler_y; oskar_element_evaluate.c: delta
phi = pi/2 - orientation): the X dipole lies along
(cos euler_x, sin euler_x, 0), the Y dipole along
(sin euler_y, -cos euler_y, 0).  At zero Euler angles these are the x and
y axes of the station (docs/theory/theory.rst, "Element Factors": X and Y
dipoles nominally along x and y), so the X feed is null towards the +-x
horizon and the Y feed towards +-y; at 90 degrees the nulls swap.
"""
import numpy as np


def feed_axes(euler_x, euler_y):
    dx = np.array([np.cos(euler_x), np.sin(euler_x), 0.0])
    dy = np.array([np.sin(euler_y), -np.cos(euler_y), 0.0])
    return dx, dy


def dipole_field(d, x, y, z, kh):
    """(E_theta, E_phi) of a dipole along d at directions (x, y, z)."""
    s = np.stack([x, y, z])
    c = d @ s
    one_m = 1.0 - c * c
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(one_m > 0, (np.cos(kh * c) - np.cos(kh)) / one_m, 0.0)
    e = -t * (d[:, None] - c * s)
    th = np.arctan2(np.hypot(x, y), z)
    ph = np.arctan2(y, x)
    that = np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph),
                     -np.sin(th)])
    phat = np.stack([-np.sin(ph), np.cos(ph), np.zeros_like(ph)])
    return (e * that).sum(0), (e * phat).sum(0)


def dipole_matrix(euler_x, euler_y, x, y, z, kh, swap_xy=0):
    """The 2x2 element pattern [[X_theta, X_phi], [Y_theta, Y_phi]] as
    complex (n, 2, 2); rows swapped when swap_xy."""
    dx, dy = feed_axes(euler_x, euler_y)
    xt, xp = dipole_field(dx, x, y, z, kh)
    yt, yp = dipole_field(dy, x, y, z, kh)
    rows = [[yt, yp], [xt, xp]] if swap_xy else [[xt, xp], [yt, yp]]
    return np.array(rows, dtype=complex).transpose(2, 0, 1)
