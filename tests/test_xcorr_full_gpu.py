"""oskar_cross_correlate at the full sizes of the two direct-correlator
workloads, through the C-ABI, against the FP64 oracle (which reproduces the
reference's correlator bit for bit, see test_oracle.py):

  64 stations x 10,000 point sources, FP64 (ENU in a 5 km disc; sources in
  a 5 deg field; Stokes I power law, Q = U = V = 0; E = CN(0, 1) + I per
  station and source; 100 MHz; no smearing): every one of the 2016
  baselines, max relative difference below 1e-10 (the reference's own
  double-versus-double bound).

  512 stations x 100,000 sources, 16 channels of 10 kHz over 50-350 MHz,
  1 s time average, RADEC phase centre (log-radial layout to 35 km; 30 %
  Gaussian sources, FWHM 5-60 arcsec; full Stokes with |pol| <= 0.2 I):
  the sky's direction cosines and Gaussian parameters come from the
  library's own GPU sky functions (checked against the oracle first), the
  Jones matrices are CN(0, 1) + I; the 276 baselines among 24 stations
  (the 8 outermost included) are compared in every channel:
    - FP64: relative RMS below 1e-12;
    - FP32: relative RMS no more than 1.5 x that of the reference's own
      FP32 algorithm (the oracle's FP32 path, Kahan-summed as the
      reference's CPU code) on the same rounded inputs, both measured
      against FP64 (measured on an MI355X: 3.3e-8 and 3.0e-8).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from gpu_util import OSKAR_COORDS_RADEC, sync_check
from oskar_amd import capi, lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_DOUBLE, OSKAR_GPU,
                           OSKAR_MATRIX, OSKAR_SINGLE, Jones, Mem, Sky,
                           Telescope, check_status)

pytestmark = pytest.mark.gpu

RA, DEC, I_, Q_, U_, V_, MAJ, MIN, PA = 1, 2, 3, 4, 5, 6, 10, 11, 12
EXT_A, EXT_B, EXT_C, L_, M_, N_ = 0x100, 0x101, 0x102, 0x103, 0x104, 0x105
C0 = 299792458.0


def power_law(rng, n, index=-1.8, lo=0.01, hi=10.0):
    """Inverse CDF of a power law on [lo, hi]
    (oskar/math/src/oskar_random_power_law.c:15-36)."""
    b0 = index + 1.0
    u = rng.uniform(0, 1, n)
    return ((hi ** b0 - lo ** b0) * u + lo ** b0) ** (1.0 / b0)


def random_jones(rng, ns, nsrc):
    """E = CN(0, 1) + I per (station, source), float32 values, as reals
    [station][source][a, b, c, d re/im]."""
    j = rng.standard_normal(ns * nsrc * 8, dtype=np.float32)
    j *= np.float32(1 / math.sqrt(2))
    j = j.reshape(ns, nsrc, 8)
    j[:, :, 0] += 1
    j[:, :, 6] += 1
    return j


def gpu_xcorr(prec, gauss, jones_reals, ns, nsrc, flux, dirs, ext, uvw, xy,
              freq, gast, bw=0.0, tavg=0.0, ra0=0.0, dec0=0.0, sky_cols=None):
    P = OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE
    dt = np.float64 if prec == "d" else np.float32
    jt = P | OSKAR_COMPLEX | OSKAR_MATRIX
    jones = Jones(jt, OSKAR_GPU, ns, nsrc)
    jones.set(np.asarray(jones_reals, dt))
    tel = Telescope(P, OSKAR_GPU, ns)
    tel.set_station_ecef(xy[0], xy[1], np.zeros(ns))
    lib.oskar_telescope_set_channel_bandwidth(tel.ptr, float(bw))
    lib.oskar_telescope_set_time_average(tel.ptr, float(tavg))
    lib.oskar_telescope_set_phase_centre(tel.ptr, OSKAR_COORDS_RADEC,
                                         float(ra0), float(dec0))
    if sky_cols is not None:
        f_m, d_m, e_m = sky_cols
    else:
        f_m = [Mem.from_numpy(x, P) for x in flux]
        d_m = [Mem.from_numpy(x, P) for x in dirs]
        e_m = [Mem.from_numpy(x, P) for x in ext]
    u_m = [Mem.from_numpy(x, P) for x in uvw]
    out = []
    for f in np.atleast_1d(freq):
        vis = Mem.from_numpy(np.zeros(ns * (ns - 1) // 2 * 8), jt)
        s = C.c_int(0)
        lib.oskar_cross_correlate(
            0, int(gauss), nsrc, jones.ptr,
            capi.MEMA(*[m.ptr for m in f_m]),
            capi.MEM3(*[m.ptr for m in d_m]),
            capi.MEM3(*[m.ptr for m in e_m]), tel.ptr,
            capi.MEM3(*[m.ptr for m in u_m]), float(gast), float(f), 0,
            vis.ptr, C.byref(s))
        check_status(s, "oskar_cross_correlate")
        sync_check("oskar_cross_correlate %s %.0f Hz" % (prec, f))
        out.append(vis.numpy().reshape(-1, 8))
    return out


def test_xcorr_full_64_stations_10k_sources_fp64(hip):
    print("direct correlator 64 x 10k FP64", flush=True)
    rng = np.random.default_rng(31)
    ns, nsrc = 64, 10_000
    r = 5000 * np.sqrt(rng.uniform(0, 1, ns))
    th = rng.uniform(0, 2 * np.pi, ns)
    uvw = [r * np.cos(th), r * np.sin(th), np.zeros(ns)]
    rho = np.radians(5.0) * np.sqrt(rng.uniform(0, 1, nsrc))
    phi = rng.uniform(0, 2 * np.pi, nsrc)
    l, m = rho * np.cos(phi), rho * np.sin(phi)
    dirs = [l, m, np.sqrt(1 - l * l - m * m)]
    z = np.zeros(nsrc)
    flux = [power_law(rng, nsrc), z, z, z]
    J = random_jones(rng, ns, nsrc).astype(np.float64)
    (got,) = gpu_xcorr("d", 0, J, ns, nsrc, flux, dirs, [z, z, z], uvw,
                       [np.zeros(ns)] * 2, 100e6, 0.0)
    want = oracle.cross_correlate(
        "double", 0, 0, 1, J.ravel(), *flux, *dirs, z, z, z, *uvw,
        np.zeros(ns), np.zeros(ns), freq=100e6).reshape(-1, 8)
    err = np.abs(got - want).max() / np.abs(want).max()
    print("  max |dV| / max |V|: %.3g" % err, flush=True)
    assert err < 1e-10


def full_size_inputs():
    rng = np.random.default_rng(32)
    ns, nsrc = 512, 100_000
    # Log-radial layout to 35 km (r_min 100 m).
    r = np.exp(rng.uniform(np.log(100.0), np.log(35000.0), ns))
    th = rng.uniform(0, 2 * np.pi, ns)
    e, n = r * np.cos(th), r * np.sin(th)
    uvw = [e, n, rng.uniform(-50, 50, ns)]
    xy = [e, n]
    ra0, dec0, gast = 0.4, -0.5, 1.3
    # Sources in a 5 deg field about the phase centre.
    rho = np.radians(5.0) * np.sqrt(rng.uniform(0, 1, nsrc))
    phi = rng.uniform(0, 2 * np.pi, nsrc)
    l, m = np.sin(rho) * np.cos(phi), np.sin(rho) * np.sin(phi)
    nn = np.sqrt(1 - l * l - m * m)
    dec = np.arcsin(nn * math.sin(dec0) + m * math.cos(dec0))
    ra = ra0 + np.arctan2(l, nn * math.cos(dec0) - m * math.sin(dec0))
    I = power_law(rng, nsrc)
    pol = rng.uniform(0, 0.2, nsrc) * I
    d = rng.normal(size=(3, nsrc))
    d /= np.linalg.norm(d, axis=0)
    gauss = rng.uniform(0, 1, nsrc) < 0.3
    maj = np.where(gauss, np.radians(rng.uniform(5, 60, nsrc) / 3600), 0.0)
    mn = maj * rng.uniform(0.2, 1.0, nsrc)
    pa = rng.uniform(0, np.pi, nsrc)
    sky = {RA: ra, DEC: dec, I_: I, Q_: pol * d[0], U_: pol * d[1],
           V_: pol * d[2], MAJ: maj, MIN: mn, PA: pa}
    # 24 stations for the oracle: the 8 outermost and 16 others.
    order = np.argsort(r)
    sub = np.sort(np.concatenate([order[-8:], rng.choice(order[:-8], 16,
                                                        replace=False)]))
    return (rng, ns, nsrc, uvw, xy, ra0, dec0, gast, sky, sub)


def subset_rows(ns, sub):
    """Rows of the full baseline list (q < p, index q(N-1) - (q-1)q/2 +
    p - q - 1 of define_correlate_utils.h:42-43) for the oracle's baselines
    among the stations in `sub`, in the oracle's order."""
    rows = []
    for i in range(len(sub)):
        for j in range(i + 1, len(sub)):
            q, p = sub[i], sub[j]
            rows.append(q * (ns - 1) - (q - 1) * q // 2 + p - q - 1)
    return np.array(rows)


def rel_rms(got, want):
    return float(np.sqrt(np.sum((got - want) ** 2) / np.sum(want ** 2)))


def test_xcorr_full_512_stations_100k_gaussian_smearing(hip):
    print("direct correlator 512 x 100k, Gaussian + smearing, 16 ch",
          flush=True)
    rng, ns, nsrc, uvw, xy, ra0, dec0, gast, sky, sub = full_size_inputs()
    freqs = np.linspace(50e6, 350e6, 16)
    bw, tavg = 10e3, 1.0

    # Sky parameters on the GPU (FP64 sky), checked against the oracle.
    gsky = Sky(OSKAR_DOUBLE, OSKAR_GPU, nsrc)
    for t, v in sky.items():
        gsky.set_column(t, v)
    s = C.c_int(0)
    lib.oskar_sky_evaluate_relative_directions(gsky.ptr, ra0, dec0,
                                               C.byref(s))
    nf = C.c_int(0)
    lib.oskar_sky_evaluate_gaussian_source_parameters(
        gsky.ptr, 0, ra0, dec0, C.byref(nf), C.byref(s))
    check_status(s, "sky parameters")
    sync_check("sky parameters")
    assert nf.value == 0
    dirs = oracle.relative_directions("d", sky[RA], sky[DEC], ra0, dec0)
    ext = oracle.gaussian_parameters("d", sky[RA], sky[DEC], sky[MAJ],
                                     sky[MIN], sky[PA], ra0, dec0)[:3]
    for col, want in zip((L_, M_, N_), dirs):
        assert np.abs(gsky.get_column(col) - want).max() < 1e-15
    for col, want in zip((EXT_A, EXT_B, EXT_C), ext):
        assert np.all(np.abs(gsky.get_column(col) - want) <=
                      1e-9 * np.abs(ext[0] + ext[2]))
    flux = [sky[k] for k in (I_, Q_, U_, V_)]
    sky_cols = ([gsky.column(t) for t in (I_, Q_, U_, V_)],
                [gsky.column(t) for t in (L_, M_, N_)],
                [gsky.column(t) for t in (EXT_A, EXT_B, EXT_C)])

    J32 = random_jones(rng, ns, nsrc)
    rows = subset_rows(ns, sub)
    Jsub = np.ascontiguousarray(J32[sub]).astype(np.float64).ravel()
    kw = dict(channel_bw=bw, time_avg=tavg, ra0=ra0, dec0=dec0, gast=gast)
    sub_uvw = [x[sub] for x in uvw]
    sub_xy = [x[sub] for x in xy]

    # FP64: the GPU sky columns feed the correlator directly.
    got64 = gpu_xcorr("d", 1, J32, ns, nsrc, None, None, None, uvw, xy,
                      freqs, gast, bw, tavg, ra0, dec0, sky_cols=sky_cols)
    worst64 = 0.0
    want64 = []
    for k, f in enumerate(freqs):
        w = oracle.cross_correlate(
            "double", 0, 1, 1, Jsub, *[x for x in flux], *dirs, *ext,
            *sub_uvw, *sub_xy, freq=f, **kw).reshape(-1, 8)
        want64.append(w)
        worst64 = max(worst64, rel_rms(got64[k][rows], w))
    del got64, gsky, sky_cols
    print("  FP64 rel RMS (worst channel): %.3g" % worst64, flush=True)
    assert worst64 < 1e-12

    # FP32: inputs rounded to float; the reference's FP32 algorithm on the
    # same values is the yardstick.
    r32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    f32 = [r32(x) for x in flux]
    d32 = [r32(x) for x in dirs]
    e32 = [r32(x) for x in ext]
    got32 = gpu_xcorr("f", 1, J32, ns, nsrc, f32, d32, e32, uvw, xy, freqs,
                      gast, bw, tavg, ra0, dec0)
    worst_ours = worst_ref = 0.0
    for k, f in enumerate(freqs):
        ref32 = oracle.cross_correlate(
            "single", 0, 1, 1, Jsub, *f32, *d32, *e32, *[r32(x) for x in
                                                      sub_uvw],
            *[r32(x) for x in sub_xy], freq=f, **kw).reshape(-1, 8)
        w = oracle.cross_correlate(
            "double", 0, 1, 1, Jsub, *f32, *d32, *e32, *[r32(x) for x in
                                                      sub_uvw],
            *[r32(x) for x in sub_xy], freq=f, **kw).reshape(-1, 8)
        e_ours = rel_rms(got32[k][rows].astype(np.float64), w)
        e_ref = rel_rms(ref32.astype(np.float64), w)
        worst_ours, worst_ref = max(worst_ours, e_ours), max(worst_ref,
                                                             e_ref)
        assert e_ours <= 1.5 * e_ref + 1e-7, (f, e_ours, e_ref)
    print("  FP32 rel RMS vs FP64 (worst channel): ours %.3g, reference "
          "FP32 algorithm %.3g" % (worst_ours, worst_ref), flush=True)
