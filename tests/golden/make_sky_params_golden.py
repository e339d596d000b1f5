"""Writes tests/golden/sky_params.npz from the reference's own
oskar_sky_evaluate_gaussian_source_parameters (compiled with what it calls
into oracle/_ref/liboskar_sky_ref.so by ``make -C oracle ref``) and the 3D
body of its lon/lat -> relative-direction conversion macro.

Run in a container that has the reference tree (the GPU box does not):
    make -C oracle ref && python tests/golden/make_sky_params_golden.py

The fixture holds inputs and outputs only:
  reftest_*   the reference test's sky
              (oskar/sky/test/Test_Sky_evaluate_gaussian_source_parameters
              .cpp:21-47: source i at RA = Dec = 0.001 i deg, FWHM 1200 x
              600 arcsec, PA 30 deg, phase centre (0, 10 deg)), every 16th of
              its 16384 sources;
  rand_*      1500 sources over the whole sphere, phase centre
              (0.3, -0.5) rad: major FWHM log-uniform in [1 arcsec, 10 deg],
              minor = major x U(0, 1), PA uniform in [0, pi); 100 sources
              with a zero minor axis, 100 with both axes zero (skipped),
              Stokes I, Q, U, V columns; evaluated with zero_failed = 1, so
              the outputs include the flux columns after failed sources
              were cleared;
  reldir_*    relative direction cosines of the random sky's positions.
Each has _f (float sky) and _d (double sky) outputs; scratch columns start
at -7 so that skipped and failed sources show their untouched values.
"""
import ctypes as C
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))
import oracle  # noqa: E402

DEG = np.pi / 180.0
ARCSEC = DEG / 3600.0
OSKAR_SINGLE, OSKAR_DOUBLE, OSKAR_CPU = 4, 8, 0
RA, DEC, I_, Q_, U_, V_, MAJ, MIN, PA = 1, 2, 3, 4, 5, 6, 10, 11, 12
EXT_A, EXT_B, EXT_C = 0x100, 0x101, 0x102


def run_reference(ref, prec, cols, ra0, dec0, zero_failed):
    """Builds a reference sky in host memory with the given columns, runs
    the reference function, returns (outputs by column, num_failed)."""
    typ = OSKAR_SINGLE if prec == "f" else OSKAR_DOUBLE
    ct = C.c_float if prec == "f" else C.c_double
    n = len(cols[RA])
    st = C.c_int(0)
    sky = ref.oskar_sky_create(typ, OSKAR_CPU, n, C.byref(st))
    assert st.value == 0

    def col(t):
        p = ref.oskar_sky_column(sky, t, 0, C.byref(st))
        assert st.value == 0
        return np.ctypeslib.as_array(
            C.cast(ref.oskar_mem_void(p), C.POINTER(ct)), (n,))

    for t, v in cols.items():
        col(t)[:] = v
    for t in (EXT_A, EXT_B, EXT_C):
        col(t)[:] = -7.0
    nf = C.c_int(0)
    ref.oskar_sky_evaluate_gaussian_source_parameters(
        sky, zero_failed, ra0, dec0, C.byref(nf), C.byref(st))
    assert st.value == 0
    out = {t: col(t).copy() for t in list(cols) + [EXT_A, EXT_B, EXT_C]}
    ref.oskar_sky_free(sky, C.byref(st))
    return out, nf.value


def main():
    ref = oracle.ref_sky()
    if ref is None:
        raise SystemExit("oracle/_ref/liboskar_sky_ref.so missing: "
                         "run `make -C oracle ref` first")
    out = {}

    # The reference test's sky.
    i = np.arange(0, 16384, 16)
    n = i.size
    rt = {RA: i * DEG * 0.001, DEC: i * DEG * 0.001,
          MAJ: np.full(n, 1200 * ARCSEC), MIN: np.full(n, 600 * ARCSEC),
          PA: np.full(n, 30 * DEG)}
    out["reftest_ra0_dec0"] = np.array([0.0, 10 * DEG])
    for k, v in rt.items():
        out["reftest_in_%d" % k] = v
    for prec in ("f", "d"):
        dt = np.float32 if prec == "f" else np.float64
        res, nf = run_reference(ref, prec, {k: v.astype(dt)
                                            for k, v in rt.items()},
                                0.0, 10 * DEG, 0)
        out["reftest_failed_%s" % prec] = nf
        for t in (EXT_A, EXT_B, EXT_C):
            out["reftest_out_%d_%s" % (t, prec)] = res[t]

    # Random all-sky case.
    rng = np.random.default_rng(11)
    n = 1500
    ra = rng.uniform(0, 2 * np.pi, n)
    dec = np.arcsin(rng.uniform(-1, 1, n))
    maj = 10 ** rng.uniform(0, np.log10(36000.0), n) * ARCSEC
    mn = maj * rng.uniform(0, 1, n)
    pa = rng.uniform(0, np.pi, n)
    mn[:100] = 0.0
    maj[100:200] = 0.0
    mn[100:200] = 0.0
    stokes = rng.uniform(0.1, 10, (4, n))
    rd = {RA: ra, DEC: dec, MAJ: maj, MIN: mn, PA: pa, I_: stokes[0],
          Q_: stokes[1], U_: stokes[2], V_: stokes[3]}
    ra0, dec0 = 0.3, -0.5
    out["rand_ra0_dec0"] = np.array([ra0, dec0])
    for k, v in rd.items():
        out["rand_in_%d" % k] = v
    for prec in ("f", "d"):
        dt = np.float32 if prec == "f" else np.float64
        res, nf = run_reference(ref, prec, {k: v.astype(dt)
                                            for k, v in rd.items()},
                                ra0, dec0, 1)
        out["rand_failed_%s" % prec] = nf
        for t in (EXT_A, EXT_B, EXT_C, I_, Q_, U_, V_):
            out["rand_out_%d_%s" % (t, prec)] = res[t]

        # Relative directions of the same positions (3D macro body, with
        # the float casts of the reference's dispatcher).
        fn = getattr(ref, "ref_lon_lat_to_relative_directions_3d_%s" % (
            "float" if prec == "f" else "double"))
        r, d = ra.astype(dt), dec.astype(dt)
        l, m, nn = (np.zeros(n, dt) for _ in range(3))
        fn(n, r.ctypes.data, d.ctypes.data, ra0, math.cos(dec0),
           math.sin(dec0), l.ctypes.data, m.ctypes.data, nn.ctypes.data)
        out["reldir_l_%s" % prec] = l
        out["reldir_m_%s" % prec] = m
        out["reldir_n_%s" % prec] = nn

    path = os.path.join(HERE, "sky_params.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
