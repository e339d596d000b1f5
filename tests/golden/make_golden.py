"""Writes the golden fixtures under tests/golden/ from the reference's own
CPU kernels, compiled from its sources by ``make -C oracle ref`` into
oracle/_ref/liboskar_ref.so.

Run in a container that has the reference tree (the GPU box does not):
    make -C oracle ref && python tests/golden/make_golden.py

Fixtures hold inputs and outputs only (numpy .npz, float64 inputs, outputs
in the precision of the kernel that produced them):

  xcorr_ref.npz      reference test shapes and value ranges
                     (oskar/correlate/test/Test_cross_correlate.cpp:53-115):
                     19 stations, 277 sources.
  xcorr_long.npz     13 stations, 199 sources on long baselines (up to
                     30 km, 150 MHz, Gaussian widths of 5-60 arcsec), so
                     that every smearing factor differs clearly from 1.
      Each holds every combination of {scalar, matrix} x {point, Gaussian}
      x {bandwidth smearing off/on} x {time smearing off/on} x
      {CASA convention 0/1} x {float, double}, produced by
      oskar_cross_correlate{,_scalar}_{point,gaussian}_omp_{f,d}.
  dftw.npz           oskar_dftw CPU bodies (define_dftw_c2c.h, define_dftw_m2m.h)
                     for {c2c, m2m} x {2D, 3D} x {float, double} x
                     {no data index, two element types}.
  beam_parts.npz     DFT weights, theta/phi, dipole patterns, Jones K.
  acorr.npz          auto-correlation CPU bodies (define_auto_correlate.h)
                     at the reference test's shapes and value ranges
                     (oskar/correlate/test/Test_auto_correlate.cpp:40-63):
                     50 stations, 277 sources, {scalar, matrix} x
                     {float, double}, added at offset 3 into random vis
                     (inputs stored as float32, exact in both precisions).
  sky_clip.npz       horizon mask bodies (define_update_horizon_mask.h) for
                     three (hour angle, Dec, latitude) cases on 3000
                     directions, 600 of them within ~1e-7 of the horizon
                     plane, {float, double}, each case alone and ORed over
                     the three; the masked row copy
                     (define_sky_copy_source_data.h) of a 7-column table
                     with its valid-column counts, {float, double}.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))
import oracle  # noqa: E402

C0 = 299792458.0
FLT_MAX = float(np.finfo(np.float32).max)


def P(a):
    return a.ctypes.data_as(C.c_void_p)


def derived(freq, channel_bw, time_avg, gast, ra0, dec0, uv_min, uv_max,
            uv_wavelengths):
    """Derived scalars of oskar_cross_correlate.c:42-71 (RADEC centre)."""
    freq = abs(freq)
    inv_wl = freq / C0
    frac_bw = channel_bw / freq
    gha0 = gast - ra0
    if not uv_wavelengths:
        uv_min *= inv_wl
        uv_max *= inv_wl
    if uv_max < 0.0 or uv_max > FLT_MAX:
        uv_max = FLT_MAX
    return inv_wl, frac_bw, time_avg, gha0, dec0, uv_min, uv_max


def xcorr_case(ref, inp, precision, matrix, gauss, casa, bs, ts):
    dt = np.float64 if precision == "d" else np.float32
    ns, nsrc = inp["u"].size, inp["l"].size
    nb = ns * (ns - 1) // 2
    chan_bw = inp["channel_bw"] if bs else 0.0
    t_avg = inp["time_avg"] if ts else 0.0
    inv_wl, frac_bw, t_avg, gha0, dec0, uv_min, uv_max = derived(
        inp["freq"], chan_bw, t_avg, inp["gast"], inp["ra0"], inp["dec0"],
        0.0, -1.0, False)
    a = {k: np.ascontiguousarray(inp[k], dtype=dt) for k in
         ("I", "Q", "U", "V", "l", "m", "n", "a", "b", "c", "u", "v", "w",
          "x", "y")}
    jones = np.ascontiguousarray(
        inp["jones_m"] if matrix else inp["jones_s"], dtype=dt)
    vis = np.zeros(nb * (8 if matrix else 2), dtype=dt)
    R = C.c_double if precision == "d" else C.c_float
    scal = [R(v) for v in (uv_min, uv_max, inv_wl, frac_bw, t_avg, gha0,
                           dec0)]
    st = [P(a[k]) for k in ("u", "v", "w", "x", "y")]
    if matrix:
        name = "oskar_cross_correlate_%s_omp_%s" % (
            "gaussian" if gauss else "point", precision)
        args = [C.c_int(casa), C.c_int(nsrc), C.c_int(ns), C.c_int(0),
                P(jones)] + [P(a[k]) for k in ("I", "Q", "U", "V", "l", "m",
                                               "n")]
    else:
        name = "oskar_cross_correlate_scalar_%s_omp_%s" % (
            "gaussian" if gauss else "point", precision)
        args = [C.c_int(casa), C.c_int(nsrc), C.c_int(ns), C.c_int(0),
                P(jones), P(a["I"]), P(a["l"]), P(a["m"]), P(a["n"])]
    if gauss:
        args += [P(a["a"]), P(a["b"]), P(a["c"])]
    args += st + scal + [P(vis)]
    getattr(ref, name)(*args)
    return vis


def make_xcorr_inputs(rng, ns, nsrc, long_baselines):
    inp = {}
    if long_baselines:
        inp["jones_m"] = rng.normal(size=ns * nsrc * 8)
        inp["jones_s"] = rng.normal(size=ns * nsrc * 2)
        for k in ("u", "v"):
            inp[k] = rng.uniform(-15000.0, 15000.0, ns)
        inp["w"] = rng.uniform(-100.0, 100.0, ns)
        inp["x"] = rng.uniform(-15000.0, 15000.0, ns)
        inp["y"] = rng.uniform(-15000.0, 15000.0, ns)
        I_ = rng.uniform(0.1, 10.0, nsrc)
        pol = rng.uniform(0.0, 0.2, nsrc) * I_
        d = rng.normal(size=(3, nsrc))
        d /= np.linalg.norm(d, axis=0)
        inp["I"], inp["Q"], inp["U"], inp["V"] = I_, pol * d[0], pol * d[1], \
            pol * d[2]
        r = np.radians(5.0) * np.sqrt(rng.uniform(size=nsrc))
        th = rng.uniform(0, 2 * np.pi, nsrc)
        inp["l"], inp["m"] = r * np.cos(th), r * np.sin(th)
        inp["n"] = np.sqrt(1.0 - inp["l"] ** 2 - inp["m"] ** 2)
        # Gaussian parameters from FWHM 5-60 arcsec (major), minor <= major,
        # PA uniform; a,b,c as oskar_sky_evaluate_gaussian_source_parameters.c
        # :171-178 in its small-angle limit.
        fwhm_maj = np.radians(rng.uniform(5.0, 60.0, nsrc) / 3600.0)
        fwhm_min = fwhm_maj * rng.uniform(0.2, 1.0, nsrc)
        pa = rng.uniform(0.0, np.pi, nsrc)
        f = np.pi ** 2 / (4.0 * np.log(2.0))
        inv_maj = 0.5 * fwhm_maj ** 2 * f
        inv_min = 0.5 * fwhm_min ** 2 * f
        s2, c2, sn2 = np.sin(pa) ** 2, np.cos(pa) ** 2, np.sin(2 * pa)
        inp["a"] = c2 * inv_min + s2 * inv_maj
        inp["b"] = -sn2 * inv_min / 2 + sn2 * inv_maj / 2
        inp["c"] = s2 * inv_min + c2 * inv_maj
        inp["freq"], inp["channel_bw"], inp["time_avg"] = 150e6, 100e3, 10.0
        inp["gast"], inp["ra0"], inp["dec0"] = 1.3, 0.4, -0.5
    else:
        # Ranges of Test_cross_correlate.cpp:93-115.
        inp["jones_m"] = rng.uniform(1.0, 5.0, ns * nsrc * 8)
        inp["jones_s"] = rng.uniform(1.0, 5.0, ns * nsrc * 2)
        for k in ("u", "v", "w"):
            inp[k] = rng.uniform(1.0, 5.0, ns)
        for k in ("x", "y"):
            inp[k] = rng.uniform(0.1, 1000.0, ns)
        inp["I"] = rng.uniform(1.0, 2.0, nsrc)
        inp["Q"] = rng.uniform(0.1, 1.0, nsrc)
        inp["U"] = rng.uniform(0.1, 0.5, nsrc)
        inp["V"] = rng.uniform(0.1, 0.2, nsrc)
        for k in ("l", "m", "n"):
            inp[k] = rng.uniform(0.1, 0.9, nsrc)
        for k in ("a", "b", "c"):
            inp[k] = rng.uniform(0.1e-6, 0.2e-6, nsrc)
        # Test_cross_correlate.cpp:143-144,256-280: gast 1, 100 MHz,
        # 10 s averaging, 10 kHz channels.
        inp["freq"], inp["channel_bw"], inp["time_avg"] = 100e6, 10e3, 10.0
        inp["gast"], inp["ra0"], inp["dec0"] = 1.0, 0.0, 0.0
    return inp


def make_xcorr(ref, path, seed, ns, nsrc, long_baselines):
    rng = np.random.default_rng(seed)
    inp = make_xcorr_inputs(rng, ns, nsrc, long_baselines)
    out = dict(inp)
    for prec in ("f", "d"):
        for matrix in (0, 1):
            for gauss in (0, 1):
                for casa in (0, 1):
                    for bs in (0, 1):
                        for ts in (0, 1):
                            key = "vis_%s_%s_%s_c%d_bs%d_ts%d" % (
                                prec, "m" if matrix else "s",
                                "g" if gauss else "p", casa, bs, ts)
                            out[key] = xcorr_case(ref, inp, prec, matrix,
                                                  gauss, casa, bs, ts)
    np.savez_compressed(path, **out)


def make_dftw(ref, path, seed):
    rng = np.random.default_rng(seed)
    nin, nout, ntypes = 16, 256, 2
    out = {}
    out["x_in"] = rng.uniform(-10, 10, nin)
    out["y_in"] = rng.uniform(-10, 10, nin)
    out["z_in"] = rng.uniform(-1, 1, nin)
    ph = rng.uniform(0, 2 * np.pi, nin)
    out["weights"] = np.stack([np.cos(ph), np.sin(ph)], 1).ravel()
    th = rng.uniform(0, np.pi / 2, nout)
    phi = rng.uniform(0, 2 * np.pi, nout)
    out["x_out"] = np.sin(th) * np.cos(phi)
    out["y_out"] = np.sin(th) * np.sin(phi)
    out["z_out"] = np.cos(th)
    out["idx"] = rng.integers(0, ntypes, nin).astype(np.int32)
    out["data_m_full"] = rng.normal(size=nin * nout * 8)
    out["data_c_full"] = rng.normal(size=nin * nout * 2)
    out["data_m_typed"] = rng.normal(size=ntypes * nout * 8)
    out["data_c_typed"] = rng.normal(size=ntypes * nout * 2)
    out["wavenumber"] = np.array(2 * np.pi * 150e6 / C0)
    for prec, dt, fp in (("f", np.float32, C.c_float),
                         ("d", np.float64, C.c_double)):
        cast = {k: np.ascontiguousarray(out[k], dtype=dt) for k in
                ("x_in", "y_in", "z_in", "weights", "x_out", "y_out", "z_out",
                 "data_m_full", "data_c_full", "data_m_typed",
                 "data_c_typed")}
        for kind in ("c2c", "m2m"):
            for is3d in (0, 1):
                for typed in (0, 1):
                    nr = 8 if kind == "m2m" else 2
                    data = cast["data_%s_%s" % ("m" if kind == "m2m" else "c",
                                                "typed" if typed else "full")]
                    res = np.zeros(nout * nr, dtype=dt)
                    fn = getattr(ref, "ref_dftw_%s_%s_%s" % (
                        kind, "3d" if is3d else "2d",
                        "double" if prec == "d" else "float"))
                    fn(C.c_int(nin), fp(float(out["wavenumber"])),
                       P(cast["weights"]), P(cast["x_in"]), P(cast["y_in"]),
                       P(cast["z_in"]) if is3d else None, C.c_int(0),
                       C.c_int(nout), P(cast["x_out"]), P(cast["y_out"]),
                       P(cast["z_out"]) if is3d else None,
                       P(out["idx"]) if typed else None, P(data),
                       C.c_int(1), C.c_int(1), C.c_int(0), P(res),
                       fp(1.0 / nin), C.c_int(0))
                    out["out_%s_%s_%dd_%s" % (prec, kind, 3 if is3d else 2,
                                              "typed" if typed else "full")] \
                        = res
    np.savez_compressed(path, **out)


def make_beam_parts(ref, path, seed):
    rng = np.random.default_rng(seed)
    out = {}
    n = 64
    x = rng.uniform(-19, 19, n)
    y = rng.uniform(-19, 19, n)
    z = rng.uniform(-0.5, 0.5, n)
    cable = rng.uniform(-0.01, 0.01, n)
    k = 2 * np.pi * 120e6 / C0
    b = np.array([0.3, -0.2, np.sqrt(1 - 0.13)])
    out.update(ew_x=x, ew_y=y, ew_z=z, ew_cable=cable, ew_k=np.array(k),
               ew_beam=b)
    w = np.zeros(2 * n)
    ref.ref_element_weights_dft_double(
        C.c_int(n), P(x), P(y), P(z), P(cable), C.c_double(k),
        C.c_double(b[0]), C.c_double(b[1]), C.c_double(b[2]), P(w))
    out["ew_weights"] = w
    # theta / phi with an extra point at the pole.
    m = 300
    th = rng.uniform(0, np.pi / 2, m)
    ph = rng.uniform(-np.pi, np.pi, m)
    dx, dy, dz = np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)
    out.update(tp_x=dx, tp_y=dy, tp_z=dz)
    theta = np.zeros(m + 1)
    p1 = np.zeros(m + 1)
    p2 = np.zeros(m + 1)
    d1, d2 = -0.3, np.pi / 2 - 0.1
    ref.ref_theta_phi_double(C.c_int(0), C.c_int(m), P(dx), P(dy), P(dz),
                             C.c_int(1), C.c_double(d1), C.c_double(d2),
                             P(theta), P(p1), P(p2))
    out.update(tp_d1=np.array(d1), tp_d2=np.array(d2), tp_theta=theta,
               tp_phi1=p1, tp_phi2=p2)
    # Dipole pattern on those directions (half-wave at 120 MHz).
    kL = 0.5 * (C0 / 120e6) * (np.pi * 120e6 / C0)
    for swap in (0, 1):
        pat = np.zeros(8 * (m + 1))
        ref.ref_dipole_pattern_double(C.c_int(m + 1), P(theta), P(p1), P(p2),
                                      C.c_double(kL), C.c_double(np.cos(kL)),
                                      C.c_int(swap), C.c_int(0), P(pat))
        out["dip_m_swap%d" % swap] = pat
    pat = np.zeros(2 * (m + 1))
    ref.ref_dipole_pattern_scalar_double(
        C.c_int(m + 1), P(theta), P(p1), P(p2), C.c_double(kL),
        C.c_double(np.cos(kL)), C.c_int(0), C.c_int(0), P(pat))
    out["dip_s"] = pat
    out["dip_kL"] = np.array(kL)
    # Jones K with a source filter.
    ns, nsrc = 7, 120
    u = rng.uniform(-3000, 3000, ns)
    v = rng.uniform(-3000, 3000, ns)
    ww = rng.uniform(-50, 50, ns)
    r = np.radians(10.0) * np.sqrt(rng.uniform(size=nsrc))
    t = rng.uniform(0, 2 * np.pi, nsrc)
    l, mm = r * np.cos(t), r * np.sin(t)
    nn = np.sqrt(1 - l * l - mm * mm)
    I_ = rng.uniform(0.0, 2.0, nsrc)
    kk = 2 * np.pi * 80e6 / C0
    jk = np.zeros(2 * ns * nsrc)
    ref.ref_jones_K_double(C.c_int(nsrc), P(l), P(mm), P(nn), C.c_int(ns),
                           P(u), P(v), P(ww), C.c_double(kk), P(I_),
                           C.c_double(0.2), C.c_double(1.8), C.c_int(0),
                           P(jk))
    out.update(k_u=u, k_v=v, k_w=ww, k_l=l, k_m=mm, k_n=nn, k_I=I_,
               k_k=np.array(kk), k_fmin=np.array(0.2), k_fmax=np.array(1.8),
               k_jones=jk)
    np.savez_compressed(path, **out)


def make_acorr(ref, path, seed):
    rng = np.random.default_rng(seed)
    ns, nsrc, off = 50, 277, 3
    out = dict(num_stations=ns, num_sources=nsrc, offset_out=off)
    f32 = np.float32
    out["I"] = rng.uniform(1.0, 2.0, nsrc).astype(f32)
    out["Q"] = rng.uniform(0.1, 1.0, nsrc).astype(f32)
    out["U"] = rng.uniform(0.1, 0.5, nsrc).astype(f32)
    out["V"] = rng.uniform(0.1, 0.2, nsrc).astype(f32)
    for matrix in (0, 1):
        nr = 8 if matrix else 2
        jones = rng.uniform(1.0, 5.0, ns * nsrc * nr).astype(f32)
        vis0 = rng.uniform(-1.0, 1.0, (off + ns + 2) * nr).astype(f32)
        out["jones_%d" % matrix] = jones
        out["vis0_%d" % matrix] = vis0
        for prec in ("f", "d"):
            dt = np.float32 if prec == "f" else np.float64
            name = "ref_acorr_%s%s" % ("" if matrix else "scalar_",
                                       "float" if prec == "f" else "double")
            fn = getattr(ref, name)
            fn.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 6
            fl = [np.ascontiguousarray(out[k], dt) for k in "IQUV"]
            j = np.ascontiguousarray(jones, dt)
            vis = np.array(vis0, dtype=dt, copy=True)
            fn(nsrc, ns, off, *[P(a) for a in fl], P(j), P(vis))
            out["vis_%d_%s" % (matrix, prec)] = vis
    np.savez_compressed(path, **out)


def horizon_factors(ha0, dec0, lat):
    """(ll, mm, nn) as oskar/sky/src/oskar_update_horizon_mask.c:37-44
    computes them (libm, double)."""
    import math
    return (math.cos(lat) * math.sin(ha0),
            math.sin(lat) * math.cos(dec0) -
            math.cos(lat) * math.cos(ha0) * math.sin(dec0),
            math.sin(lat) * math.sin(dec0) +
            math.cos(lat) * math.cos(ha0) * math.cos(dec0))


def make_sky_clip(ref, path, seed):
    rng = np.random.default_rng(seed)
    f32 = np.float32
    cases = np.array([[0.7, -0.45, -0.47], [2.9, 0.2, 0.65],
                      [-1.3, -1.1, -0.9]])
    num = 3000
    v = rng.normal(size=(num, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    # 200 directions near each case's horizon plane: project out the plane
    # normal, renormalise, tilt by up to 1e-7.
    for k, (ha0, dec0, lat) in enumerate(cases):
        f = np.array(horizon_factors(ha0, dec0, lat))
        w = v[k * 200:(k + 1) * 200]
        w -= np.outer(w @ f, f)
        w /= np.linalg.norm(w, axis=1)[:, None]
        w += np.outer(rng.uniform(-1e-7, 1e-7, 200), f)
        v[k * 200:(k + 1) * 200] = w
    v = v.astype(f32)
    out = dict(l=v[:, 0], m=v[:, 1], n=v[:, 2], cases=cases)
    for prec in ("f", "d"):
        dt = f32 if prec == "f" else np.float64
        fn = getattr(ref, "ref_update_horizon_mask_%s" % (
            "float" if prec == "f" else "double"))
        ct = C.c_float if prec == "f" else C.c_double
        fn.argtypes = [C.c_int] + [C.c_void_p] * 3 + [ct] * 3 + [C.c_void_p]
        l, m, n = (np.ascontiguousarray(out[k], dt) for k in "lmn")
        acc = np.zeros(num, np.int32)
        for k, (ha0, dec0, lat) in enumerate(cases):
            ff = horizon_factors(ha0, dec0, lat)
            one = np.zeros(num, np.int32)
            fn(num, P(l), P(m), P(n), *[ct(x) for x in ff], P(one))
            fn(num, P(l), P(m), P(n), *[ct(x) for x in ff], P(acc))
            out["mask_%s_%d" % (prec, k)] = one
        out["mask_%s_or" % prec] = acc
    # Row copy: 7 columns, input capacity 3072, output capacity 4096.
    ncol, cap_in, cap_out, nfix = 7, 3072, 4096, 20
    mask = out["mask_d_or"].copy()
    mask[rng.integers(0, num, 300)] = 0
    idx = np.zeros(num + 1, np.int32)
    idx[1:] = np.cumsum(mask)
    out.update(copy_mask=mask, copy_indices=idx, ncol=ncol, cap_in=cap_in,
               cap_out=cap_out, nfix=nfix)
    out["table_in"] = rng.uniform(-10, 10, ncol * cap_in).astype(f32)
    out["valid_in"] = rng.integers(0, 4, cap_in * nfix).astype(np.int32)
    for prec in ("f", "d"):
        dt = f32 if prec == "f" else np.float64
        fn = getattr(ref, "ref_sky_copy_source_data_%s" % (
            "float" if prec == "f" else "double"))
        fn.argtypes = [C.c_int] * 3 + [C.c_void_p] * 2 + [C.c_int] + \
            [C.c_void_p] * 4
        tin = np.ascontiguousarray(out["table_in"], dt)
        tout = np.full(ncol * cap_out, -7.0, dt)
        vout = np.full(cap_out * nfix, -1, np.int32)
        fn(num, cap_in, cap_out, P(mask), P(idx), ncol, P(tin), P(tout),
           P(out["valid_in"]), P(vout))
        out["table_out_%s" % prec] = tout
        out["valid_out_%s" % prec] = vout
    np.savez_compressed(path, **out)


def main():
    ref = oracle.ref()
    if ref is None:
        raise SystemExit("oracle/_ref/liboskar_ref.so missing: "
                         "run `make -C oracle ref` first")
    make_xcorr(ref, os.path.join(HERE, "xcorr_ref.npz"), 2, 19, 277, False)
    make_xcorr(ref, os.path.join(HERE, "xcorr_long.npz"), 3, 13, 199, True)
    make_dftw(ref, os.path.join(HERE, "dftw.npz"), 4)
    make_beam_parts(ref, os.path.join(HERE, "beam_parts.npz"), 5)
    make_acorr(ref, os.path.join(HERE, "acorr.npz"), 6)
    make_sky_clip(ref, os.path.join(HERE, "sky_clip.npz"), 7)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
