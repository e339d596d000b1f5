"""Drives the C-ABI hot-path functions from numpy arrays (GPU tests)."""
import ctypes as C

import numpy as np

from oskar_amd import capi, lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_DOUBLE, OSKAR_GPU,
                           OSKAR_MATRIX, OSKAR_SINGLE, OSKAR_INT, Jones, Mem,
                           Station, StationWork, Telescope, check_status)

OSKAR_COORDS_RADEC = 2


def prec_code(prec):
    return OSKAR_DOUBLE if prec in ("d", "double") else OSKAR_SINGLE


def sync_check(where):
    lib.oskar_device_synchronize(OSKAR_GPU)
    s = C.c_int(0)
    lib.oskar_device_check_error_hip(C.byref(s))
    check_status(s, where)


def gpu_cross_correlate(prec, casa, gauss, matrix, g, bs, ts, vis_init=None,
                        offset_out=0, uv_filter=None, num_sources=None,
                        entry="oskar_cross_correlate"):
    """oskar_cross_correlate (or the entry point named by `entry`, which
    takes the same arguments) on the GPU for the inputs in dict g (golden
    fixture layout); returns vis as a flat numpy array of reals."""
    P = prec_code(prec)
    ns = g["u"].size
    nsrc = g["l"].size if num_sources is None else num_sources
    jt = P | OSKAR_COMPLEX | (OSKAR_MATRIX if matrix else 0)
    jones = Jones(jt, OSKAR_GPU, ns, g["l"].size)
    jones.set(g["jones_m"] if matrix else g["jones_s"])
    tel = Telescope(P, OSKAR_GPU, ns)
    tel.set_station_ecef(g["x"], g["y"], np.zeros(ns))
    lib.oskar_telescope_set_channel_bandwidth(
        tel.ptr, float(g["channel_bw"]) if bs else 0.0)
    lib.oskar_telescope_set_time_average(
        tel.ptr, float(g["time_avg"]) if ts else 0.0)
    lib.oskar_telescope_set_phase_centre(tel.ptr, OSKAR_COORDS_RADEC,
                                         float(g["ra0"]), float(g["dec0"]))
    if uv_filter is not None:
        s = C.c_int(0)
        lib.oskar_telescope_set_uv_filter(tel.ptr, uv_filter[0], uv_filter[1],
                                          b"Metres", C.byref(s))
        check_status(s, "set_uv_filter")
    flux = [Mem.from_numpy(g[k], P) for k in ("I", "Q", "U", "V")]
    dirs = [Mem.from_numpy(g[k], P) for k in ("l", "m", "n")]
    ext = [Mem.from_numpy(g[k], P) for k in ("a", "b", "c")]
    uvw = [Mem.from_numpy(g[k], P) for k in ("u", "v", "w")]
    nb = ns * (ns - 1) // 2
    nr = 8 if matrix else 2
    if vis_init is None:
        vis_init = np.zeros((offset_out + nb) * nr)
    vis = Mem.from_numpy(vis_init, jt)
    s = C.c_int(0)
    getattr(lib, entry)(
        int(casa), int(gauss), int(nsrc), jones.ptr,
        capi.MEMA(*[m.ptr for m in flux]), capi.MEM3(*[m.ptr for m in dirs]),
        capi.MEM3(*[m.ptr for m in ext]), tel.ptr,
        capi.MEM3(*[m.ptr for m in uvw]), float(g["gast"]), float(g["freq"]),
        int(offset_out), vis.ptr, C.byref(s))
    check_status(s, entry)
    sync_check(entry)
    return vis.numpy()


def gpu_dftw(prec, matrix, normalise, wavenumber, weights, x_in, y_in, z_in,
             x_out, y_out, z_out, data_idx, data, eval_x=1, eval_y=1,
             offset_coord_out=0, num_out=None, offset_out=0, output=None):
    P = prec_code(prec)
    ct = P | OSKAR_COMPLEX
    dt = ct | (OSKAR_MATRIX if matrix else 0)
    if num_out is None:
        num_out = x_out.size - offset_coord_out
    nr = 8 if matrix else 2
    if output is None:
        output = np.zeros((offset_out + num_out) * nr)
    mw = Mem.from_numpy(weights, ct)
    mi = [Mem.from_numpy(a, P) if a is not None else None
          for a in (x_in, y_in, z_in)]
    mo = [Mem.from_numpy(a, P) if a is not None else None
          for a in (x_out, y_out, z_out)]
    md = Mem.from_numpy(data, dt)
    mx = Mem.from_numpy(data_idx, OSKAR_INT) if data_idx is not None else None
    out = Mem.from_numpy(output, dt)
    s = C.c_int(0)
    lib.oskar_dftw(int(normalise), int(x_in.size), float(wavenumber), mw.ptr,
                   *[m.ptr if m else None for m in mi], int(offset_coord_out),
                   int(num_out), *[m.ptr if m else None for m in mo],
                   mx.ptr if mx else None, md.ptr, int(eval_x), int(eval_y),
                   int(offset_out), out.ptr, C.byref(s))
    check_status(s, "oskar_dftw")
    sync_check("oskar_dftw")
    return out.numpy()


def gpu_factorised(prec, mode, casa, beam, I_, Q, U, V, l, m, n, u, v, w,
                   freq, vis_init=None, offset_out=0, uv_filter=None,
                   num_sources=None):
    """oskar_cross_correlate_factorised on the GPU; returns vis reals."""
    P = prec_code(prec)
    ns = u.size
    nsrc = l.size if num_sources is None else num_sources
    mt = P | OSKAR_COMPLEX | OSKAR_MATRIX
    tel = Telescope(P, OSKAR_GPU, ns)
    if uv_filter is not None:
        s = C.c_int(0)
        lib.oskar_telescope_set_uv_filter(tel.ptr, uv_filter[0], uv_filter[1],
                                          b"Metres", C.byref(s))
        check_status(s, "set_uv_filter")
    flux = [Mem.from_numpy(a, P) for a in (I_, Q, U, V)]
    dirs = [Mem.from_numpy(a, P) for a in (l, m, n)]
    uvw = [Mem.from_numpy(a, P) for a in (u, v, w)]
    mb = Mem.from_numpy(beam, mt) if beam is not None else None
    nb = ns * (ns - 1) // 2
    if vis_init is None:
        vis_init = np.zeros((offset_out + nb) * 8)
    vis = Mem.from_numpy(vis_init, mt)
    s = C.c_int(0)
    lib.oskar_cross_correlate_factorised(
        int(casa), int(nsrc), mb.ptr if mb else None,
        capi.MEMA(*[x.ptr for x in flux]), capi.MEM3(*[x.ptr for x in dirs]),
        tel.ptr, capi.MEM3(*[x.ptr for x in uvw]), float(freq), int(mode),
        int(offset_out), vis.ptr, C.byref(s))
    check_status(s, "oskar_cross_correlate_factorised")
    sync_check("oskar_cross_correlate_factorised")
    return vis.numpy()


def gpu_factorised_jones(prec, mode, casa, E, I_, Q, U, V, l, m, n, u, v, w,
                         freq, vis_init=None, offset_out=0, uv_filter=None,
                         num_sources=None, expect_status=0):
    """oskar_cross_correlate_factorised_jones on the GPU with one beam per
    station, E (stations, sources, 8 reals); returns vis reals."""
    P = prec_code(prec)
    ns = u.size
    nsrc = l.size if num_sources is None else num_sources
    mt = P | OSKAR_COMPLEX | OSKAR_MATRIX
    tel = Telescope(P, OSKAR_GPU, ns)
    if uv_filter is not None:
        s = C.c_int(0)
        lib.oskar_telescope_set_uv_filter(tel.ptr, uv_filter[0], uv_filter[1],
                                          b"Metres", C.byref(s))
        check_status(s, "set_uv_filter")
    flux = [Mem.from_numpy(a, P) for a in (I_, Q, U, V)]
    dirs = [Mem.from_numpy(a, P) for a in (l, m, n)]
    uvw = [Mem.from_numpy(a, P) for a in (u, v, w)]
    jones = Jones(mt, OSKAR_GPU, E.shape[0], E.shape[1])
    jones.set(np.ascontiguousarray(E).ravel())
    nb = ns * (ns - 1) // 2
    if vis_init is None:
        vis_init = np.zeros((offset_out + nb) * 8)
    vis = Mem.from_numpy(vis_init, mt)
    s = C.c_int(0)
    lib.oskar_cross_correlate_factorised_jones(
        int(casa), int(nsrc), jones.ptr, capi.MEMA(*[x.ptr for x in flux]),
        capi.MEM3(*[x.ptr for x in dirs]), tel.ptr,
        capi.MEM3(*[x.ptr for x in uvw]), float(freq), int(mode),
        int(offset_out), vis.ptr, C.byref(s))
    if expect_status:
        assert s.value == expect_status, s.value
        return None
    check_status(s, "oskar_cross_correlate_factorised_jones")
    sync_check("oskar_cross_correlate_factorised_jones")
    return vis.numpy()
