"""FP64 visibilities on every baseline, for checking the factorised MFMA
path at full size.

The reference is the plain pipeline of the reference's sim_baselines
(oskar/interferometer/src/oskar_interferometer_run_block.c:251-333) run in
double precision through this library's C-ABI:
    K = oskar_evaluate_jones_K, J = K E (oskar_jones_join),
    V += oskar_cross_correlate(J)
with the direct FP64 correlator kernel (pinned to the CPU oracle at 1e-10
and re-checked against the long-double oracle by the tests that use this).
Sources are taken in chunks so that J stays a few GB.
"""
import ctypes as C

import numpy as np

from oskar_amd import capi, lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_DOUBLE, OSKAR_GPU,
                           OSKAR_MATRIX, Jones, Mem, Telescope, check_status)

CM64 = OSKAR_DOUBLE | OSKAR_COMPLEX | OSKAR_MATRIX


def f32(a):
    """Values as an FP32 path sees them, promoted to FP64."""
    return np.asarray(a, np.float32).astype(np.float64)


def direct_fp64_vis(u, v, w, l, m, n, I, freq, Q=None, U=None, V=None,
                    beam=None, station_beams=None, chunk=1 << 18, casa=0,
                    uv_filter=None):
    """vis (num_baselines x 8 reals) = sum over sources of J_p B J_q^H with
    J_p(s) = K_p(s) E_p(s) in FP64.

    beam: (num_sources, 8) reals of one station beam shared by all
    stations, or None for E = identity.  station_beams: a function
    (s0, s1, mem) that writes E_p(s) for sources [s0, s1) of every station
    into the device FP64 matrix array mem ([station][source] order, s1 - s0
    sources per station); it replaces beam."""
    N, S = u.size, l.size
    NB = N * (N - 1) // 2
    z = np.zeros(S)
    Q = z if Q is None else Q
    U = z if U is None else U
    V = z if V is None else V
    chunk = min(chunk, max(S, 1))
    tel = Telescope(OSKAR_DOUBLE, OSKAR_GPU, N)
    if uv_filter is not None:
        s = C.c_int(0)
        lib.oskar_telescope_set_uv_filter(tel.ptr, uv_filter[0],
                                          uv_filter[1], b"Metres", C.byref(s))
        check_status(s, "set_uv_filter")
    uvw = [Mem.from_numpy(a, OSKAR_DOUBLE) for a in (u, v, w)]
    vis = Mem(CM64, OSKAR_GPU, NB)
    vis.clear()
    E = Jones(CM64, OSKAR_GPU, N, chunk)
    K = Jones(OSKAR_DOUBLE | OSKAR_COMPLEX, OSKAR_GPU, N, chunk)
    Emem = lib.oskar_jones_mem(E.ptr)
    ext = [Mem(OSKAR_DOUBLE, OSKAR_GPU, 1) for _ in range(3)]
    s = C.c_int(0)
    for s0 in range(0, S, chunk):
        s1 = min(S, s0 + chunk)
        ns = s1 - s0
        flux = [Mem.from_numpy(a[s0:s1], OSKAR_DOUBLE) for a in (I, Q, U, V)]
        dirs = [Mem.from_numpy(a[s0:s1], OSKAR_DOUBLE) for a in (l, m, n)]
        if ns != chunk:
            E = Jones(CM64, OSKAR_GPU, N, ns)
            K = Jones(OSKAR_DOUBLE | OSKAR_COMPLEX, OSKAR_GPU, N, ns)
            Emem = lib.oskar_jones_mem(E.ptr)
        if station_beams is not None:
            station_beams(s0, s1, Emem)
        else:
            if beam is None:
                b = np.zeros((ns, 8))
                b[:, 0] = b[:, 6] = 1.0
            else:
                b = beam[s0:s1]
            bm = Mem.from_numpy(b, CM64)
            for a in range(N):
                lib.oskar_mem_copy_contents(Emem, bm.ptr, a * ns, 0, ns,
                                            C.byref(s))
            check_status(s, "beam copies")
        lib.oskar_evaluate_jones_K(K.ptr, ns, dirs[0].ptr, dirs[1].ptr,
                                   dirs[2].ptr, uvw[0].ptr, uvw[1].ptr,
                                   uvw[2].ptr, float(freq), flux[0].ptr,
                                   -1e300, 1e300, 0, C.byref(s))
        check_status(s, "oskar_evaluate_jones_K")
        lib.oskar_jones_join(E.ptr, K.ptr, E.ptr, C.byref(s))
        check_status(s, "oskar_jones_join")
        lib.oskar_cross_correlate(
            int(casa), 0, ns, E.ptr, capi.MEMA(*[x.ptr for x in flux]),
            capi.MEM3(*[x.ptr for x in dirs]),
            capi.MEM3(*[x.ptr for x in ext]), tel.ptr,
            capi.MEM3(*[x.ptr for x in uvw]), 0.0, float(freq), 0, vis.ptr,
            C.byref(s))
        check_status(s, "oskar_cross_correlate")
        lib.oskar_device_synchronize(OSKAR_GPU)
    out = vis.numpy()
    e = C.c_int(0)
    lib.oskar_device_check_error_hip(C.byref(e))
    check_status(e, "direct_fp64_vis")
    return out


def rel_rms(got, ref):
    """sqrt(sum |V - V64|^2) / sqrt(sum |V64|^2) over every real."""
    g = np.asarray(got, np.float64)
    return float(np.sqrt(np.sum((g - ref) ** 2) / np.sum(ref * ref)))


def sample_baselines(rng, N, count):
    """count distinct (p > q) pairs and their baseline indices."""
    seen, p, q = set(), [], []
    while len(p) < count:
        a, b = sorted(rng.integers(0, N, 2))[::-1]
        if a != b and (a, b) not in seen:
            seen.add((a, b))
            p.append(a)
            q.append(b)
    p, q = np.array(p, np.int32), np.array(q, np.int32)
    return p, q, q * (N - 1) - (q - 1) * q // 2 + p - q - 1
