"""The backend seam: the eight raw-pointer launchers that replace the
reference's oskar_cross_correlate_{point,gaussian}_cuda_{f,d} and
oskar_cross_correlate_scalar_{point,gaussian}_cuda_{f,d}
(oskar/correlate/oskar_cross_correlate_cuda.h:60-244,
oskar_cross_correlate_scalar_cuda.h), called with device pointers and the
derived scalars oskar_cross_correlate computes
(oskar/correlate/src/oskar_cross_correlate.c:42-71), against the golden
vectors of the reference's own CPU kernels for every variant (FP64 max rel
< 1e-10; FP32 rel RMS < 1e-5, max rel < 1e-4); a launch error would reach
oskar_device_check_error_hip.

Concurrency: correlations on two HIP streams at once (each with its own
split-sum workspace) equal the same calls made one after the other, bit
for bit, for the direct and the factorised correlators.
"""
import ctypes as C

import numpy as np
import pytest

from gpu_util import gpu_cross_correlate, sync_check
from helpers import XCORR_CASES, load_golden, max_rel, rel_rms, xcorr_key
from oskar_amd import capi, lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_DOUBLE, OSKAR_GPU,
                           OSKAR_MATRIX, OSKAR_SINGLE, Jones, Mem,
                           Telescope, check_status)

pytestmark = pytest.mark.gpu
C0 = 299792458.0
FLT_MAX = 3.4028234663852886e38


def seam_call(g, prec, matrix, gauss, casa, bs, ts):
    P = OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE
    ns = g["u"].size
    nsrc = g["l"].size
    jt = P | OSKAR_COMPLEX | (OSKAR_MATRIX if matrix else 0)
    jones = Jones(jt, OSKAR_GPU, ns, nsrc)
    jones.set(g["jones_m"] if matrix else g["jones_s"])
    keep = [Mem.from_numpy(g[k], P) for k in ("I", "Q", "U", "V", "l", "m",
                                              "n", "a", "b", "c", "u", "v",
                                              "w", "x", "y")]
    d = {k: lib.oskar_mem_void(m.ptr) for k, m in zip(
        ("I", "Q", "U", "V", "l", "m", "n", "a", "b", "c", "u", "v", "w",
         "x", "y"), keep)}
    nb = ns * (ns - 1) // 2
    nr = 8 if matrix else 2
    vis = Mem.from_numpy(np.zeros(nb * nr), jt)
    freq = float(g["freq"])
    scal = [0.0, FLT_MAX, freq / C0,
            float(g["channel_bw"]) / freq if bs else 0.0,
            float(g["time_avg"]) if ts else 0.0,
            float(g["gast"]) - float(g["ra0"]), float(g["dec0"])]
    name = "oskar_cross_correlate_%s%s_hip_%s" % (
        "" if matrix else "scalar_", "gaussian" if gauss else "point", prec)
    fn = getattr(lib, name)
    jp = lib.oskar_mem_void(lib.oskar_jones_mem(jones.ptr))
    src = [d["I"]] + ([d["Q"], d["U"], d["V"]] if matrix else []) + \
        [d["l"], d["m"], d["n"]] + ([d["a"], d["b"], d["c"]] if gauss else [])
    fn(int(casa), nsrc, ns, 0, jp, *src, d["u"], d["v"], d["w"], d["x"],
       d["y"], *scal, lib.oskar_mem_void(vis.ptr))
    s = C.c_int(0)
    lib.oskar_device_synchronize(OSKAR_GPU)
    lib.oskar_device_check_error_hip(C.byref(s))
    check_status(s, name)
    return vis.numpy()


@pytest.mark.parametrize("case", XCORR_CASES,
                         ids=[xcorr_key(*c) for c in XCORR_CASES])
def test_seam_launchers_match_golden(hip, case):
    g = load_golden("xcorr_long.npz")
    got = seam_call(g, *case)
    want = g[xcorr_key(*case)]
    if case[0] == "d":
        assert max_rel(got, want) < 1e-10
    else:
        assert rel_rms(got, want) < 1e-5 and max_rel(got, want) < 1e-4


def _hip():
    return C.CDLL("libamdhip64.so")


def _stream_create():
    s = C.c_void_p()
    assert _hip().hipStreamCreate(C.byref(s)) == 0
    return s


def _stream_destroy(s):
    _hip().hipStreamDestroy(s)


def _direct_inputs(seed, ns, nsrc):
    rng = np.random.default_rng(seed)
    P = OSKAR_SINGLE
    jones = Jones(P | OSKAR_COMPLEX | OSKAR_MATRIX, OSKAR_GPU, ns, nsrc)
    jones.set(rng.normal(size=ns * nsrc * 8))
    flux = [Mem.from_numpy(rng.uniform(0.1, 2, nsrc), P)] + [
        Mem.from_numpy(rng.normal(scale=0.05, size=nsrc), P)
        for _ in range(3)]
    l, m = rng.uniform(-0.3, 0.3, nsrc), rng.uniform(-0.3, 0.3, nsrc)
    dirs = [Mem.from_numpy(a, P) for a in (l, m, np.sqrt(1 - l * l - m * m))]
    ext = [Mem(P, OSKAR_GPU, 1) for _ in range(3)]
    uvw = [Mem.from_numpy(rng.uniform(-3e3, 3e3, ns), P) for _ in range(3)]
    tel = Telescope(P, OSKAR_GPU, ns)
    vis = Mem(P | OSKAR_COMPLEX | OSKAR_MATRIX, OSKAR_GPU,
              ns * (ns - 1) // 2)
    return dict(jones=jones, flux=flux, dirs=dirs, ext=ext, uvw=uvw,
                tel=tel, vis=vis, nsrc=nsrc)


def _direct(x, status):
    lib.oskar_cross_correlate(
        0, 0, x["nsrc"], x["jones"].ptr,
        capi.MEMA(*[m.ptr for m in x["flux"]]),
        capi.MEM3(*[m.ptr for m in x["dirs"]]),
        capi.MEM3(*[m.ptr for m in x["ext"]]), x["tel"].ptr,
        capi.MEM3(*[m.ptr for m in x["uvw"]]), 0.0, 150e6, 0,
        x["vis"].ptr, C.byref(status))


def _fact(x, status):
    lib.oskar_cross_correlate_factorised(
        0, x["nsrc"], None, capi.MEMA(*[m.ptr for m in x["flux"]]),
        capi.MEM3(*[m.ptr for m in x["dirs"]]), x["tel"].ptr,
        capi.MEM3(*[m.ptr for m in x["uvw"]]), 150e6, 0, 0, x["vis"].ptr,
        C.byref(status))


@pytest.mark.parametrize("which", ["direct", "factorised"])
def test_two_streams_equal_serial(hip, which):
    fn = _direct if which == "direct" else _fact
    a = _direct_inputs(1, 256, 30000)
    b = _direct_inputs(2, 200, 40000)
    s = C.c_int(0)
    serial = []
    for x in (a, b):
        x["vis"].clear()
        fn(x, s)
        check_status(s, which)
        sync_check(which)
        serial.append(x["vis"].numpy())
    sa, sb = _stream_create(), _stream_create()
    try:
        for rep in range(3):
            for x in (a, b):
                x["vis"].clear()
            sync_check("clear")
            lib.oskar_hip_set_stream(sa)
            fn(a, s)
            lib.oskar_hip_set_stream(sb)
            fn(b, s)
            lib.oskar_hip_set_stream(None)
            check_status(s, which)
            sync_check(which + " on two streams")
            assert np.array_equal(a["vis"].numpy(), serial[0])
            assert np.array_equal(b["vis"].numpy(), serial[1])
    finally:
        lib.oskar_hip_set_stream(None)
        _stream_destroy(sa)
        _stream_destroy(sb)
