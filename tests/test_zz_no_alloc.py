"""oskar_cross_correlate_no_alloc: the allocation-free variant of
oskar_cross_correlate, whose split-source partial sums live in a fixed
device array instead of a workspace the library allocates.

Checked against oskar_cross_correlate on the same inputs, on two calls in a
row, the second with a smaller input (the array keeps the first call's
partials, which must not leak into the second's result):
  1. 512 stations x 3,000 sources, FP64 matrix Jones, Gaussian sources,
     bandwidth and time smearing (8 source splits, which fill the 64 MiB
     array): every value identical to oskar_cross_correlate's;
  2. 64 stations x 1,000 sources, FP32 scalar Jones, point sources: every
     value identical.
Both also accumulate into non-zero visibilities at an offset, and the
check order and error codes of oskar_cross_correlate hold (an FP32 vis for
FP64 Jones is a type mismatch, with vis untouched).

At 600 stations the usual 6 splits of FP64 matrix partials (69 MB) do not
fit the array, so the entry point splits the sources 5 times: the result
equals oskar_cross_correlate's to the FP64 summation order (max difference
below 1e-12 of the largest value).
"""
import ctypes as C
import math

import numpy as np
import pytest

from gpu_util import gpu_cross_correlate
from oskar_amd import capi, lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_DOUBLE, OSKAR_GPU,
                           OSKAR_MATRIX, OSKAR_SINGLE, Jones, Mem,
                           Telescope)

pytestmark = pytest.mark.gpu


def inputs(rng, ns, nsrc):
    """Random stations, sources and Jones in the golden-fixture layout."""
    r = 5000.0 * np.sqrt(rng.uniform(0, 1, ns))
    t = rng.uniform(0, 2 * math.pi, ns)
    rho = math.radians(5.0) * np.sqrt(rng.uniform(0, 1, nsrc))
    phi = rng.uniform(0, 2 * math.pi, nsrc)
    l, m = rho * np.cos(phi), rho * np.sin(phi)
    I_ = rng.uniform(0.01, 10.0, nsrc)
    pol = 0.2 * I_ * rng.uniform(0, 1, nsrc)
    d = rng.standard_normal((3, nsrc))
    d /= np.linalg.norm(d, axis=0)
    j = rng.standard_normal(ns * nsrc * 8) / math.sqrt(2)
    return dict(
        u=r * np.cos(t), v=r * np.sin(t), w=rng.uniform(-5, 5, ns),
        x=r * np.cos(t + 0.3), y=r * np.sin(t + 0.3),
        l=l, m=m, n=np.sqrt(1 - l * l - m * m),
        I=I_, Q=pol * d[0], U=pol * d[1], V=pol * d[2],
        a=rng.uniform(1e-9, 1e-7, nsrc), b=rng.uniform(-1e-8, 1e-8, nsrc),
        c=rng.uniform(1e-9, 1e-7, nsrc),
        jones_m=j, jones_s=j[: ns * nsrc * 2],
        channel_bw=10e3, time_avg=1.0, ra0=0.4, dec0=-0.7, gast=1.3,
        freq=150e6)


def both(prec, gauss, matrix, g, bs, ts, offset_out):
    nb = g["u"].size * (g["u"].size - 1) // 2
    nr = 8 if matrix else 2
    init = np.random.default_rng(7).standard_normal((offset_out + nb) * nr)
    want = gpu_cross_correlate(prec, 0, gauss, matrix, g, bs, ts,
                               vis_init=init, offset_out=offset_out)
    got = gpu_cross_correlate(prec, 0, gauss, matrix, g, bs, ts,
                              vis_init=init, offset_out=offset_out,
                              entry="oskar_cross_correlate_no_alloc")
    return got, want, init


def test_no_alloc_equals_oskar_cross_correlate_two_calls(hip):
    rng = np.random.default_rng(2025)
    print("no_alloc call 1: 512 st x 3000 src FP64 matrix Gaussian BS TS",
          flush=True)
    got, want, init = both("d", 1, 1, inputs(rng, 512, 3000), 1, 1, 5)
    assert np.array_equal(got, want)
    assert not np.array_equal(got, init)
    print("no_alloc call 2: 64 st x 1000 src FP32 scalar point", flush=True)
    got, want, init = both("f", 0, 0, inputs(rng, 64, 1000), 0, 0, 3)
    assert np.array_equal(got, want)
    assert not np.array_equal(got, init)


def test_no_alloc_checks_like_oskar_cross_correlate(hip):
    ns, nsrc = 8, 16
    jt = OSKAR_DOUBLE | OSKAR_COMPLEX | OSKAR_MATRIX
    jones = Jones(jt, OSKAR_GPU, ns, nsrc)
    tel = Telescope(OSKAR_DOUBLE, OSKAR_GPU, ns)
    src = [Mem.from_numpy(np.ones(nsrc), OSKAR_DOUBLE) for _ in range(10)]
    uvw = [Mem.from_numpy(np.zeros(ns), OSKAR_DOUBLE) for _ in range(3)]
    init = np.arange(ns * (ns - 1) // 2 * 8, dtype=np.float64)
    vis = Mem.from_numpy(init, OSKAR_SINGLE | OSKAR_COMPLEX | OSKAR_MATRIX)
    s = C.c_int(0)
    lib.oskar_cross_correlate_no_alloc(
        0, 0, nsrc, jones.ptr, capi.MEMA(*[m.ptr for m in src[:4]]),
        capi.MEM3(*[m.ptr for m in src[4:7]]),
        capi.MEM3(*[m.ptr for m in src[7:]]), tel.ptr,
        capi.MEM3(*[m.ptr for m in uvw]), 0.0, 100e6, 0, vis.ptr,
        C.byref(s))
    assert s.value == -9            # OSKAR_ERR_TYPE_MISMATCH
    assert np.array_equal(vis.numpy(), init.astype(np.float32))


def test_no_alloc_fewer_splits_when_scratch_is_short(hip):
    rng = np.random.default_rng(600)
    print("no_alloc capped: 600 st x 2000 src FP64 matrix Gaussian BS TS",
          flush=True)
    got, want, init = both("d", 1, 1, inputs(rng, 600, 2000), 1, 1, 0)
    scale = float(np.max(np.abs(want - init)))
    assert scale > 0.0
    assert float(np.max(np.abs(got - want))) <= 1e-12 * scale
