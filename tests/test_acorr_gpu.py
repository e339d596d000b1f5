"""GPU parity of oskar_auto_correlate against the reference's golden vectors
and the CPU oracle, through the C-ABI.

Tolerances:
  double: max |V - V_ref| / max |V_ref| < 1e-12 (FP64 sums in another
          order than the reference's sequential ones);
  single: against the FP64 result of the same (float32-exact) inputs,
          max rel < 2e-6 (FP32 sums over a few sources per thread and over
          a wave, FP64 beyond: about 10 FP32 roundings), and against the
          reference's own single-precision output, max rel < 1e-5.
Entries outside [offset_out, offset_out + stations) and the imaginary parts
that the reference leaves alone must come back bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from gpu_util import prec_code, sync_check
from helpers import load_golden, max_rel
from oskar_amd import capi, lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_GPU, OSKAR_MATRIX, Jones, Mem,
                           check_status)

pytestmark = pytest.mark.gpu


def gpu_auto_correlate(prec, matrix, ns, jones, I_, Q, U, V, vis_init,
                       offset_out=0, num_sources=None, jones_sources=None):
    P = prec_code(prec)
    jt = P | OSKAR_COMPLEX | (OSKAR_MATRIX if matrix else 0)
    nsrc = I_.size if num_sources is None else num_sources
    js = I_.size if jones_sources is None else jones_sources
    J = Jones(jt, OSKAR_GPU, ns, js)
    J.set(jones)
    flux = [Mem.from_numpy(x, P) for x in (I_, Q, U, V)]
    vis = Mem.from_numpy(vis_init, jt)
    s = C.c_int(0)
    lib.oskar_auto_correlate(int(nsrc), J.ptr,
                             capi.MEMA(*[m.ptr for m in flux]),
                             int(offset_out), vis.ptr, C.byref(s))
    check_status(s, "oskar_auto_correlate")
    sync_check("oskar_auto_correlate")
    return vis.numpy()


def untouched(got, vis0, nr, off, ns, matrix):
    v, v0 = got.reshape(-1, nr), vis0.reshape(-1, nr).astype(got.dtype)
    keep = [1, 7] if matrix else [1]
    return (np.array_equal(v[:off], v0[:off]) and
            np.array_equal(v[off + ns:], v0[off + ns:]) and
            np.array_equal(v[off:off + ns, keep], v0[off:off + ns, keep]))


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("matrix", [0, 1])
def test_acorr_gpu_matches_golden(hip, matrix, prec):
    print("acorr golden", matrix, prec, flush=True)
    g = load_golden("acorr.npz")
    ns, off = int(g["num_stations"]), int(g["offset_out"])
    nr = 8 if matrix else 2
    vis0 = g["vis0_%d" % matrix]
    got = gpu_auto_correlate(prec, matrix, ns, g["jones_%d" % matrix],
                             g["I"], g["Q"], g["U"], g["V"], vis0,
                             offset_out=off)
    ref_d = g["vis_%d_d" % matrix]
    if prec == "d":
        assert max_rel(got, ref_d) < 1e-12
    else:
        assert max_rel(got, ref_d) < 2e-6
        assert max_rel(got, g["vis_%d_f" % matrix]) < 1e-5
    assert untouched(got, vis0, nr, off, ns, matrix)


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("matrix", [0, 1])
@pytest.mark.parametrize("shape", [(1, 1), (3, 255), (9, 257), (13, 4099),
                                   (257, 12345)])
def test_acorr_gpu_ragged_shapes(hip, matrix, prec, shape):
    """Station counts that do not fill a group of 8 (4 for FP64 matrices),
    source counts below one workgroup's stride and not a multiple of it;
    (13, 4099) and (257, 12345) split the sources over several chunks, so
    the FP64 partial sums and the chunk reduction are exercised."""
    ns, nsrc = shape
    print("acorr ragged", shape, matrix, prec, flush=True)
    rng = np.random.default_rng(ns * 7919 + nsrc)
    nr = 8 if matrix else 2
    f32 = np.float32
    jones = rng.normal(size=ns * nsrc * nr).astype(f32)
    I_ = rng.uniform(0.1, 3.0, nsrc).astype(f32)
    Q, U, V = (rng.uniform(-0.1, 0.1, nsrc).astype(f32) for _ in range(3))
    off = 2
    vis0 = rng.uniform(-1, 1, (off + ns + 1) * nr).astype(f32)
    got = gpu_auto_correlate(prec, matrix, ns, jones, I_, Q, U, V, vis0,
                             offset_out=off)
    want = oracle.auto_correlate("double", matrix, ns, jones, I_, Q, U, V,
                                 vis=vis0.astype(np.float64), offset_out=off)
    assert max_rel(got, want) < (1e-12 if prec == "d" else 2e-6)
    assert untouched(got, vis0, nr, off, ns, matrix)


@pytest.mark.parametrize("matrix", [0, 1])
def test_acorr_gpu_fewer_sources_than_jones_and_zero_sources(hip, matrix):
    """num_sources below the Jones source count reads the Jones values with
    a source stride of num_sources, as the reference's kernels do; zero
    sources leave vis unchanged."""
    rng = np.random.default_rng(5)
    ns, js, nsrc = 6, 40, 33
    nr = 8 if matrix else 2
    jones = rng.normal(size=ns * js * nr)
    I_ = rng.uniform(0.5, 2.0, js)
    Q, U, V = (rng.uniform(-0.2, 0.2, js) for _ in range(3))
    vis0 = rng.uniform(-1, 1, ns * nr)
    got = gpu_auto_correlate("d", matrix, ns, jones, I_, Q, U, V, vis0,
                             num_sources=nsrc)
    want = oracle.auto_correlate("double", matrix, ns,
                                 jones[:ns * nsrc * nr], I_[:nsrc], Q[:nsrc],
                                 U[:nsrc], V[:nsrc], vis=vis0.copy())
    assert max_rel(got, want) < 1e-12
    got = gpu_auto_correlate("d", matrix, ns, jones, I_, Q, U, V, vis0,
                             num_sources=0)
    assert np.array_equal(got, vis0)


def test_acorr_gpu_accumulates_linearly(hip):
    """Two calls over the two halves of a sky add up to one call over the
    whole sky (accumulation into vis, never a clear)."""
    rng = np.random.default_rng(8)
    ns, nsrc = 20, 3000
    h = nsrc // 2
    jones = rng.normal(size=(ns, nsrc, 8))
    I_ = rng.uniform(0.5, 2.0, nsrc)
    Q, U, V = (rng.uniform(-0.2, 0.2, nsrc) for _ in range(3))
    vis0 = np.zeros(ns * 8)
    whole = gpu_auto_correlate("d", 1, ns, jones.ravel(), I_, Q, U, V, vis0)
    a = gpu_auto_correlate("d", 1, ns, jones[:, :h].ravel(), I_[:h], Q[:h],
                           U[:h], V[:h], vis0)
    b = gpu_auto_correlate("d", 1, ns, jones[:, h:].ravel(), I_[h:], Q[h:],
                           U[h:], V[h:], a)
    assert max_rel(b, whole) < 1e-13


def test_acorr_gpu_errors(hip):
    """The reference's check order (location, type, dimension), plus a vis
    too short for offset_out + stations."""
    rng = np.random.default_rng(9)
    ns, nsrc = 4, 10
    P = prec_code("d")
    jt = P | OSKAR_COMPLEX | OSKAR_MATRIX
    J = Jones(jt, OSKAR_GPU, ns, nsrc)
    J.set(rng.normal(size=ns * nsrc * 8))
    flux = [Mem.from_numpy(rng.uniform(size=nsrc), P) for _ in range(4)]
    fl = capi.MEMA(*[m.ptr for m in flux])

    def call(nsources, vis, off=0):
        s = C.c_int(0)
        lib.oskar_auto_correlate(nsources, J.ptr, fl, off, vis.ptr,
                                 C.byref(s))
        return s.value

    assert call(nsrc, Mem(P | OSKAR_COMPLEX, OSKAR_GPU, ns)) == -9
    assert call(nsrc + 1, Mem(jt, OSKAR_GPU, ns)) == -11
    assert call(nsrc, Mem(jt, OSKAR_GPU, ns - 1)) == -11
    assert call(nsrc, Mem(jt, OSKAR_GPU, ns + 1), off=2) == -11
    assert call(nsrc, Mem(jt, OSKAR_GPU, ns + 1), off=1) == 0
    sync_check("acorr errors")
