"""GPU parity of oskar_evaluate_jones_K and oskar_jones_join (the step
before the correlator) through the C-ABI.

Tolerances:
  K, double: max |K - K_ref| < 1e-12 against the reference's own CPU body
      (golden vectors) and the oracle (phases up to ~1e3 rad; the sines
      differ from glibc's by a few ulp of the phase);
  K, single: |K - K_64| <= 4 eps_32 k (|u l| + |v m| + |w (n - 1)|) + 1e-6
      per element, where K_64 is the FP64 value for the same float32
      inputs: the FP32 phase carries about one rounding of its largest
      term (more or less with fused multiply-adds) and sincosf adds ~1 ulp;
  join: bit for bit against numpy in the same precision and term order
      (the kernels round every product and sum on its own: no fused
      multiply-add), and within 8 eps |a| |b| of complex128 products; the
      whole-matrix and per-element kernels agree bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from gpu_util import prec_code, sync_check
from helpers import load_golden
from oskar_amd import lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_GPU, OSKAR_MATRIX, Jones, Mem,
                           check_status)

pytestmark = pytest.mark.gpu

C0 = 299792458.0


def gpu_K(prec, ns, l, m, n, u, v, w, freq, filt, fmin, fmax, ignore_w=0):
    P = prec_code(prec)
    K = Jones(P | OSKAR_COMPLEX, OSKAR_GPU, ns, l.size)
    mems = [Mem.from_numpy(x, P) for x in (l, m, n, u, v, w, filt)]
    s = C.c_int(0)
    lib.oskar_evaluate_jones_K(K.ptr, l.size, *[x.ptr for x in mems[:6]],
                               float(freq), mems[6].ptr, float(fmin),
                               float(fmax), int(ignore_w), C.byref(s))
    check_status(s, "oskar_evaluate_jones_K")
    sync_check("oskar_evaluate_jones_K")
    return K.mem.numpy()


def k_bound(l, m, n, u, v, w, k):
    """Per-element FP32 bound (see the module docstring)."""
    t = (np.abs(np.outer(u, l)) + np.abs(np.outer(v, m)) +
         np.abs(np.outer(w, n - 1)))
    return (4 * 2.0 ** -24 * k * t + 1e-6).ravel()


@pytest.mark.parametrize("prec", ["f", "d"])
def test_jones_K_matches_golden(hip, prec):
    print("jones K golden", prec, flush=True)
    g = load_golden("beam_parts.npz")
    k = float(g["k_k"])
    freq = k * C0 / (2 * np.pi)
    dt = np.float64 if prec == "d" else np.float32
    x = {c: g["k_" + c].astype(dt) for c in "lmnuvwI"}
    got = gpu_K(prec, x["u"].size, x["l"], x["m"], x["n"], x["u"], x["v"],
                x["w"], freq, x["I"], float(g["k_fmin"]), float(g["k_fmax"]))
    if prec == "d":
        assert np.max(np.abs(got - g["k_jones"])) < 1e-12
    else:
        want = oracle.jones_K(*[x[c].astype(np.float64) for c in "lmnuvw"],
                              float(np.float32(k)), x["I"].astype(np.float64),
                              float(g["k_fmin"]), float(g["k_fmax"]))
        err = np.abs(got - want).reshape(-1, 2).max(axis=1)
        assert np.all(err <= k_bound(*[x[c] for c in "lmnuvw"], k))
    # Filtered sources are exactly zero.
    z = ~((g["k_I"] > g["k_fmin"]) & (g["k_I"] <= g["k_fmax"]))
    assert z.any()
    kk = got.reshape(x["u"].size, -1, 2)
    assert np.all(kk[:, z, :] == 0)


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("ns,nsrc,ignore_w", [(1, 1, 0), (15, 255, 0),
                                              (16, 256, 1), (37, 1001, 0),
                                              (70, 4099, 1)])
def test_jones_K_ragged_shapes(hip, prec, ns, nsrc, ignore_w):
    """Station counts off the kernel's group of 16 and source counts off
    its 256-thread blocks."""
    print("jones K", prec, ns, nsrc, ignore_w, flush=True)
    rng = np.random.default_rng(ns * 7 + nsrc)
    dt = np.float64 if prec == "d" else np.float32
    u, v = (rng.uniform(-20000, 20000, ns).astype(dt) for _ in range(2))
    w = rng.uniform(-100, 100, ns).astype(dt)
    r = np.radians(20.0) * np.sqrt(rng.uniform(size=nsrc))
    t = rng.uniform(0, 2 * np.pi, nsrc)
    l, m = (r * np.cos(t)).astype(dt), (r * np.sin(t)).astype(dt)
    n = np.sqrt(1 - l.astype(np.float64) ** 2 -
                m.astype(np.float64) ** 2).astype(dt)
    filt = rng.uniform(0, 2, nsrc).astype(dt)
    freq = 120e6
    k = 2 * np.pi * freq / C0
    got = gpu_K(prec, ns, l, m, n, u, v, w, freq, filt, 0.1, 1.9, ignore_w)
    f8 = [x.astype(np.float64) for x in (l, m, n, u, v, w)]
    kk = float(np.float32(k)) if prec == "f" else k
    if ignore_w:
        f8[5] = np.zeros(ns)
    want = oracle.jones_K(*f8, kk, filt.astype(np.float64), 0.1, 1.9)
    err = np.abs(got - want).reshape(-1, 2).max(axis=1)
    if prec == "d":
        assert err.max() < 1e-9       # phases up to ~1e4 rad
    else:
        ww = f8[5].astype(dt)
        assert np.all(err <= k_bound(l, m, n, u, v, ww, k))


def rand_jones(rng, dt, matrix, ns, nsrc):
    return rng.normal(size=ns * nsrc * (8 if matrix else 2)).astype(dt)


def np_join(a, b, am, bm):
    """J1 J2 per element in complex arithmetic (numpy complex128)."""
    ca = a.astype(np.float64).view(np.complex128)
    cb = b.astype(np.float64).view(np.complex128)
    if am and bm:
        A, B = ca.reshape(-1, 2, 2), cb.reshape(-1, 2, 2)
        return (A @ B).reshape(-1).view(np.float64)
    if am:
        return (ca.reshape(-1, 4) * cb[:, None]).reshape(-1).view(np.float64)
    if bm:
        return (ca[:, None] * cb.reshape(-1, 4)).reshape(-1).view(np.float64)
    return (ca * cb).view(np.float64)


def np_join_exact(a, b, am, bm):
    """The kernels' arithmetic in a's precision: each complex product is
    (ax bx - ay by, ax by + ay bx), the 2x2 product sums two of them."""
    def cm(x, y):
        return (x[..., 0] * y[..., 0] - x[..., 1] * y[..., 1],
                x[..., 0] * y[..., 1] + x[..., 1] * y[..., 0])
    A = a.reshape(-1, 4, 2) if am else a.reshape(-1, 1, 2)
    B = b.reshape(-1, 4, 2) if bm else b.reshape(-1, 1, 2)
    if am and bm:
        out = np.empty_like(A)
        for r, (i, j, k, l) in enumerate(((0, 0, 1, 2), (0, 1, 1, 3),
                                          (2, 0, 3, 2), (2, 1, 3, 3))):
            t0, t1 = cm(A[:, i], B[:, j]), cm(A[:, k], B[:, l])
            out[:, r, 0] = t0[0] + t1[0]
            out[:, r, 1] = t0[1] + t1[1]
        return out.reshape(-1)
    re, im = cm(A, B)
    return np.stack([re, im], axis=-1).reshape(-1)


def join_bound(a, b, am, bm, eps):
    """4 eps |a| |b| per output real (two products and a sum per term)."""
    ca = np.abs(a.astype(np.float64).view(np.complex128))
    cb = np.abs(b.astype(np.float64).view(np.complex128))
    if am and bm:
        A, B = ca.reshape(-1, 2, 2), cb.reshape(-1, 2, 2)
        mag = (A @ B).reshape(-1)
    elif am:
        mag = (ca.reshape(-1, 4) * cb[:, None]).reshape(-1)
    elif bm:
        mag = (ca[:, None] * cb.reshape(-1, 4)).reshape(-1)
    else:
        mag = ca * cb
    return np.repeat(8 * eps * mag, 2)


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("am,bm", [(1, 1), (1, 0), (0, 1), (0, 0)])
@pytest.mark.parametrize("in_place", [0, 1])
def test_jones_join(hip, prec, am, bm, in_place):
    print("join", prec, am, bm, in_place, flush=True)
    if in_place and bm and not am:
        pytest.skip("j3 may alias j1 only when j1 has j3's type")
    P = prec_code(prec)
    dt = np.float64 if prec == "d" else np.float32
    eps = 2.0 ** -52 if prec == "d" else 2.0 ** -23
    ns, nsrc = 13, 1001
    rng = np.random.default_rng(am * 2 + bm + 10 * in_place)
    a = rand_jones(rng, dt, am, ns, nsrc)
    b = rand_jones(rng, dt, bm, ns, nsrc)
    ta = P | OSKAR_COMPLEX | (OSKAR_MATRIX if am else 0)
    tb = P | OSKAR_COMPLEX | (OSKAR_MATRIX if bm else 0)
    to = P | OSKAR_COMPLEX | (OSKAR_MATRIX if (am or bm) else 0)
    J1, J2 = Jones(ta, OSKAR_GPU, ns, nsrc), Jones(tb, OSKAR_GPU, ns, nsrc)
    J1.set(a)
    J2.set(b)
    J3 = J1 if in_place else Jones(to, OSKAR_GPU, ns, nsrc)
    s = C.c_int(0)
    lib.oskar_jones_join(J3.ptr, J1.ptr, J2.ptr, C.byref(s))
    check_status(s, "oskar_jones_join")
    sync_check("oskar_jones_join")
    got = J3.mem.numpy()
    assert np.array_equal(got, np_join_exact(a, b, am, bm))
    want = np_join(a, b, am, bm)
    assert np.all(np.abs(got - want) <= join_bound(a, b, am, bm, eps))


@pytest.mark.parametrize("am,bm", [(1, 1), (1, 0), (0, 1)])
def test_join_vector_and_element_paths_agree(hip, am, bm):
    """The launcher takes whole-matrix 16-byte accesses when every matrix
    pointer is aligned to a matrix, and element accesses otherwise; both
    give the same bits.  Offsetting the pointers by one complex element
    (inside buffers allocated one matrix larger) selects the element path."""
    print("join paths", am, bm, flush=True)
    fn = lib.oskar_amd_jones_join_launch
    fn.restype = C.c_int
    fn.argtypes = [C.c_int, C.c_int, C.c_int, C.c_size_t, C.c_void_p,
                   C.c_void_p, C.c_void_p, C.c_void_p]
    cur = lib.oskar_amd_current_stream
    cur.restype = C.c_void_p
    cur.argtypes = []
    P = prec_code("f")
    rng = np.random.default_rng(3 + am + bm)
    n = 5003
    a = rng.normal(size=n * (8 if am else 2)).astype(np.float32)
    b = rng.normal(size=n * (8 if bm else 2)).astype(np.float32)
    got = []
    for off in (0, 2):                    # 0 or 8 bytes (2 floats)
        pad = np.zeros(off + 8, np.float32)
        ma = Mem.from_numpy(np.concatenate([pad[:off], a, pad]), P)
        mb = Mem.from_numpy(np.concatenate([pad[:off], b, pad]), P)
        mo = Mem.from_numpy(np.zeros(off + 8 * n + 8, np.float32), P)
        pa = lib.oskar_mem_void(ma.ptr) + 4 * off
        pb = lib.oskar_mem_void(mb.ptr) + 4 * off
        po = lib.oskar_mem_void(mo.ptr) + 4 * off
        assert fn(4, am, bm, n, po, pa, pb, cur()) == 0
        sync_check("join paths")
        got.append(mo.numpy()[off:off + 8 * n])
    assert np.array_equal(got[0], got[1])
    assert np.array_equal(got[0], np_join_exact(a, b, am, bm))
    want = np_join(a, b, am, bm)
    assert np.all(np.abs(got[0] - want) <=
                  join_bound(a, b, am, bm, 2.0 ** -23))
