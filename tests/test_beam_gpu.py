"""GPU parity of oskar_dftw and the aperture-array station beam against the
CPU oracle and the reference's golden vectors, through the C-ABI.

Tolerances: double max rel < 1e-10; single relative RMS < 1e-5 and max rel
< 2e-4 (the hardware sine/cosine of an FP32 phase; the reference's own GPU
test asserts double precision only, Test_evaluate_array_pattern.cpp:29-42).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from gpu_util import gpu_dftw, sync_check
from helpers import load_golden, max_rel, rel_rms
from oskar_amd import lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_DOUBLE, OSKAR_GPU,
                           OSKAR_MATRIX, OSKAR_SINGLE, Mem, Station,
                           StationWork, check_status)

pytestmark = pytest.mark.gpu
C0 = 299792458.0


def check(prec, got, ref):
    if prec == "d":
        assert max_rel(got, ref) < 1e-10
    else:
        assert rel_rms(got, ref) < 1e-5
        assert max_rel(got, ref) < 2e-4


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("kind", ["c2c", "m2m"])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("typed", ["full", "typed"])
def test_dftw_gpu_matches_golden(hip, prec, kind, dim, typed):
    print("dftw", prec, kind, dim, typed, flush=True)
    g = load_golden("dftw.npz")
    matrix = kind == "m2m"
    data = g["data_%s_%s" % ("m" if matrix else "c", typed)]
    got = gpu_dftw(prec, matrix, 1, float(g["wavenumber"]), g["weights"],
                   g["x_in"], g["y_in"], g["z_in"] if dim == 3 else None,
                   g["x_out"], g["y_out"], g["z_out"] if dim == 3 else None,
                   g["idx"] if typed == "typed" else None, data)
    check(prec, got, g["out_%s_%s_%dd_%s" % (prec, kind, dim, typed)])


@pytest.mark.parametrize("prec", ["f", "d"])
def test_dftw_gpu_offsets_eval_flags_and_chunks(hip, prec):
    """600 elements (three LDS chunks, element types in runs), an output
    offset, a coordinate offset, and eval_x / eval_y selecting pairs."""
    rng = np.random.default_rng(3)
    nin, nout, ntypes = 600, 700, 3
    k = 2 * np.pi * 200e6 / C0
    xi, yi, zi = (rng.uniform(-19, 19, nin) for _ in range(3))
    ph = rng.uniform(0, 2 * np.pi, nin)
    w = np.stack([np.cos(ph), np.sin(ph)], 1).ravel()
    th = rng.uniform(0, 1.5, nout + 5)
    phi = rng.uniform(0, 2 * np.pi, nout + 5)
    xo, yo, zo = np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi), \
        np.cos(th)
    idx = np.sort(rng.integers(0, ntypes, nin)).astype(np.int32)
    data = rng.normal(size=ntypes * nout * 8)
    for ex, ey in ((1, 1), (1, 0), (0, 1)):
        init = rng.normal(size=(nout + 9) * 8)
        got = gpu_dftw(prec, 1, 0, k, w, xi, yi, zi, xo, yo, zo, idx, data,
                       eval_x=ex, eval_y=ey, offset_coord_out=5,
                       num_out=nout, offset_out=9, output=init)
        want = oracle.dftw("double" if prec == "d" else "single", 1, 0, k, w,
                           xi, yi, zi, xo, yo, zo, idx, data, eval_x=ex,
                           eval_y=ey, offset_coord_out=5, num_out=nout,
                           offset_out=9, output=init.astype(
                               np.float64 if prec == "d" else np.float32))
        check(prec, got, want)


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("layout", ["one_type", "type_per_chunk"])
def test_dftw_gpu_uniform_chunks(hip, prec, layout):
    """1031 elements: four full LDS chunks and a 7-element tail (the
    unrolled loop's remainder), with one data row for everything, or a
    different data row in each chunk (a flush at every chunk boundary)."""
    rng = np.random.default_rng(4)
    nin, nout = 1031, 500
    k = 2 * np.pi * 300e6 / C0
    xi, yi = (rng.uniform(-40, 40, nin) for _ in range(2))
    ph = rng.uniform(0, 2 * np.pi, nin)
    w = np.stack([np.cos(ph), np.sin(ph)], 1).ravel()
    th = rng.uniform(0, 1.5, nout)
    phi = rng.uniform(0, 2 * np.pi, nout)
    xo, yo = np.sin(th) * np.cos(phi), np.sin(th) * np.sin(phi)
    if layout == "one_type":
        idx = np.zeros(nin, np.int32)
    else:
        idx = (np.arange(nin) // 256).astype(np.int32)
    data = rng.normal(size=(int(idx.max()) + 1) * nout * 8)
    got = gpu_dftw(prec, 1, 1, k, w, xi, yi, None, xo, yo, None, idx, data)
    want = oracle.dftw("double" if prec == "d" else "single", 1, 1, k, w,
                       xi, yi, None, xo, yo, None, idx, data)
    check(prec, got, want)


def make_station(prec, ne, rng, radius=19.0, kinds=(0,), types=None,
                 is_3d=False, cable=None):
    P = OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE
    st = Station(P, OSKAR_GPU, ne)
    r = radius * np.sqrt(rng.uniform(size=ne))
    t = rng.uniform(0, 2 * np.pi, ne)
    xt = np.stack([r * np.cos(t), r * np.sin(t),
                   rng.uniform(-0.3, 0.3, ne) if is_3d else np.zeros(ne)])
    xm = xt + rng.normal(scale=0.01, size=xt.shape)
    st.set_coords(xt, xm)
    s = C.c_int(0)
    lib.oskar_station_resize_element_types(st.ptr, len(kinds), C.byref(s))
    for i, kind in enumerate(kinds):
        e = lib.oskar_station_element(st.ptr, i)
        lib.oskar_element_set_element_type(
            e, b"Isotropic" if kind == 1 else b"Dipole", C.byref(s))
    if types is not None:
        for i, ty in enumerate(types):
            lib.oskar_station_set_element_type(st.ptr, i, int(ty), C.byref(s))
    if cable is not None:
        for i, c in enumerate(cable):
            lib.oskar_station_set_element_cable_length_error(
                st.ptr, 0, i, float(c), C.byref(s))
    check_status(s, "make_station")
    return st, xt, xm


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("variant", ["dipole", "isotropic_norm", "two_types",
                                     "3d_cable_swap"])
def test_station_beam_aa_gpu_matches_oracle(hip, prec, variant):
    print("station beam", prec, variant, flush=True)
    seeds = {"dipole": 1, "isotropic_norm": 2, "two_types": 3,
             "3d_cable_swap": 4}
    rng = np.random.default_rng(seeds[variant])
    ne, npts, freq = 64, 2000, 110e6
    kinds, types, is_3d, cable = (0,), None, False, None
    norm_array = norm_element = swap = 0
    if variant == "isotropic_norm":
        kinds, norm_array = (1,), 1
    elif variant == "two_types":
        kinds = (0, 1)
        types = np.sort(rng.integers(0, 2, ne))
    elif variant == "3d_cable_swap":
        is_3d, cable, swap, norm_element = True, rng.uniform(
            -0.02, 0.02, ne), 1, 1
    st, xt, xm = make_station(prec, ne, rng, kinds=kinds, types=types,
                              is_3d=is_3d, cable=cable)
    lib.oskar_station_set_normalise_array_pattern(st.ptr, norm_array)
    lib.oskar_station_set_normalise_element_pattern(st.ptr, norm_element)
    lib.oskar_station_set_swap_xy(st.ptr, swap)
    # Beam pointed at az 30 deg, el 60 deg.
    lib.oskar_station_set_phase_centre(st.ptr, 4, np.radians(30.0),
                                       np.radians(60.0))
    th = np.arccos(rng.uniform(-0.2, 1.0, npts))
    phi = rng.uniform(0, 2 * np.pi, npts)
    x, y, z = np.sin(th) * np.sin(phi), np.sin(th) * np.cos(phi), np.cos(th)
    P = OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE
    mx, my, mz = (Mem.from_numpy(a, P) for a in (x, y, z))
    beam = Mem(P | OSKAR_COMPLEX | OSKAR_MATRIX, OSKAR_GPU, npts)
    work = StationWork(P, OSKAR_GPU)
    s = C.c_int(0)
    lib.oskar_evaluate_station_beam_aperture_array(
        st.ptr, work.ptr, npts, mx.ptr, my.ptr, mz.ptr, 0, 0.0, freq,
        beam.ptr, C.byref(s))
    check_status(s, "station beam")
    sync_check("station beam")
    got = beam.numpy()
    bd = np.array([np.cos(np.radians(60)) * np.sin(np.radians(30)),
                   np.cos(np.radians(60)) * np.cos(np.radians(30)),
                   np.sin(np.radians(60))])
    if prec == "f":
        # The oracle runs in double on the single-precision inputs.
        xt, xm = xt.astype(np.float32), xm.astype(np.float32)
        x, y, z = (a.astype(np.float32) for a in (x, y, z))
        cable = None if cable is None else cable.astype(np.float32)
    want = oracle.station_beam_aa(
        1, xt[0], xt[1], xt[2], xm[0], xm[1], xm[2],
        np.zeros(ne) if cable is None else cable,
        np.zeros(ne, np.int32) if types is None else types,
        np.array(kinds, np.int32), np.full(len(kinds), 0.5 * C0 / freq), bd,
        x, y, z, freq, swap_xy=swap, is_3d=int(is_3d), norm_array=norm_array,
        norm_element=norm_element)
    if prec == "d":
        assert max_rel(got, want) < 1e-10
    else:
        assert rel_rms(got, want) < 2e-5
    # Blanking below the horizon.
    below = z < 0
    assert below.any()
    assert np.all(got.reshape(-1, 8)[below] == 0)
