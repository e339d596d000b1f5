"""CPU-side checks of the C-ABI library: it loads, exports every function
that include/*.h declares, and keeps the reference's error contract
(oskar/oskar_global.h:20-55; check order location -> type -> dimension of
oskar/correlate/src/oskar_cross_correlate.c:73-110; a non-zero status on
entry makes every call a no-op).  No kernel is launched here.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oskar_amd import capi, lib
from oskar_amd.mem import (OSKAR_CPU, OSKAR_DOUBLE, OSKAR_SINGLE,
                           OSKAR_COMPLEX, OSKAR_MATRIX, Mem)

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in sorted(os.listdir(inc)):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(inc, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = "\n".join(ln for ln in text.splitlines()
                         if not ln.lstrip().startswith("#"))
        for m in re.finditer(r"OSKAR_EXPORT\s+([^;{]*?)\b(\w+)\s*\(", text):
            names.add(m.group(2))
    return sorted(names)


def test_headers_declare_functions():
    names = declared_functions()
    assert "oskar_cross_correlate" in names
    assert "oskar_dftw" in names
    assert "oskar_evaluate_station_beam_aperture_array" in names
    assert len(names) > 100


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports_declared_symbol(name):
    assert hasattr(lib.dll, name), name
    assert name in capi.SIGNATURES, "no ctypes prototype for " + name


def st():
    return C.c_int(0)


def test_error_codes_and_strings():
    assert lib.oskar_get_error_string(-9) == b"data type mismatch"
    assert lib.oskar_get_error_string(-10) == b"data location mismatch"
    assert lib.oskar_get_error_string(-11) == b"data dimension mismatch"


def test_mem_host_roundtrip_and_types():
    a = np.arange(12, dtype=np.float64)
    m = Mem.from_numpy(a, OSKAR_DOUBLE | OSKAR_COMPLEX, OSKAR_CPU)
    assert len(m) == 6
    assert lib.oskar_mem_is_complex(m.ptr) and not lib.oskar_mem_is_matrix(
        m.ptr)
    assert np.array_equal(m.numpy(), a)
    assert lib.oskar_mem_element_size(OSKAR_SINGLE | OSKAR_COMPLEX |
                                      OSKAR_MATRIX) == 32
    assert lib.oskar_mem_element_size(OSKAR_DOUBLE | OSKAR_COMPLEX |
                                      OSKAR_MATRIX) == 64
    s = st()
    bad = lib.oskar_mem_create(0x7F, OSKAR_CPU, 3, C.byref(s))
    assert bad is None and s.value == -13


def make_xcorr_args(location, prec=OSKAR_DOUBLE, ns=5, nsrc=7):
    s = st()
    jones = lib.oskar_jones_create(prec | OSKAR_COMPLEX | OSKAR_MATRIX,
                                   location, ns, nsrc, C.byref(s))
    tel = lib.oskar_telescope_create(prec, location, ns, C.byref(s))
    flux = [lib.oskar_mem_create(prec, location, nsrc, C.byref(s))
            for _ in range(4)]
    dirs = [lib.oskar_mem_create(prec, location, nsrc, C.byref(s))
            for _ in range(3)]
    ext = [lib.oskar_mem_create(prec, location, nsrc, C.byref(s))
           for _ in range(3)]
    uvw = [lib.oskar_mem_create(prec, location, ns, C.byref(s))
           for _ in range(3)]
    vis = lib.oskar_mem_create(prec | OSKAR_COMPLEX | OSKAR_MATRIX, location,
                               ns * (ns - 1) // 2, C.byref(s))
    assert s.value == 0
    return jones, tel, flux, dirs, ext, uvw, vis


def call_xcorr(jones, tel, flux, dirs, ext, uvw, vis, status, nsrc=7):
    lib.oskar_cross_correlate(0, 0, nsrc, jones, capi.MEMA(*flux),
                              capi.MEM3(*dirs), capi.MEM3(*ext), tel,
                              capi.MEM3(*uvw), 1.0, 100e6, 0, vis,
                              C.byref(status))


def test_cross_correlate_error_contract_on_host_memory():
    jones, tel, flux, dirs, ext, uvw, vis = make_xcorr_args(OSKAR_CPU)
    # Pre-set status: no-op, status unchanged.
    s = C.c_int(-3)
    call_xcorr(jones, tel, flux, dirs, ext, uvw, vis, s)
    assert s.value == -3
    # Host memory: every check passes, then the HIP-only library reports
    # that it has no CPU path.
    s = st()
    call_xcorr(jones, tel, flux, dirs, ext, uvw, vis, s)
    assert s.value == -14
    # Type mismatch: uvw in the wrong precision.
    s2 = st()
    bad_uvw = [lib.oskar_mem_create(OSKAR_SINGLE, OSKAR_CPU, 5, C.byref(s2))
               for _ in range(3)]
    s = st()
    call_xcorr(jones, tel, flux, dirs, ext, bad_uvw, vis, s)
    assert s.value == -9
    # Dimension mismatch: more sources requested than the Jones holds.
    s = st()
    call_xcorr(jones, tel, flux, dirs, ext, uvw, vis, s, nsrc=8)
    assert s.value == -11
    # Type mismatch: vis type differs from the Jones type.
    vis_s = lib.oskar_mem_create(OSKAR_DOUBLE | OSKAR_COMPLEX, OSKAR_CPU, 10,
                                 C.byref(s2))
    s = st()
    call_xcorr(jones, tel, flux, dirs, ext, uvw, vis_s, s)
    assert s.value == -9


def test_auto_correlate_error_contract_on_host_memory():
    """oskar_auto_correlate.c:30-46 order: location, type, dimension."""
    jones, tel, flux, dirs, ext, uvw, vis = make_xcorr_args(OSKAR_CPU)
    s2 = st()
    vis_a = lib.oskar_mem_create(OSKAR_DOUBLE | OSKAR_COMPLEX | OSKAR_MATRIX,
                                 OSKAR_CPU, 5, C.byref(s2))

    def call(nsrc, v, status):
        lib.oskar_auto_correlate(nsrc, jones, capi.MEMA(*flux), 0, v,
                                 C.byref(status))

    s = C.c_int(-3)
    call(7, vis_a, s)
    assert s.value == -3
    s = st()
    call(7, vis_a, s)
    assert s.value == -14        # host memory: no CPU path in this library
    vis_s = lib.oskar_mem_create(OSKAR_DOUBLE | OSKAR_COMPLEX, OSKAR_CPU, 5,
                                 C.byref(s2))
    s = st()
    call(7, vis_s, s)
    assert s.value == -9
    s = st()
    call(8, vis_a, s)
    assert s.value == -11
    # Location is checked before type: a CPU Jones with a (CPU) vis of the
    # wrong type reports the type; a GPU vis would report the location, but
    # needs a device.
    assert s2.value == 0


def test_dftw_error_contract_on_host_memory():
    s = st()
    f8, c16 = OSKAR_DOUBLE, OSKAR_DOUBLE | OSKAR_COMPLEX
    w = lib.oskar_mem_create(c16, OSKAR_CPU, 4, C.byref(s))
    xi = [lib.oskar_mem_create(f8, OSKAR_CPU, 4, C.byref(s)) for _ in range(3)]
    xo = [lib.oskar_mem_create(f8, OSKAR_CPU, 9, C.byref(s)) for _ in range(3)]
    data = lib.oskar_mem_create(c16, OSKAR_CPU, 36, C.byref(s))
    out = lib.oskar_mem_create(c16, OSKAR_CPU, 9, C.byref(s))
    real_out = lib.oskar_mem_create(f8, OSKAR_CPU, 9, C.byref(s))
    assert s.value == 0
    s = st()
    lib.oskar_dftw(0, 4, 1.0, w, *xi, 0, 9, *xo, None, data, 1, 1, 0,
                   real_out, C.byref(s))
    assert s.value == -13
    xs = lib.oskar_mem_create(OSKAR_SINGLE, OSKAR_CPU, 4, C.byref(s))
    s = st()
    lib.oskar_dftw(0, 4, 1.0, w, xs, xi[1], xi[2], 0, 9, *xo, None, data, 1,
                   1, 0, out, C.byref(s))
    assert s.value == -9
    s = st()
    lib.oskar_dftw(0, 4, 1.0, w, *xi, 0, 9, *xo, None, data, 1, 1, 0, out,
                   C.byref(s))
    assert s.value == -14


def test_telescope_defaults_follow_reference():
    s = st()
    t = lib.oskar_telescope_create(OSKAR_DOUBLE, OSKAR_CPU, 4, C.byref(s))
    assert lib.oskar_telescope_uv_filter_max(t) == pytest.approx(
        np.finfo(np.float32).max)
    assert lib.oskar_telescope_uv_filter_units(t) == 0
    assert lib.oskar_telescope_num_baselines(t) == 6
    lib.oskar_telescope_set_uv_filter(t, 1.0, 2.0, b"Wavelengths",
                                      C.byref(s))
    assert s.value == 0 and lib.oskar_telescope_uv_filter_units(t) == 1
    lib.oskar_telescope_set_uv_filter(t, 1.0, 2.0, b"furlongs", C.byref(s))
    assert s.value == -15


def test_station_defaults_follow_reference():
    s = st()
    stn = lib.oskar_station_create(OSKAR_DOUBLE, OSKAR_CPU, 3, C.byref(s))
    assert s.value == 0
    assert lib.oskar_station_type(stn) == 0
    assert lib.oskar_station_normalise_final_beam(stn) == 1
    assert lib.oskar_station_enable_array_pattern(stn) == 1
    assert lib.oskar_station_beam_coord_type(stn) == 2     # RADEC
    e = lib.oskar_station_element(stn, 0)
    assert lib.oskar_element_type(e) == 0                    # dipole
    assert lib.oskar_element_dipole_length(e) == 0.5
    assert lib.oskar_element_dipole_length_units(e) == 1     # wavelengths
    lib.oskar_station_set_element_coords(
        stn, 0, 1, (C.c_double * 3)(1, 2, 0.5), (C.c_double * 3)(1, 2, 0.5),
        C.byref(s))
    assert s.value == 0 and lib.oskar_station_array_is_3d(stn) == 1
    lib.oskar_station_set_element_coords(
        stn, 0, 3, (C.c_double * 3)(0, 0, 0), (C.c_double * 3)(0, 0, 0),
        C.byref(s))
    assert s.value == -5
