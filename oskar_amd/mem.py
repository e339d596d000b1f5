"""Python handles over the C-ABI structures (oskar_Mem, oskar_Jones,
oskar_Telescope, oskar_Station, oskar_StationWork, oskar_Sky,
oskar_FactorisedBatch).

Mirrors the way the reference's own tests drive the C API
(oskar/correlate/test/Test_cross_correlate.cpp:60-120): create typed
buffers in a location, fill them, call the hot-path function, read back.
Data cross the boundary as numpy arrays copied through host aliases.
"""
import ctypes as C

import numpy as np

from . import capi
from .capi import lib

OSKAR_CPU, OSKAR_GPU = 0, 1
OSKAR_CHAR, OSKAR_INT, OSKAR_SINGLE, OSKAR_DOUBLE = 0x01, 0x02, 0x04, 0x08
OSKAR_COMPLEX, OSKAR_MATRIX = 0x20, 0x40


class OskarError(RuntimeError):
    def __init__(self, status, where=""):
        msg = lib.oskar_get_error_string(status)
        msg = msg.decode() if msg else "unknown"
        super().__init__("%s: status %d (%s)" % (where, status, msg))
        self.status = status


def check_status(status, where=""):
    v = status.value if isinstance(status, C.c_int) else int(status)
    if v != 0:
        raise OskarError(v, where)


def _np_dtype(type_code):
    prec = type_code & 0x0F
    if prec == OSKAR_INT:
        return np.int32
    if prec == OSKAR_CHAR:
        return np.int8
    return np.float64 if prec == OSKAR_DOUBLE else np.float32


def _reals_per_element(type_code):
    n = 1
    if type_code & OSKAR_COMPLEX:
        n *= 2
    if type_code & OSKAR_MATRIX:
        n *= 4
    return n


class Mem:
    """Owning handle to an oskar_Mem."""

    def __init__(self, type_code, location, num_elements, ptr=None):
        self.type = type_code
        self.location = location
        if ptr is None:
            st = C.c_int(0)
            ptr = lib.oskar_mem_create(type_code, location, num_elements,
                                       C.byref(st))
            check_status(st, "oskar_mem_create")
        self.ptr = ptr

    @classmethod
    def from_numpy(cls, arr, type_code, location=OSKAR_GPU):
        """Copy a numpy array (reals, interleaved for complex) into a new
        oskar_Mem of the given type and location."""
        dt = _np_dtype(type_code)
        a = np.ascontiguousarray(arr, dtype=dt).ravel()
        per = _reals_per_element(type_code)
        if a.size % per:
            raise ValueError("array size not a multiple of element size")
        n = a.size // per
        m = cls(type_code, location, n)
        if n:
            st = C.c_int(0)
            alias = lib.oskar_mem_create_alias_from_raw(
                a.ctypes.data_as(C.c_void_p), type_code, OSKAR_CPU, n,
                C.byref(st))
            lib.oskar_mem_copy_contents(m.ptr, alias, 0, 0, n, C.byref(st))
            lib.oskar_mem_free(alias, C.byref(C.c_int(0)))
            check_status(st, "copy to oskar_Mem")
        return m

    def __len__(self):
        return int(lib.oskar_mem_length(self.ptr))

    def numpy(self):
        """Copy contents to a new flat numpy array of reals."""
        n = len(self)
        out = np.zeros(n * _reals_per_element(self.type),
                       dtype=_np_dtype(self.type))
        if n:
            st = C.c_int(0)
            alias = lib.oskar_mem_create_alias_from_raw(
                out.ctypes.data_as(C.c_void_p), self.type, OSKAR_CPU, n,
                C.byref(st))
            lib.oskar_mem_copy_contents(alias, self.ptr, 0, 0, n, C.byref(st))
            lib.oskar_mem_free(alias, C.byref(C.c_int(0)))
            check_status(st, "copy from oskar_Mem")
        return out

    def clear(self):
        st = C.c_int(0)
        lib.oskar_mem_clear_contents(self.ptr, C.byref(st))
        check_status(st, "oskar_mem_clear_contents")

    def data_ptr(self):
        return lib.oskar_mem_void(self.ptr)

    def free(self):
        if self.ptr:
            lib.oskar_mem_free(self.ptr, C.byref(C.c_int(0)))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _Borrowed(Mem):
    """Non-owning view of an oskar_Mem held by another structure."""

    def __init__(self, ptr):
        self.ptr = ptr
        self.type = lib.oskar_mem_type(ptr)
        self.location = lib.oskar_mem_location(ptr)

    def free(self):
        self.ptr = None


class Jones:
    def __init__(self, type_code, location, num_stations, num_sources):
        st = C.c_int(0)
        self.ptr = lib.oskar_jones_create(type_code, location, num_stations,
                                          num_sources, C.byref(st))
        check_status(st, "oskar_jones_create")

    @property
    def mem(self):
        return _Borrowed(lib.oskar_jones_mem(self.ptr))

    def set(self, arr):
        """Fill from a numpy array of reals ([station][source] order)."""
        src = Mem.from_numpy(arr, lib.oskar_jones_type(self.ptr), OSKAR_CPU)
        st = C.c_int(0)
        lib.oskar_mem_copy_contents(lib.oskar_jones_mem(self.ptr), src.ptr,
                                    0, 0, len(src), C.byref(st))
        check_status(st, "Jones.set")

    def __del__(self):
        try:
            if self.ptr:
                lib.oskar_jones_free(self.ptr, C.byref(C.c_int(0)))
                self.ptr = None
        except Exception:
            pass


class Telescope:
    def __init__(self, precision, location, num_stations):
        st = C.c_int(0)
        self.ptr = lib.oskar_telescope_create(precision, location,
                                              num_stations, C.byref(st))
        check_status(st, "oskar_telescope_create")

    def set_station_ecef(self, x, y, z):
        for dim, arr in enumerate((x, y, z)):
            dst = _Borrowed(
                lib.oskar_telescope_station_true_offset_ecef_metres(
                    self.ptr, dim))
            src = Mem.from_numpy(arr, dst.type, OSKAR_CPU)
            st = C.c_int(0)
            lib.oskar_mem_copy_contents(dst.ptr, src.ptr, 0, 0, len(src),
                                        C.byref(st))
            check_status(st, "set_station_ecef")

    def __del__(self):
        try:
            if self.ptr:
                lib.oskar_telescope_free(self.ptr, C.byref(C.c_int(0)))
                self.ptr = None
        except Exception:
            pass


class Station:
    def __init__(self, precision, location, num_elements):
        st = C.c_int(0)
        self.ptr = lib.oskar_station_create(precision, location, num_elements,
                                            C.byref(st))
        check_status(st, "oskar_station_create")
        self.precision = precision

    def set_coords(self, true_xyz, measured_xyz=None):
        """Bulk-set element ENU coordinates (3 x N arrays), then analyse."""
        if measured_xyz is None:
            measured_xyz = true_xyz
        for dim in range(3):
            for getter, arr in (
                    (lib.oskar_station_element_true_enu_metres, true_xyz),
                    (lib.oskar_station_element_measured_enu_metres,
                     measured_xyz)):
                dst = _Borrowed(getter(self.ptr, 0, dim))
                src = Mem.from_numpy(arr[dim], self.precision, OSKAR_CPU)
                st = C.c_int(0)
                lib.oskar_mem_copy_contents(dst.ptr, src.ptr, 0, 0, len(src),
                                            C.byref(st))
                check_status(st, "Station.set_coords")
        st = C.c_int(0)
        lib.oskar_station_analyse(self.ptr, C.byref(C.c_int(0)), C.byref(st))
        check_status(st, "oskar_station_analyse")

    def __del__(self):
        try:
            if self.ptr:
                lib.oskar_station_free(self.ptr, C.byref(C.c_int(0)))
                self.ptr = None
        except Exception:
            pass


class StationView(Station):
    """Non-owning handle to a station held by another (a child station):
    freed with its parent."""

    def __init__(self, ptr, precision):
        self.ptr = ptr
        self.precision = precision

    def __del__(self):
        pass


class StationWork:
    def __init__(self, precision, location):
        st = C.c_int(0)
        self.ptr = lib.oskar_station_work_create(precision, location,
                                                 C.byref(st))
        check_status(st, "oskar_station_work_create")

    def __del__(self):
        try:
            if self.ptr:
                lib.oskar_station_work_free(self.ptr, C.byref(C.c_int(0)))
                self.ptr = None
        except Exception:
            pass


class Sky:
    """Handle to an oskar_Sky column table (oskar/sky/private_sky.h)."""

    # oskar_SkyAttribInt / oskar_SkyAttribDouble
    PRECISION, MEM_LOCATION, CAPACITY, NUM_SOURCES, NUM_COLUMNS, \
        USE_EXTENDED = range(6)
    REF_RA_RAD, REF_DEC_RAD = 0, 1

    def __init__(self, precision, location, num_sources):
        st = C.c_int(0)
        self.ptr = lib.oskar_sky_create(precision, location, num_sources,
                                        C.byref(st))
        check_status(st, "oskar_sky_create")

    def attr(self, a):
        return lib.oskar_sky_int(self.ptr, a)

    def set_reference(self, ra_rad, dec_rad):
        lib.oskar_sky_set_double(self.ptr, self.REF_RA_RAD, float(ra_rad))
        lib.oskar_sky_set_double(self.ptr, self.REF_DEC_RAD, float(dec_rad))

    def column(self, column_type, attribute=0):
        """The column (created if missing) as a non-owning Mem view."""
        st = C.c_int(0)
        p = lib.oskar_sky_column(self.ptr, column_type, attribute,
                                 C.byref(st))
        check_status(st, "oskar_sky_column")
        return _Borrowed(p)

    def set_column(self, column_type, values, attribute=0):
        """Writes values into rows [0, len(values)) of a column."""
        col = self.column(column_type, attribute)
        src = Mem.from_numpy(values, col.type, OSKAR_CPU)
        st = C.c_int(0)
        lib.oskar_mem_copy_contents(col.ptr, src.ptr, 0, 0, len(src),
                                    C.byref(st))
        check_status(st, "Sky.set_column")

    def get_column(self, column_type, attribute=0):
        """Rows [0, NUM_SOURCES) of an existing column (None if absent)."""
        p = lib.oskar_sky_column_const(self.ptr, column_type, attribute)
        if not p:
            return None
        return _Borrowed(p).numpy()[:self.attr(self.NUM_SOURCES)]

    def __del__(self):
        try:
            if self.ptr:
                lib.oskar_sky_free(self.ptr, C.byref(C.c_int(0)))
                self.ptr = None
        except Exception:
            pass


class FactorisedBatch:
    """oskar_FactorisedBatch: many (channel, time) units in one persistent
    launch.  Keeps a reference to every unit's Mem objects, which the C side
    only borrows."""

    def __init__(self, precision, mode, num_stations, max_units):
        st = C.c_int(0)
        self.ptr = lib.oskar_factorised_batch_create(
            precision, mode, num_stations, max_units, C.byref(st))
        check_status(st, "oskar_factorised_batch_create")
        self._keep = {}

    def set_unit(self, unit, num_sources, beam, flux, dirs, uvw, freq_hz,
                 offset_out, vis):
        """beam: Mem or None; flux: 4 Mem (I, Q, U, V); dirs: 3 Mem (l, m,
        n); uvw: 3 Mem (metres); vis: Mem (complex matrix)."""
        st = C.c_int(0)
        lib.oskar_factorised_batch_set_unit(
            self.ptr, int(unit), int(num_sources),
            beam.ptr if beam is not None else None,
            capi.MEMA(*[x.ptr for x in flux]),
            capi.MEM3(*[x.ptr for x in dirs]),
            capi.MEM3(*[x.ptr for x in uvw]), float(freq_hz),
            int(offset_out), vis.ptr, C.byref(st))
        check_status(st, "oskar_factorised_batch_set_unit")
        self._keep[unit] = (beam, tuple(flux), tuple(dirs), tuple(uvw), vis)

    def run(self, num_units, use_casa, tel):
        st = C.c_int(0)
        lib.oskar_factorised_batch_run(self.ptr, int(num_units),
                                       int(use_casa), tel.ptr, C.byref(st))
        check_status(st, "oskar_factorised_batch_run")

    def autotune(self, num_units):
        st = C.c_int(0)
        lib.oskar_factorised_batch_autotune(self.ptr, int(num_units),
                                            C.byref(st))
        check_status(st, "oskar_factorised_batch_autotune")
        return self.config()

    def set_config(self, tile_shape, src_per_part, num_workgroups=0,
                   use_graph=1):
        st = C.c_int(0)
        lib.oskar_factorised_batch_set_config(
            self.ptr, int(tile_shape), int(src_per_part),
            int(num_workgroups), int(use_graph), C.byref(st))
        check_status(st, "oskar_factorised_batch_set_config")

    def config(self):
        out = (C.c_int * 4)()
        lib.oskar_factorised_batch_config(self.ptr, out)
        return dict(tile_shape=out[0], src_per_part=out[1],
                    workgroups=out[2], use_graph=out[3])

    def last_run_ms(self):
        return lib.oskar_factorised_batch_last_run_ms(self.ptr)

    def free(self):
        if self.ptr:
            lib.oskar_factorised_batch_free(self.ptr, C.byref(C.c_int(0)))
            self.ptr = None
            self._keep = {}

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BeamSurrogate:
    """oskar_BeamSurrogate: the learned station-beam network."""

    def __init__(self, freq_min, freq_max, l0, m0, station_radius, seed=1):
        st = C.c_int(0)
        self.ptr = lib.oskar_beam_surrogate_create(
            float(freq_min), float(freq_max), float(l0), float(m0),
            float(station_radius), int(seed), C.byref(st))
        check_status(st, "oskar_beam_surrogate_create")
        self.num_params = lib.oskar_beam_surrogate_num_params()

    def fourier(self):
        n = lib.oskar_beam_surrogate_num_fourier()
        out = np.zeros(2 * n, np.float32)
        lib.oskar_beam_surrogate_fourier(
            self.ptr, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out.reshape(n, 2)

    def get_params(self):
        out = np.zeros(self.num_params, np.float32)
        st = C.c_int(0)
        lib.oskar_beam_surrogate_get_params(
            self.ptr, out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(st))
        check_status(st, "oskar_beam_surrogate_get_params")
        return out

    def set_params(self, theta):
        a = np.ascontiguousarray(theta, np.float32)
        st = C.c_int(0)
        lib.oskar_beam_surrogate_set_params(
            self.ptr, a.ctypes.data_as(C.POINTER(C.c_float)), C.byref(st))
        check_status(st, "oskar_beam_surrogate_set_params")

    @property
    def peak(self):
        return lib.oskar_beam_surrogate_peak(self.ptr)

    def train(self, n, l, m, f, beam, batch, steps, lr0, lr1, seed=1):
        hist = np.zeros(max(steps, 1))
        st = C.c_int(0)
        lib.oskar_beam_surrogate_train(
            self.ptr, int(n), l.ptr, m.ptr, f.ptr, beam.ptr, int(batch),
            int(steps), float(lr0), float(lr1), int(seed),
            hist.ctypes.data_as(C.POINTER(C.c_double)), C.byref(st))
        check_status(st, "oskar_beam_surrogate_train")
        return hist[:steps]

    def loss_and_gradient(self, features, targets):
        """features [B, 128], targets [B, 8] (numpy); returns (loss, grad)."""
        B = features.shape[0]
        fm = Mem.from_numpy(features.ravel(), OSKAR_SINGLE)
        tm = Mem.from_numpy(targets.ravel(), OSKAR_SINGLE)
        gm = Mem(OSKAR_SINGLE, OSKAR_GPU, self.num_params)
        loss = C.c_double(0)
        st = C.c_int(0)
        lib.oskar_beam_surrogate_loss_and_gradient(
            self.ptr, int(B), fm.ptr, tm.ptr, C.byref(loss), gm.ptr,
            C.byref(st))
        check_status(st, "oskar_beam_surrogate_loss_and_gradient")
        return loss.value, gm.numpy()

    def evaluate(self, n, l, m, freq, beam):
        st = C.c_int(0)
        lib.oskar_beam_surrogate_evaluate(self.ptr, int(n), l.ptr, m.ptr,
                                          float(freq), beam.ptr, C.byref(st))
        check_status(st, "oskar_beam_surrogate_evaluate")

    def evaluate_with_fallback(self, station, work, n, x, y, z, freq,
                               threshold, beam):
        st = C.c_int(0)
        num = C.c_int(0)
        lib.oskar_beam_surrogate_evaluate_with_fallback(
            self.ptr, station.ptr, work.ptr, int(n), x.ptr, y.ptr, z.ptr,
            float(freq), float(threshold), beam.ptr, C.byref(num),
            C.byref(st))
        check_status(st, "oskar_beam_surrogate_evaluate_with_fallback")
        return num.value

    def __del__(self):
        try:
            if self.ptr:
                lib.oskar_beam_surrogate_free(self.ptr, C.byref(C.c_int(0)))
                self.ptr = None
        except Exception:
            pass


def beam_training_set(station, work, num_freqs, dirs_per_freq, freq_min,
                      freq_max, l0, m0, seed=1):
    """Exact-beam samples (GPU Mems l, m, f, beam) for surrogate training."""
    l, m, f = (Mem(OSKAR_SINGLE, OSKAR_GPU, 1) for _ in range(3))
    beam = Mem(OSKAR_SINGLE | OSKAR_COMPLEX | OSKAR_MATRIX, OSKAR_GPU, 1)
    st = C.c_int(0)
    lib.oskar_beam_surrogate_training_set(
        station.ptr, work.ptr, int(num_freqs), int(dirs_per_freq),
        float(freq_min), float(freq_max), float(l0), float(m0), int(seed),
        l.ptr, m.ptr, f.ptr, beam.ptr, C.byref(st))
    check_status(st, "oskar_beam_surrogate_training_set")
    return l, m, f, beam
