"""CPU oracle for the visibility hot path (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboskar_oracle.so`` (the C restatement in oskar_oracle.c)
and, when it has been built, ``oracle/_ref/liboskar_ref.so`` (the reference's
own CPU kernels compiled from its sources by ``make -C oracle ref``).
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
this package, as the checker; the product library never calls it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(_HERE, "liboskar_oracle.so")
REF_PATH = os.path.join(_HERE, "_ref", "liboskar_ref.so")

_lib = None
_ref = None

P = C.c_void_p
I = C.c_int
D = C.c_double


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise ImportError("oracle not built: make -C oracle")
        _lib = C.CDLL(ORACLE_PATH)
        xc = [I, I, I, I, I] + [P] * 16 + [D, D, I, D, D, D, D, I, D, D, I, P]
        _lib.oracle_cross_correlate_d.argtypes = xc
        _lib.oracle_cross_correlate_f.argtypes = xc
        dft = [I, I, I, D] + [P] * 4 + [I, I] + [P] * 5 + [I, I, I, P]
        _lib.oracle_dftw_d.argtypes = dft
        _lib.oracle_dftw_f.argtypes = dft
        _lib.oracle_element_weights_dft_d.argtypes = [I] + [P] * 4 + [D] * 4 + [P]
        _lib.oracle_element_weights_dft_f.argtypes = [I] + [P] * 4 + [D] * 4 + [P]
        _lib.oracle_theta_phi_d.argtypes = [I, I, P, P, P, I, D, D, P, P, P]
        _lib.oracle_dipole_pattern_d.argtypes = [I, I, P, P, P, D, D, I, I, P]
        _lib.oracle_station_beam_aa_d.argtypes = (
            [I, I] + [P] * 7 + [P, I, P, P, D, D, I, I, I, I, I, D, D, D, I,
                                P, P, P, D, P, P, P])
        _lib.oracle_station_beam_aa_d.restype = I
        _lib.oracle_evaluate_jones_K_d.argtypes = (
            [I, P, P, P, I, P, P, P, D, P, D, D, P])
        _lib.oracle_factorised_subset_ld.argtypes = (
            [I, P, P, P, P, P, I, P, P, P, D, I, P, P, P])
        _lib.oracle_update_horizon_mask_d.argtypes = [I, P, P, P, D, D, D, P]
        _lib.oracle_update_horizon_mask_f.argtypes = [I, P, P, P, D, D, D, P]
        _lib.oracle_prefix_sum_int.argtypes = [I, P, P]
        sc = [I, I, I, P, I, P, P, I, P, P]
        _lib.oracle_sky_copy_source_data_d.argtypes = sc
        _lib.oracle_sky_copy_source_data_f.argtypes = sc
        _lib.oracle_sky_copy_source_data_d.restype = I
        _lib.oracle_sky_copy_source_data_f.restype = I
        ac = [I, I, I, I] + [P] * 6
        _lib.oracle_auto_correlate_d.argtypes = ac
        _lib.oracle_auto_correlate_f.argtypes = ac
        rd = [I, P, P, D, D, P, P, P]
        _lib.oracle_relative_directions_d.argtypes = rd
        _lib.oracle_relative_directions_f.argtypes = rd
        gp = [I] + [P] * 5 + [D, D] + [P] * 4
        _lib.oracle_gaussian_parameters_d.argtypes = gp
        _lib.oracle_gaussian_parameters_f.argtypes = gp
        _lib.oracle_gaussian_parameters_d.restype = I
        _lib.oracle_gaussian_parameters_f.restype = I
        _lib.oracle_angular_distance.argtypes = [D, D, D, D]
        _lib.oracle_angular_distance.restype = D
        DP = C.POINTER(C.c_double)
        _lib.oracle_snap_beam_direction.argtypes = [DP, DP, DP, I, P, P]
        _lib.oracle_snap_beam_direction.restype = I
        U = C.c_uint
        _lib.oracle_philox4x32_10.argtypes = [P, P, P]
        _lib.oracle_random_gaussian_d.argtypes = [U, U, U, U, U, P]
        _lib.oracle_element_errors_d.argtypes = [I, U, I, I, P, P, P, P, P]
    return _lib


def ref():
    """The reference's compiled CPU kernels, or None if not built here."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        _ref = C.CDLL(REF_PATH)
    return _ref


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=dt)


def cross_correlate(precision, use_casa, source_type, matrix, jones, I_, Q,
                    U, V, l, m, n, a, b, c, st_u, st_v, st_w, st_x, st_y,
                    channel_bw=0.0, time_avg=0.0, azel=False, ra0=0.0,
                    dec0=0.0, uv_min=0.0, uv_max=-1.0, uv_wavelengths=False,
                    gast=0.0, freq=100e6, vis=None, offset_out=0):
    """oskar_cross_correlate semantics on host arrays; returns vis (reals,
    accumulated into the given vis if any)."""
    dt = np.float64 if precision == "double" else np.float32
    ns = st_u.size
    nb = ns * (ns - 1) // 2
    nr = 8 if matrix else 2
    if vis is None:
        vis = np.zeros((offset_out + nb) * nr, dtype=dt)
    vis = np.ascontiguousarray(vis, dtype=dt)
    num_sources = l.size
    arrs = [_c(x, dt) for x in (jones, I_, Q, U, V, l, m, n, a, b, c, st_u,
                                st_v, st_w, st_x, st_y)]
    fn = (lib().oracle_cross_correlate_d if precision == "double"
          else lib().oracle_cross_correlate_f)
    fn(int(use_casa), int(source_type), int(matrix), int(num_sources), ns,
       *[_p(x) for x in arrs], float(channel_bw), float(time_avg),
       int(bool(azel)), float(ra0), float(dec0), float(uv_min), float(uv_max),
       int(bool(uv_wavelengths)), float(gast), float(freq), int(offset_out),
       _p(vis))
    return vis


def auto_correlate(precision, matrix, num_stations, jones, I_, Q=None,
                   U=None, V=None, vis=None, offset_out=0):
    """oskar_auto_correlate semantics on host arrays (jones: stations x
    sources values, 8 reals per matrix, 2 per scalar); returns vis
    (accumulated into the given vis if any)."""
    dt = np.float64 if precision == "double" else np.float32
    nr = 8 if matrix else 2
    if vis is None:
        vis = np.zeros((offset_out + num_stations) * nr, dtype=dt)
    vis = np.ascontiguousarray(vis, dtype=dt)
    num_sources = I_.size
    z = np.zeros(num_sources)
    arrs = [_c(x, dt) for x in (jones, I_, z if Q is None else Q,
                                z if U is None else U, z if V is None else V)]
    fn = (lib().oracle_auto_correlate_d if precision == "double"
          else lib().oracle_auto_correlate_f)
    fn(int(matrix), int(num_sources), int(num_stations), int(offset_out),
       *[_p(x) for x in arrs], _p(vis))
    return vis


def update_horizon_mask(precision, l, m, n, ha0, dec0, lat, mask=None):
    """mask |= (l ll + m mm + n nn > 0), oskar_update_horizon_mask
    semantics; returns the int32 mask."""
    dt = np.float64 if precision == "double" else np.float32
    arrs = [_c(x, dt) for x in (l, m, n)]
    num = arrs[0].size
    mask = (np.zeros(num, np.int32) if mask is None
            else np.ascontiguousarray(mask, np.int32).copy())
    fn = (lib().oracle_update_horizon_mask_d if precision == "double"
          else lib().oracle_update_horizon_mask_f)
    fn(num, *[_p(x) for x in arrs], float(ha0), float(dec0), float(lat),
       _p(mask))
    return mask


def prefix_sum(values):
    """Exclusive prefix sum of int32 values with the total appended."""
    v = np.ascontiguousarray(values, np.int32)
    out = np.zeros(v.size + 1, np.int32)
    lib().oracle_prefix_sum_int(v.size, _p(v), _p(out))
    return out


def sky_copy_source_data(precision, mask, table_in, cap_in, cap_out,
                         valid_in, nfix, table_out=None, valid_out=None):
    """Order-keeping masked row copy of a column table (num_col columns of
    cap_in values) and its per-row valid-column counts; returns (rows
    copied, table_out, valid_out)."""
    dt = np.float64 if precision == "double" else np.float32
    mask = np.ascontiguousarray(mask, np.int32)
    tin = np.ascontiguousarray(table_in, dt)
    num_col = tin.size // cap_in
    vin = np.ascontiguousarray(valid_in, np.int32)
    tout = (np.zeros(num_col * cap_out, dt) if table_out is None
            else np.ascontiguousarray(table_out, dt).copy())
    vout = (np.zeros(cap_out * nfix, np.int32) if valid_out is None
            else np.ascontiguousarray(valid_out, np.int32).copy())
    fn = (lib().oracle_sky_copy_source_data_d if precision == "double"
          else lib().oracle_sky_copy_source_data_f)
    r = fn(mask.size, cap_in, cap_out, _p(mask), num_col, _p(tin), _p(tout),
           nfix, _p(vin), _p(vout))
    return r, tout, vout


def dftw(precision, matrix, normalise, wavenumber, weights, x_in, y_in, z_in,
         x_out, y_out, z_out, data_idx, data, eval_x=1, eval_y=1,
         offset_coord_out=0, num_out=None, offset_out=0, output=None):
    dt = np.float64 if precision == "double" else np.float32
    num_in = x_in.size
    if num_out is None:
        num_out = x_out.size - offset_coord_out
    nr = 8 if matrix else 2
    if output is None:
        output = np.zeros((offset_out + num_out) * nr, dtype=dt)
    args = [_c(x, dt) for x in (weights, x_in, y_in, z_in)]
    outs = [_c(x, dt) for x in (x_out, y_out, z_out)]
    idx = _c(data_idx, np.int32)
    d = _c(data, dt)
    fn = lib().oracle_dftw_d if precision == "double" else lib().oracle_dftw_f
    fn(int(matrix), int(normalise), int(num_in), float(wavenumber),
       *[_p(x) for x in args], int(offset_coord_out), int(num_out),
       *[_p(x) for x in outs], _p(idx), _p(d), int(eval_x), int(eval_y),
       int(offset_out), _p(output))
    return output


def station_beam_aa(matrix, x_true, y_true, z_true, x_meas, y_meas, z_meas,
                    cable, element_types, type_kind, dipole_length_m,
                    beam_dir, x, y, z, frequency_hz, euler_x0=0.0,
                    euler_y0=0.0, swap_xy=0, is_3d=0, norm_array=0,
                    norm_element=0, enable_array=1, err=None, apod=None):
    """Element error factors (err, from element_errors) and apodisation
    (apod), complex per element as interleaved reals, multiply the DFT
    weights when given."""
    f8 = np.float64
    ne = x_true.size
    npnt = x.size
    beam = np.zeros(npnt * (8 if matrix else 2), dtype=f8)
    arr = [_c(v, f8) for v in (x_true, y_true, z_true, x_meas, y_meas,
                               z_meas, cable)]
    et = _c(element_types, np.int32)
    tk = _c(type_kind, np.int32)
    dl = _c(dipole_length_m, f8)
    xs = [_c(v, f8) for v in (x, y, z)]
    r = lib().oracle_station_beam_aa_d(
        int(matrix), ne, *[_p(v) for v in arr], _p(et), int(tk.size), _p(tk),
        _p(dl), float(euler_x0), float(euler_y0), int(swap_xy), int(is_3d),
        int(norm_array), int(norm_element), int(enable_array),
        float(beam_dir[0]), float(beam_dir[1]), float(beam_dir[2]), npnt,
        *[_p(v) for v in xs], float(frequency_hz),
        *[_p(_c(v, f8)) for v in (err, apod)], _p(beam))
    if r != 0:
        raise RuntimeError("oracle_station_beam_aa_d returned %d" % r)
    return beam


def jones_K(l, m, n, u, v, w, wavenumber, I_, fmin=-np.inf, fmax=np.inf):
    f8 = np.float64
    out = np.zeros(2 * u.size * l.size, dtype=f8)
    arr = [_c(x, f8) for x in (l, m, n, u, v, w, I_)]
    lib().oracle_evaluate_jones_K_d(
        l.size, _p(arr[0]), _p(arr[1]), _p(arr[2]), u.size, _p(arr[3]),
        _p(arr[4]), _p(arr[5]), float(wavenumber), _p(arr[6]), float(fmin),
        float(fmax), _p(out))
    return out


def factorised_subset(l, m, n, I_, E, u, v, w, wavenumber, p, q):
    f8 = np.float64
    arr = [_c(x, f8) for x in (l, m, n, I_, E, u, v, w)]
    pi = _c(p, np.int32)
    qi = _c(q, np.int32)
    out = np.zeros(8 * pi.size, dtype=f8)
    lib().oracle_factorised_subset_ld(
        l.size, _p(arr[0]), _p(arr[1]), _p(arr[2]), _p(arr[3]), _p(arr[4]),
        u.size, _p(arr[5]), _p(arr[6]), _p(arr[7]), float(wavenumber),
        pi.size, _p(pi), _p(qi), _p(out))
    return out


def station_to_baseline(u, v, w, num_stations, casa=False, offset_in=0,
                        offset_out=0, out=None):
    """Baseline (u, v, w) from station (u, v, w): the pair loop of
    oskar/convert/src/oskar_convert_station_uvw_to_baseline_uvw.c:18-36
    (s1 ascending, s2 from s1 + 1, b counting up from offset_out, value
    factor (x[s2] - x[s1]) with factor -1 under the CASA convention).
    Pure-Python loops: small cases only.  Returns (uu, vv, ww) in the input
    dtype (arithmetic in that dtype, as the reference's FP template)."""
    dt = np.asarray(u).dtype
    nb = num_stations * (num_stations - 1) // 2
    if out is None:
        out = [np.zeros(offset_out + nb, dtype=dt) for _ in range(3)]
    factor = dt.type(-1 if casa else 1)
    b = 0
    for s1 in range(num_stations):
        i1 = s1 + offset_in
        for s2 in range(s1 + 1, num_stations):
            i2 = s2 + offset_in
            for src, dst in zip((u, v, w), out):
                dst[b + offset_out] = factor * (src[i2] - src[i1])
            b += 1
    return tuple(out)


REF_BINARY_PATH = os.path.join(_HERE, "_ref", "liboskar_binary_ref.so")
_ref_binary = None


def ref_binary():
    """The reference's binary container library (oskar/binary/src/*.c,
    built by ``make -C oracle ref``) with prototypes set, or None."""
    global _ref_binary
    if _ref_binary is None and os.path.exists(REF_BINARY_PATH):
        L = C.CDLL(REF_BINARY_PATH)
        UB, SZ, IP, CP = C.c_ubyte, C.c_size_t, C.POINTER(C.c_int), C.c_char_p
        protos = {
            "oskar_binary_create": (P, [CP, C.c_char, IP]),
            "oskar_binary_free": (None, [P]),
            "oskar_binary_write": (None, [P, UB, UB, UB, I, SZ, P, IP]),
            "oskar_binary_write_ext": (None, [P, UB, CP, CP, I, SZ, P, IP]),
            "oskar_binary_write_int": (None, [P, UB, UB, I, I, IP]),
            "oskar_binary_write_double": (None, [P, UB, UB, I, D, IP]),
            "oskar_binary_num_tags": (I, [P]),
            "oskar_binary_query": (I, [P, UB, UB, UB, I,
                                       C.POINTER(C.c_size_t), IP]),
            "oskar_binary_query_ext": (I, [P, UB, CP, CP, I,
                                           C.POINTER(C.c_size_t), IP]),
            "oskar_binary_read": (None, [P, UB, UB, UB, I, SZ, P, IP]),
            "oskar_binary_read_ext": (None, [P, UB, CP, CP, I, SZ, P, IP]),
            "oskar_binary_read_int": (None, [P, UB, UB, I, IP, IP]),
            "oskar_binary_read_double": (None, [P, UB, UB, I,
                                                C.POINTER(C.c_double), IP]),
            "oskar_crc_create": (P, [I]),
            "oskar_crc_free": (None, [P]),
            "oskar_crc_compute": (C.c_ulong, [P, P, SZ]),
        }
        for name, (res, args) in protos.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _ref_binary = L
    return _ref_binary


def relative_directions(precision, ra, dec, ra0, dec0):
    """(l, m, n) of (ra, dec) about (ra0, dec0) in the precision given."""
    dt = np.float64 if precision == "d" else np.float32
    ra, dec = _c(ra, dt), _c(dec, dt)
    n = ra.size
    l, m, nn = (np.zeros(n, dt) for _ in range(3))
    getattr(lib(), "oracle_relative_directions_" + precision)(
        n, _p(ra), _p(dec), ra0, dec0, _p(l), _p(m), _p(nn))
    return l, m, nn


def gaussian_parameters(precision, ra, dec, maj, min_, pa, ra0, dec0,
                        a=None, b=None, c=None):
    """Gaussian a, b, c and the per-source fit-failure flags; a, b, c start
    from the given arrays (zeros if None) and keep their values where a
    source is skipped or fails.  Returns (a, b, c, failed)."""
    dt = np.float64 if precision == "d" else np.float32
    ra, dec, maj, min_, pa = (_c(x, dt) for x in (ra, dec, maj, min_, pa))
    n = ra.size
    a, b, c = (np.zeros(n, dt) if x is None else np.array(x, dt)
               for x in (a, b, c))
    failed = np.zeros(n, np.int32)
    getattr(lib(), "oracle_gaussian_parameters_" + precision)(
        n, _p(ra), _p(dec), _p(maj), _p(min_), _p(pa), ra0, dec0, _p(a),
        _p(b), _p(c), _p(failed))
    return a, b, c, failed


REF_SKY_PATH = os.path.join(_HERE, "_ref", "liboskar_sky_ref.so")
_ref_sky = None


def ref_sky():
    """The reference's Gaussian source parameter function with the sky
    table, oskar_Mem core, ellipse fit and conversions it calls, compiled
    from its sources by ``make -C oracle ref``; None if not built."""
    global _ref_sky
    if _ref_sky is None and os.path.exists(REF_SKY_PATH):
        L = C.CDLL(REF_SKY_PATH)
        IP = C.POINTER(C.c_int)
        protos = {
            "oskar_sky_create": (P, [I, I, I, IP]),
            "oskar_sky_free": (None, [P, IP]),
            "oskar_sky_column": (P, [P, I, I, IP]),
            "oskar_mem_void": (P, [P]),
            "oskar_sky_evaluate_gaussian_source_parameters": (
                None, [P, I, D, D, IP, IP]),
            "oskar_angular_distance": (D, [D, D, D, D]),
            "ref_lon_lat_to_relative_directions_3d_float": (
                None, [I, P, P, C.c_float, C.c_float, C.c_float, P, P, P]),
            "ref_lon_lat_to_relative_directions_3d_double": (
                None, [I, P, P, D, D, D, P, P, P]),
        }
        for name, (res, args) in protos.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _ref_sky = L
    return _ref_sky


def snap_beam_direction(xyz, az, el):
    """(x, y, z) snapped to the nearest permitted (az, el); returns
    ((x, y, z), index)."""
    x, y, z = (C.c_double(float(v)) for v in xyz)
    az = _c(az, np.float64)
    el = _c(el, np.float64)
    k = lib().oracle_snap_beam_direction(C.byref(x), C.byref(y), C.byref(z),
                                         az.size, _p(az), _p(el))
    return (x.value, y.value, z.value), k


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().oracle_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def random_gaussian(n, seed, c1, c2, c3):
    out = np.zeros(n)
    lib().oracle_random_gaussian_d(n, seed, c1, c2, c3, _p(out))
    return out


def element_errors(seed, time_index, station_id, gain, gain_error, phase,
                   phase_error):
    """Interleaved complex error factors of the elements."""
    f8 = np.float64
    arr = [_c(v, f8) for v in (gain, gain_error, phase, phase_error)]
    err = np.zeros(2 * arr[0].size)
    lib().oracle_element_errors_d(arr[0].size, seed, time_index, station_id,
                                  *[_p(v) for v in arr], _p(err))
    return err
