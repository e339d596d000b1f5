"""Visibility headers and blocks on the host (include/oskar_vis.h,
oskar_amd/csrc/vis.c): the reference's own .vis test file (written by the
reference's binary library, tests/golden/vis_sequence.vis) read through
oskar_vis_header_read / oskar_vis_block_read; header and block round trips
through this library's writer; this writer's chunks read back with the
reference's binary library in the order of the reference's header reader
(oskar/vis/src/oskar_vis_header_read.c); block and header bookkeeping.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from oskar_amd import lib
from oskar_amd.mem import OSKAR_CPU, OSKAR_DOUBLE, OSKAR_SINGLE, Mem
from oskar_amd.vis import Binary, VisBlock, VisHeader

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
OSKAR_CHAR, OSKAR_INT = 0x01, 0x02
CPX, MAT = 0x20, 0x40
GRP_META, GRP_HDR, GRP_BLK = 1, 11, 12
POL_STOKES_I, POL_STOKES_Q, POL_LINEAR = 1, 2, 10


def test_read_reference_vis_file():
    """Values of oskar/binary/test/Test_binary_vis_read_write.c:93-158."""
    b = Binary(os.path.join(GOLDEN, "vis_sequence.vis"), "r")
    hdr = VisHeader.read(b)
    g = hdr.get
    assert g("amp_type") == OSKAR_DOUBLE | CPX | MAT
    assert g("coord_precision") == OSKAR_DOUBLE
    assert g("max_times_per_block") == 10
    assert g("num_times_total") == 33
    assert g("max_channels_per_block") == 2
    assert g("num_channels_total") == 2
    assert g("num_stations") == 4
    assert g("num_tags_per_block") == 5
    assert g("write_auto_correlations") == 0
    assert g("write_cross_correlations") == 1
    # The file has no phase convention or polarisation chunks: defaults.
    assert g("casa_phase_convention") == 0
    assert g("pol_type") == POL_LINEAR
    assert (g("phase_centre_ra_deg"), g("phase_centre_dec_deg")) == (10., 20.)
    assert (g("freq_start_hz"), g("freq_inc_hz")) == (100e6, 10e6)
    assert g("channel_bandwidth_hz") == 4e3
    assert (g("time_start_mjd_utc"), g("time_inc_sec")) == (51544.0, 10.0)
    assert g("time_average_sec") == 0.08
    assert (g("telescope_lon_deg"), g("telescope_lat_deg"),
            g("telescope_alt_metres")) == (-30.0, 20.0, 10.0)
    for k, o in enumerate((0.1, 0.2, 0.3)):
        x = hdr.mem("station_offset_ecef_metres", k).numpy()
        assert np.array_equal(x, np.arange(4) + o)
    path = hdr.mem("telescope_path").numpy()
    assert bytes(path).rstrip(b"\0") == b"my_telescope_model"
    # 25 header chunks, counted exactly, so the block search starts right.
    assert g("num_tags_header") == 25

    blk = VisBlock(OSKAR_CPU, header=hdr)
    assert blk.get("num_baselines") == 6 and blk.get("num_pols") == 4
    i = np.arange(10 * 6 * 2, dtype=np.float64)
    want = (i[:, None] + np.array([0.111, 0.222, 0.333, 0.444, 0.555, 0.666,
                                   0.777, 0.888])).ravel()
    for block in range(4):
        blk.read(hdr, b, block)
        assert blk.get("start_time_index") == 10 * block
        assert blk.get("num_times") == (10 if block < 3 else 3)
        assert blk.get("num_channels") == 2
        assert np.array_equal(blk.cross().numpy(), want)
        for k, o in enumerate((0.11, 0.22, 0.33)):
            assert np.array_equal(blk.baseline_uvw(k).numpy(),
                                  np.arange(60) + o)
    b.close()


def fill(dst, arr, type_code):
    """Copies a numpy array into an existing oskar_Mem (resizing it)."""
    src = Mem.from_numpy(arr, type_code, OSKAR_CPU)
    st = C.c_int(0)
    lib.oskar_mem_copy(dst.ptr, src.ptr, C.byref(st))
    assert st.value == 0
    src.free()


def make_header(tmp_path, amp_type, coord, ns, auto, cross, casa=1):
    hdr = VisHeader(amp_type, coord, 3, 7, 2, 5, ns, auto, cross)
    lib.oskar_vis_header_set_casa_phase_convention(hdr.ptr, casa)
    lib.oskar_vis_header_set_freq_start_hz(hdr.ptr, 1.5e8)
    lib.oskar_vis_header_set_freq_inc_hz(hdr.ptr, 2.5e5)
    lib.oskar_vis_header_set_channel_bandwidth_hz(hdr.ptr, 1e4)
    lib.oskar_vis_header_set_time_start_mjd_utc(hdr.ptr, 60000.25)
    lib.oskar_vis_header_set_time_inc_sec(hdr.ptr, 7.5)
    lib.oskar_vis_header_set_time_average_sec(hdr.ptr, 0.9)
    lib.oskar_vis_header_set_phase_centre(hdr.ptr, 0, 201.365, -43.019)
    lib.oskar_vis_header_set_telescope_centre(hdr.ptr, 116.7, -26.8, 377.0)
    rng = np.random.default_rng(11)
    dt = np.float64 if coord == OSKAR_DOUBLE else np.float32
    for k in range(3):
        fill(hdr.mem("station_offset_ecef_metres", k),
             rng.normal(size=ns).astype(dt), coord)
        for s in range(ns):
            fill(hdr.mem("element_enu_metres", k, s),
                 rng.normal(size=s + 1).astype(dt), coord)
    for s in range(ns):
        fill(hdr.mem("station_name", s),
             np.frombuffer(("station_%03d\0" % s).encode(), np.int8),
             OSKAR_CHAR)
        for d in range(2):
            for a in range(3):
                fill(hdr.mem("element_feed_angle_rad", d, a, s),
                     np.full(s % 3, 0.1 * (d + 2 * a)), OSKAR_DOUBLE)
    fill(hdr.mem("station_diameter_m"), np.linspace(30, 40, ns), OSKAR_DOUBLE)
    fill(hdr.mem("settings"), np.frombuffer(
        b"[interferometer]\noversample=1\n\0", np.int8), OSKAR_CHAR)
    fill(hdr.mem("telescope_path"), np.frombuffer(b"tel.tm\0", np.int8),
         OSKAR_CHAR)
    return hdr


HEADER_SCALARS = ("amp_type", "coord_precision", "max_times_per_block",
                  "num_times_total", "max_channels_per_block",
                  "num_channels_total", "num_stations", "num_tags_per_block",
                  "write_auto_correlations", "write_cross_correlations",
                  "pol_type", "casa_phase_convention", "freq_start_hz",
                  "freq_inc_hz", "channel_bandwidth_hz",
                  "time_start_mjd_utc", "time_inc_sec", "time_average_sec",
                  "phase_centre_type", "phase_centre_ra_deg",
                  "phase_centre_dec_deg", "telescope_lon_deg",
                  "telescope_lat_deg", "telescope_alt_metres")


@pytest.mark.parametrize("amp,coord,auto,cross", [
    (OSKAR_SINGLE | CPX | MAT, OSKAR_SINGLE, 1, 1),
    (OSKAR_DOUBLE | CPX, OSKAR_DOUBLE, 0, 1),
    (OSKAR_DOUBLE | CPX | MAT, OSKAR_DOUBLE, 1, 0)])
def test_header_round_trip(tmp_path, amp, coord, auto, cross):
    ns = 5
    hdr = make_header(tmp_path, amp, coord, ns, auto, cross)
    p = str(tmp_path / "h.vis")
    hdr.write(p).close()
    b = Binary(p, "r")
    got = VisHeader.read(b)
    for name in HEADER_SCALARS:
        assert got.get(name) == hdr.get(name), name
    assert got.get("num_tags_per_block") == 1 + 4 * cross + auto
    for name, args in [("station_offset_ecef_metres", (k,)) for k in
                       range(3)] + [("element_enu_metres", (k, s)) for k in
                                    range(3) for s in range(ns)] + \
            [("station_name", (s,)) for s in range(ns)] + \
            [("element_feed_angle_rad", (d, a, s)) for d in range(2) for a
             in range(3) for s in range(ns)] + \
            [("station_diameter_m", ()), ("settings", ()),
             ("telescope_path", ())]:
        x = hdr.mem(name, *args).numpy()
        y = got.mem(name, *args).numpy()
        assert x.dtype == y.dtype and np.array_equal(x, y), (name, args)
    # Every chunk after the metadata (date, version, cwd, user) belongs to
    # the header, and all of them are counted.
    n_meta = sum(1 for i in range(b.num_tags())
                 if _chunk_group(p, i) == GRP_META)
    assert got.get("num_tags_header") == b.num_tags() - n_meta
    b.close()


def _chunk_group(path, index):
    """Group byte of chunk index (walks the 20-byte tags)."""
    raw = open(path, "rb").read()
    off = 64
    for _ in range(index):
        off += 20 + int.from_bytes(raw[off + 12:off + 20], "little")
    return raw[off + 6]


@pytest.mark.skipif(oracle.ref_binary() is None,
                    reason="reference binary library not built here")
def test_header_chunks_read_by_reference(tmp_path):
    """The reference header reader's chunk requests
    (oskar_vis_header_read.c:44-150) against this writer's file, through
    the reference's binary library."""
    ref = oracle.ref_binary()
    ns = 4
    amp = OSKAR_DOUBLE | CPX | MAT
    hdr = make_header(tmp_path, amp, OSKAR_DOUBLE, ns, 1, 1)
    p = str(tmp_path / "h.vis")
    hdr.write(p).close()
    st = C.c_int(0)
    h = ref.oskar_binary_create(p.encode(), b"r", C.byref(st))
    v = C.c_int(0)
    ints = {3: 1, 4: 1, 5: amp, 6: OSKAR_DOUBLE, 7: 3, 8: 7, 9: 2, 10: 5,
            11: ns, 2: 6, 13: 1, 12: POL_LINEAR, 21: 0}
    for tag, want in ints.items():
        ref.oskar_binary_read_int(h, GRP_HDR, tag, 0, C.byref(v), C.byref(st))
        assert st.value == 0 and v.value == want, tag
    d = C.c_double(0)
    reals = {23: 1.5e8, 24: 2.5e5, 25: 1e4, 26: 60000.25, 27: 7.5, 28: 0.9,
             29: 116.7, 30: -26.8, 31: 377.0}
    for tag, want in reals.items():
        ref.oskar_binary_read_double(h, GRP_HDR, tag, 0, C.byref(d),
                                     C.byref(st))
        assert st.value == 0 and d.value == want, tag
    pc = np.zeros(2)
    ref.oskar_binary_read(h, OSKAR_DOUBLE, GRP_HDR, 22, 0, 16,
                          pc.ctypes.data, C.byref(st))
    assert st.value == 0 and tuple(pc) == (201.365, -43.019)
    size = C.c_size_t(0)
    for k in range(3):
        ref.oskar_binary_query(h, OSKAR_DOUBLE, GRP_HDR, 32 + k, 0,
                               C.byref(size), C.byref(st))
        assert st.value == 0 and size.value == 8 * ns
        for s in range(ns):
            ref.oskar_binary_query(h, OSKAR_DOUBLE, GRP_HDR, 35 + k, s,
                                   C.byref(size), C.byref(st))
            assert st.value == 0 and size.value == 8 * (s + 1)
    for s in range(ns):
        ref.oskar_binary_query(h, OSKAR_CHAR, GRP_HDR, 41, s, C.byref(size),
                               C.byref(st))
        assert st.value == 0 and size.value == len("station_000") + 1
        for tag in range(43, 49):
            ref.oskar_binary_query(h, OSKAR_DOUBLE, GRP_HDR, tag, s,
                                   C.byref(size), C.byref(st))
            assert st.value == 0 and size.value == 8 * (s % 3)
    ref.oskar_binary_query(h, OSKAR_CHAR, 3, 2, 0, C.byref(size), C.byref(st))
    assert st.value == 0                        # settings, group 3 tag 2
    ref.oskar_binary_query(h, OSKAR_CHAR, GRP_META, 2, 0, C.byref(size),
                           C.byref(st))
    assert st.value == 0                        # version string
    ref.oskar_binary_free(h)


@pytest.mark.parametrize("amp", [OSKAR_SINGLE | CPX | MAT, OSKAR_DOUBLE | CPX,
                                 OSKAR_SINGLE | CPX])
def test_block_round_trip_host(tmp_path, amp):
    """Blocks with auto- and cross-correlations and baseline (u, v, w)
    (the form written when no station (u, v, w) is held)."""
    ns, nt, nc = 6, 3, 2
    hdr = make_header(tmp_path, amp, amp & 0x0C, ns, 1, 1)
    p = str(tmp_path / "b.vis")
    h = hdr.write(p)
    blk = VisBlock(OSKAR_CPU, header=hdr)
    nb = ns * (ns - 1) // 2
    assert blk.get("num_baselines") == nb
    assert blk.get("num_pols") == (4 if amp & MAT else 1)
    rng = np.random.default_rng(5)
    prec = amp & 0x0C
    dt = np.float64 if prec == OSKAR_DOUBLE else np.float32
    per = (8 if amp & MAT else 2)
    written = []
    for i in range(3):
        lib.oskar_vis_block_set_start_time_index(blk.ptr, 3 * i)
        cross = rng.normal(size=nt * nc * nb * per).astype(dt)
        auto = rng.normal(size=nt * nc * ns * per).astype(dt)
        uvw = [rng.normal(size=nt * nb).astype(dt) for _ in range(3)]
        fill(blk.cross(), cross, amp)
        fill(blk.auto(), auto, amp)
        for k in range(3):
            fill(blk.baseline_uvw(k), uvw[k], prec)
        blk.write(h, i)
        written.append((cross, auto, uvw))
    h.close()
    b = Binary(p, "r")
    got_hdr = VisHeader.read(b)
    got = VisBlock(OSKAR_CPU, header=got_hdr)
    for i in (2, 0, 1):
        got.read(got_hdr, b, i)
        cross, auto, uvw = written[i]
        assert got.get("start_time_index") == 3 * i
        assert np.array_equal(got.cross().numpy(), cross)
        assert np.array_equal(got.auto().numpy(), auto)
        for k in range(3):
            assert np.array_equal(got.baseline_uvw(k).numpy(), uvw[k])
    b.close()


def test_block_bookkeeping():
    blk = VisBlock(OSKAR_CPU, OSKAR_DOUBLE | CPX | MAT, 4, 3, 10, 1, 1)
    assert [blk.get(n) for n in ("num_times", "num_channels", "num_baselines",
                                 "num_stations", "num_pols")] == \
        [4, 3, 45, 10, 4]
    assert len(blk.cross()) == 4 * 3 * 45 and len(blk.auto()) == 4 * 3 * 10
    assert len(blk.baseline_uvw(0)) == 0 and len(blk.station_uvw(2)) == 0
    st = C.c_int(0)
    lib.oskar_vis_block_resize(blk.ptr, 2, 5, 7, C.byref(st))
    assert st.value == 0
    assert len(blk.cross()) == 2 * 5 * 21 and len(blk.auto()) == 2 * 5 * 7
    lib.oskar_vis_block_clear(blk.ptr, C.byref(st))
    assert not blk.cross().numpy().any()
    auto_only = VisBlock(OSKAR_CPU, OSKAR_SINGLE | CPX, 2, 2, 5, 0, 1)
    assert auto_only.get("num_baselines") == 0
    assert auto_only.get("has_cross_correlations") == 0
    assert auto_only.get("has_auto_correlations") == 1
    for bad in (OSKAR_DOUBLE, OSKAR_INT | CPX):
        st.value = 0
        p = lib.oskar_vis_block_create(OSKAR_CPU, bad, 1, 1, 2, 1, 0,
                                       C.byref(st))
        assert st.value == -13 and not p        # BAD_DATA_TYPE


def test_header_pol_types_and_errors():
    st = C.c_int(0)
    mat = VisHeader(OSKAR_SINGLE | CPX | MAT, OSKAR_SINGLE, 0, 8, 0, 4, 3)
    assert mat.get("pol_type") == POL_LINEAR
    assert mat.get("max_times_per_block") == 8       # <= 0 means the total
    assert mat.get("max_channels_per_block") == 4
    lib.oskar_vis_header_set_pol_type(mat.ptr, POL_STOKES_Q, C.byref(st))
    assert st.value == -9                            # TYPE_MISMATCH
    st.value = 0
    lib.oskar_vis_header_set_pol_type(mat.ptr, 0, C.byref(st))
    assert st.value == 0 and mat.get("pol_type") == 0
    sc = VisHeader(OSKAR_DOUBLE | CPX, OSKAR_DOUBLE, 2, 8, 1, 4, 3)
    assert sc.get("pol_type") == POL_STOKES_I
    lib.oskar_vis_header_set_pol_type(sc.ptr, POL_LINEAR, C.byref(st))
    assert st.value == -9
    st.value = 0
    p = lib.oskar_vis_header_create(OSKAR_DOUBLE, OSKAR_DOUBLE, 1, 1, 1, 1,
                                    2, 0, 1, C.byref(st))
    assert st.value == -13 and not p


def test_station_to_baseline_is_hip_only():
    st = C.c_int(0)
    u = [Mem.from_numpy(np.arange(4.0), OSKAR_DOUBLE, OSKAR_CPU)
         for _ in range(3)]
    uu = [Mem(OSKAR_DOUBLE, OSKAR_CPU, 6) for _ in range(3)]
    lib.oskar_convert_station_uvw_to_baseline_uvw(
        0, 4, 0, u[0].ptr, u[1].ptr, u[2].ptr, 0, uu[0].ptr, uu[1].ptr,
        uu[2].ptr, C.byref(st))
    assert st.value == -14                           # BAD_LOCATION
