"""Writes the .vis container fixtures under tests/golden/ with the
reference's own binary library (oskar/binary/src/*.c), compiled from its
sources by ``make -C oracle ref`` into oracle/_ref/liboskar_binary_ref.so.

Run in a container that has the reference tree:
    make -C oracle ref && python tests/golden/make_vis_golden.py

  binary_sequence.dat  the chunk sequence and values of the reference's
                       oskar/binary/test/Test_binary.c:54-104 (ints, a
                       22-double and two 17-int arrays under assorted
                       groups, tags and user indices), then one extended
                       (named) chunk and one empty chunk.
  vis_sequence.vis     the file of oskar/binary/test/Test_binary_vis_read_
                       write.c:93-249: a visibility header (4 stations,
                       2 channels, 33 times in blocks of 10, double complex
                       matrix amplitudes) and 4 blocks of cross-correlations
                       with baseline (u, v, w), values i + 0.111 ... i +
                       0.888 and i + 0.11 ... i + 0.33.
Both files hold no time stamps, so they are reproducible byte for byte.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))
import oracle  # noqa: E402

OSKAR_CHAR, OSKAR_INT, OSKAR_DOUBLE = 0x01, 0x02, 0x08
OSKAR_DOUBLE_COMPLEX_MATRIX = 0x08 | 0x20 | 0x40
GRP_HDR, GRP_BLK = 11, 12


class Writer:
    def __init__(self, lib, path):
        self.lib = lib
        self.st = C.c_int(0)
        self.h = lib.oskar_binary_create(path.encode(), b"w",
                                         C.byref(self.st))

    def arr(self, dtype_code, grp, tag, idx, a):
        a = np.ascontiguousarray(a)
        self.lib.oskar_binary_write(self.h, dtype_code, grp, tag, idx,
                                    a.nbytes, a.ctypes.data, C.byref(self.st))

    def int(self, grp, tag, idx, v):
        self.lib.oskar_binary_write_int(self.h, grp, tag, idx, v,
                                        C.byref(self.st))

    def dbl(self, grp, tag, idx, v):
        self.lib.oskar_binary_write_double(self.h, grp, tag, idx, v,
                                           C.byref(self.st))

    def close(self):
        self.lib.oskar_binary_free(self.h)
        if self.st.value:
            raise RuntimeError("reference writer status %d" % self.st.value)


def binary_sequence(w):
    """The sequence of Test_binary.c:54-104, plus an extended and an empty
    chunk."""
    w.int(0, 0, 12345, 65)
    w.arr(OSKAR_DOUBLE, 1, 10, 987654321, np.arange(22) + 1000.0)
    w.arr(OSKAR_INT, 2, 20, 1, np.arange(17, dtype=np.int32) * 10)
    w.int(0, 0, 2, 67)
    w.arr(OSKAR_INT, 14, 5, 6, np.arange(17, dtype=np.int32) * 75)
    w.int(12, 0, 0, 66)
    w.arr(OSKAR_DOUBLE, 4, 0, 3, np.arange(22) * 1234.0)
    a = np.linspace(-1.0, 1.0, 9)
    w.lib.oskar_binary_write_ext(w.h, OSKAR_DOUBLE, b"my_group",
                                 b"my_tag_name", 42, a.nbytes, a.ctypes.data,
                                 C.byref(w.st))
    w.lib.oskar_binary_write(w.h, OSKAR_CHAR, 3, 1, 0, 0, None,
                             C.byref(w.st))


VIS_SEQUENCE = dict(
    telescope_path=b"my_telescope_model\0", num_tags_per_block=5,
    amp_type=OSKAR_DOUBLE_COMPLEX_MATRIX, coord_precision=OSKAR_DOUBLE,
    max_times_per_block=10, num_times_total=33, num_channels_total=2,
    num_stations=4, phase_centre_deg=(10.0, 20.0), lon=-30.0, lat=20.0,
    alt=10.0, freq_start_hz=100e6, freq_inc_hz=10e6,
    channel_bandwidth_hz=4e3, time_start_mjd_utc=51544.0, time_inc_sec=10.0,
    time_average_sec=0.08)


def vis_arrays(s):
    """Station, baseline and amplitude arrays of Test_binary_vis_read_
    write.c:129-158."""
    ns = s["num_stations"]
    nb = ns * (ns - 1) // 2
    ntb = s["max_times_per_block"] * nb
    st = [np.arange(ns) + o for o in (0.1, 0.2, 0.3)]
    uvw = [np.arange(ntb) + o for o in (0.11, 0.22, 0.33)]
    i = np.arange(ntb * s["num_channels_total"], dtype=np.float64)
    vis = i[:, None] + np.array([0.111, 0.222, 0.333, 0.444, 0.555, 0.666,
                                 0.777, 0.888])[None, :]
    return st, uvw, vis


def vis_sequence(w):
    s = VIS_SEQUENCE
    st, uvw, vis = vis_arrays(s)
    ns = s["num_stations"]
    nb = ns * (ns - 1) // 2
    w.arr(OSKAR_CHAR, GRP_HDR, 1, 0,
          np.frombuffer(s["telescope_path"], dtype=np.uint8))
    w.int(GRP_HDR, 2, 0, s["num_tags_per_block"])
    for tag, v in ((3, 0), (4, 1), (5, s["amp_type"]),
                   (6, s["coord_precision"]), (7, s["max_times_per_block"]),
                   (8, s["num_times_total"]), (9, s["num_channels_total"]),
                   (10, s["num_channels_total"]), (11, ns), (21, 0)):
        w.int(GRP_HDR, tag, 0, v)
    w.arr(OSKAR_DOUBLE, GRP_HDR, 22, 0, np.array(s["phase_centre_deg"]))
    for tag, v in ((23, s["freq_start_hz"]), (24, s["freq_inc_hz"]),
                   (25, s["channel_bandwidth_hz"]),
                   (26, s["time_start_mjd_utc"]), (27, s["time_inc_sec"]),
                   (28, s["time_average_sec"]), (29, s["lon"]),
                   (30, s["lat"]), (31, s["alt"])):
        w.dbl(GRP_HDR, tag, 0, v)
    for k in range(3):
        w.arr(OSKAR_DOUBLE, GRP_HDR, 32 + k, 0, st[k])
    mt, nt = s["max_times_per_block"], s["num_times_total"]
    for b in range((nt + mt - 1) // mt):
        times = min(mt, nt - b * mt)
        w.arr(OSKAR_INT, GRP_BLK, 1, b, np.array(
            [b * mt, 0, times, s["num_channels_total"], nb, ns],
            dtype=np.int32))
        w.arr(s["amp_type"], GRP_BLK, 3, b, vis)
        for k in range(3):
            w.arr(OSKAR_DOUBLE, GRP_BLK, 4 + k, b, uvw[k])


def main():
    lib = oracle.ref_binary()
    if lib is None:
        raise SystemExit("oracle/_ref/liboskar_binary_ref.so missing: "
                         "run `make -C oracle ref` first")
    for name, fn in (("binary_sequence.dat", binary_sequence),
                     ("vis_sequence.vis", vis_sequence)):
        path = os.path.join(HERE, name)
        w = Writer(lib, path)
        fn(w)
        w.close()
        print(name, os.path.getsize(path))


if __name__ == "__main__":
    main()
