"""The .vis container (OSKAR binary format) on the host: CRC-32C, the chunk
writer and the indexed reader (include/oskar_binary.h, oskar_amd/csrc/
binary.c), against the reference's own binary library where it is built
(oracle/_ref/liboskar_binary_ref.so, compiled from oskar/binary/src/*.c)
and against golden files that library wrote (tests/golden/
make_vis_golden.py).
"""
import ctypes as C
import os
import shutil

import numpy as np
import pytest

import oracle
from oskar_amd import lib
from oskar_amd.vis import Binary

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
OSKAR_CHAR, OSKAR_INT, OSKAR_DOUBLE = 0x01, 0x02, 0x08
TAG_NOT_FOUND, CRC_FAIL = -115, -118


def crc(data, start=0):
    b = bytes(data)
    return lib.oskar_binary_crc32c(start, b, len(b))


def test_crc32c_known_answers():
    """CRC-32C check value and the iSCSI vectors of RFC 3720 B.4."""
    assert crc(b"123456789") == 0xE3069283
    assert crc(b"") == 0
    assert crc(bytes(32)) == 0x8A9136AA
    assert crc(b"\xff" * 32) == 0x62A8AB43
    assert crc(bytes(range(32))) == 0x46DD794E
    assert crc(bytes(range(31, -1, -1))) == 0x113FDB5C


def test_crc32c_continuation():
    rng = np.random.default_rng(1)
    a = rng.integers(0, 256, 1001, dtype=np.uint8).tobytes()
    for cut in (0, 1, 7, 8, 500, 1001):
        assert crc(a[cut:], crc(a[:cut])) == crc(a)


@pytest.mark.skipif(oracle.ref_binary() is None,
                    reason="reference binary library not built here")
def test_crc32c_matches_reference():
    ref = oracle.ref_binary()
    h = ref.oskar_crc_create(2)                  # OSKAR_CRC_32C
    rng = np.random.default_rng(2)
    for n in (0, 1, 3, 8, 9, 63, 64, 1000):
        a = rng.integers(0, 256, n, dtype=np.uint8)
        assert ref.oskar_crc_compute(h, a.ctypes.data, n) == crc(a)
    ref.oskar_crc_free(h)


def write_sequence(path):
    """The golden file's chunk sequence (Test_binary.c:54-104 and two more
    chunks), written here."""
    st = C.c_int(0)
    h = lib.oskar_binary_create(path.encode(), b"w", C.byref(st))
    lib.oskar_binary_write_int(h, 0, 0, 12345, 65, C.byref(st))
    for t, g, tag, idx, a in (
            (OSKAR_DOUBLE, 1, 10, 987654321, np.arange(22) + 1000.0),
            (OSKAR_INT, 2, 20, 1, np.arange(17, dtype=np.int32) * 10)):
        lib.oskar_binary_write(h, t, g, tag, idx, a.nbytes,
                               a.ctypes.data, C.byref(st))
    lib.oskar_binary_write_int(h, 0, 0, 2, 67, C.byref(st))
    a = np.arange(17, dtype=np.int32) * 75
    lib.oskar_binary_write(h, OSKAR_INT, 14, 5, 6, a.nbytes, a.ctypes.data,
                           C.byref(st))
    lib.oskar_binary_write_int(h, 12, 0, 0, 66, C.byref(st))
    a = np.arange(22) * 1234.0
    lib.oskar_binary_write(h, OSKAR_DOUBLE, 4, 0, 3, a.nbytes, a.ctypes.data,
                           C.byref(st))
    a = np.linspace(-1.0, 1.0, 9)
    lib.oskar_binary_write_ext(h, OSKAR_DOUBLE, b"my_group", b"my_tag_name",
                               42, a.nbytes, a.ctypes.data, C.byref(st))
    lib.oskar_binary_write(h, OSKAR_CHAR, 3, 1, 0, 0, None, C.byref(st))
    lib.oskar_binary_free(h)
    assert st.value == 0


def test_writer_bytes_identical_to_reference(tmp_path):
    p = str(tmp_path / "seq.dat")
    write_sequence(p)
    with open(p, "rb") as f, \
            open(os.path.join(GOLDEN, "binary_sequence.dat"), "rb") as g:
        assert f.read() == g.read()


def test_read_reference_sequence():
    """The reads and checks of Test_binary.c:106-175 on the golden file."""
    b = Binary(os.path.join(GOLDEN, "binary_sequence.dat"), "r")
    assert b.num_tags() == 9
    st = C.c_int(0)
    v = C.c_int(0)
    for grp, idx, want in ((0, 12345, 65), (12, 0, 66), (0, 2, 67)):
        lib.oskar_binary_read_int(b.ptr, grp, 0, idx, C.byref(v),
                                  C.byref(st))
        assert st.value == 0 and v.value == want
    assert np.array_equal(b.read(OSKAR_INT, 2, 20, 1, np.int32),
                          np.arange(17) * 10)
    assert np.array_equal(b.read(OSKAR_DOUBLE, 4, 0, 3, np.float64),
                          np.arange(22) * 1234.0)
    assert np.array_equal(b.read(OSKAR_INT, 14, 5, 6, np.int32),
                          np.arange(17) * 75)
    assert np.array_equal(b.read(OSKAR_DOUBLE, 1, 10, 987654321, np.float64),
                          np.arange(22) + 1000.0)
    # Type 0 matches any type.
    assert b.read(0, 1, 10, 987654321, np.float64)[0] == 1000.0
    # The extended chunk, by name.
    size = C.c_size_t(0)
    i = lib.oskar_binary_query_ext(b.ptr, OSKAR_DOUBLE, b"my_group",
                                   b"my_tag_name", 42, C.byref(size),
                                   C.byref(st))
    assert st.value == 0 and i == 7 and size.value == 72
    out = np.zeros(9)
    lib.oskar_binary_read_ext(b.ptr, OSKAR_DOUBLE, b"my_group",
                              b"my_tag_name", 42, out.nbytes,
                              out.ctypes.data, C.byref(st))
    assert st.value == 0 and np.array_equal(out, np.linspace(-1, 1, 9))
    # The empty chunk.
    assert b.read(OSKAR_CHAR, 3, 1, 0, np.uint8).size == 0
    # A tag that is not there (Test_binary.c:165-171).
    d = C.c_double(0)
    lib.oskar_binary_read_double(b.ptr, 255, 0, 0, C.byref(d), C.byref(st))
    assert st.value == TAG_NOT_FOUND
    st.value = 0
    lib.oskar_binary_query_ext(b.ptr, OSKAR_DOUBLE, b"my_group", b"other",
                               42, None, C.byref(st))
    assert st.value == TAG_NOT_FOUND
    b.close()


def test_corrupt_payload_fails_crc(tmp_path):
    p = str(tmp_path / "bad.dat")
    shutil.copy(os.path.join(GOLDEN, "binary_sequence.dat"), p)
    # Second chunk (22 doubles, user index 987654321) starts after the
    # 64-byte file header and the 24-byte first chunk; flip its 5th
    # payload byte.
    raw = bytearray(open(p, "rb").read())
    off = 64 + 20 + 4 + 4 + 20 + 4
    raw[off] ^= 0x10
    open(p, "wb").write(bytes(raw))
    b = Binary(p, "r")
    st = C.c_int(0)
    out = np.zeros(22)
    lib.oskar_binary_read(b.ptr, OSKAR_DOUBLE, 1, 10, 987654321, out.nbytes,
                          out.ctypes.data, C.byref(st))
    assert st.value == CRC_FAIL
    # Other chunks still read.
    assert np.array_equal(b.read(OSKAR_INT, 2, 20, 1, np.int32),
                          np.arange(17) * 10)
    b.close()
    ref = oracle.ref_binary()
    if ref is not None:
        st = C.c_int(0)
        h = ref.oskar_binary_create(p.encode(), b"r", C.byref(st))
        ref.oskar_binary_read(h, OSKAR_DOUBLE, 1, 10, 987654321, out.nbytes,
                              out.ctypes.data, C.byref(st))
        assert st.value == CRC_FAIL
        ref.oskar_binary_free(h)


def test_error_contract(tmp_path):
    st = C.c_int(0)
    h = lib.oskar_binary_create(str(tmp_path / "none.dat").encode(), b"r",
                                C.byref(st))
    assert st.value == -100 and not h           # OPEN_FAIL
    p = str(tmp_path / "w.dat")
    st.value = 0
    h = lib.oskar_binary_create(p.encode(), b"w", C.byref(st))
    v = C.c_int(0)
    lib.oskar_binary_read_int(h, 1, 1, 0, C.byref(v), C.byref(st))
    assert st.value == -104                     # NOT_OPEN_FOR_READ
    lib.oskar_binary_free(h)
    st.value = 0
    h = lib.oskar_binary_create(
        os.path.join(GOLDEN, "binary_sequence.dat").encode(), b"r",
        C.byref(st))
    lib.oskar_binary_write_int(h, 1, 1, 0, 5, C.byref(st))
    assert st.value == -105                     # NOT_OPEN_FOR_WRITE
    st.value = 0
    lib.oskar_binary_set_query_search_start(h, 9, C.byref(st))
    assert st.value == -117                     # TAG_OUT_OF_RANGE
    st.value = 0
    # Searching from chunk 4 no longer finds chunk 0.
    lib.oskar_binary_set_query_search_start(h, 4, C.byref(st))
    lib.oskar_binary_read_int(h, 0, 0, 12345, C.byref(v), C.byref(st))
    assert st.value == TAG_NOT_FOUND
    st.value = 0
    lib.oskar_binary_read_int(h, 12, 0, 0, C.byref(v), C.byref(st))
    assert st.value == 0 and v.value == 66
    lib.oskar_binary_free(h)
    # A file that is not an OSKAR binary file.
    q = str(tmp_path / "junk.dat")
    open(q, "wb").write(b"not a binary file" * 8)
    h = lib.oskar_binary_create(q.encode(), b"r", C.byref(st))
    assert st.value != 0
    lib.oskar_binary_free(h)
    # Truncated inside the last chunk.
    raw = open(os.path.join(GOLDEN, "binary_sequence.dat"), "rb").read()
    open(q, "wb").write(raw[:-30])
    st.value = 0
    h = lib.oskar_binary_create(q.encode(), b"r", C.byref(st))
    assert st.value != 0
    lib.oskar_binary_free(h)


def test_append_mode(tmp_path):
    p = str(tmp_path / "app.dat")
    st = C.c_int(0)
    for k in range(3):
        h = lib.oskar_binary_create(p.encode(), b"a", C.byref(st))
        lib.oskar_binary_write_double(h, 5, 9, k, 1.5 * k, C.byref(st))
        lib.oskar_binary_free(h)
    assert st.value == 0
    assert os.path.getsize(p) == 64 + 3 * (20 + 8 + 4)
    b = Binary(p, "r")
    assert b.num_tags() == 3
    for k in range(3):
        assert b.read(OSKAR_DOUBLE, 5, 9, k, np.float64)[0] == 1.5 * k
    b.close()


def random_chunks(rng, n):
    chunks = []
    for k in range(n):
        t = int(rng.choice([OSKAR_CHAR, OSKAR_INT, 0x04, OSKAR_DOUBLE,
                            0x04 | 0x20, 0x08 | 0x20 | 0x40]))
        es = {OSKAR_CHAR: 1, OSKAR_INT: 4, 0x04: 4, OSKAR_DOUBLE: 8,
              0x24: 8, 0x68: 64}[t]
        data = rng.integers(0, 256, es * int(rng.integers(0, 50)),
                            dtype=np.uint8)
        if rng.random() < 0.3:
            names = (("g%d" % rng.integers(0, 9)).encode(),
                     ("tag_%d" % k).encode())
        else:
            names = None
        chunks.append((t, int(rng.integers(0, 256)),
                       int(rng.integers(0, 256)), int(rng.integers(-5, 1 << 30)),
                       names, data))
    return chunks


def write_with(L, path, chunks):
    st = C.c_int(0)
    h = L.oskar_binary_create(path.encode(), b"w", C.byref(st))
    for t, g, tag, idx, names, data in chunks:
        ptr = data.ctypes.data if data.size else None
        if names:
            L.oskar_binary_write_ext(h, t, names[0], names[1], idx,
                                     data.nbytes, ptr, C.byref(st))
        else:
            L.oskar_binary_write(h, t, g, tag, idx, data.nbytes, ptr,
                                 C.byref(st))
    L.oskar_binary_free(h)
    assert st.value == 0


def read_with(L, path, chunks):
    st = C.c_int(0)
    h = L.oskar_binary_create(path.encode(), b"r", C.byref(st))
    assert st.value == 0
    assert L.oskar_binary_num_tags(h) == len(chunks)
    for t, g, tag, idx, names, data in chunks:
        out = np.zeros(max(data.size, 1), dtype=np.uint8)
        size = C.c_size_t(0)
        if names:
            L.oskar_binary_query_ext(h, t, names[0], names[1], idx,
                                     C.byref(size), C.byref(st))
            L.oskar_binary_read_ext(h, t, names[0], names[1], idx,
                                    out.nbytes, out.ctypes.data, C.byref(st))
        else:
            L.oskar_binary_query(h, t, g, tag, idx, C.byref(size),
                                 C.byref(st))
            L.oskar_binary_read(h, t, g, tag, idx, out.nbytes,
                                out.ctypes.data, C.byref(st))
        assert st.value == 0
        assert size.value == data.size
        assert np.array_equal(out[:data.size], data)
    L.oskar_binary_free(h)


@pytest.mark.skipif(oracle.ref_binary() is None,
                    reason="reference binary library not built here")
@pytest.mark.parametrize("seed", [3, 4])
def test_interchange_with_reference(tmp_path, seed):
    """Random chunk sequences (unique keys): files written here are byte
    for byte the reference's, and each library reads the other's."""
    rng = np.random.default_rng(seed)
    chunks = random_chunks(rng, 40)
    keys = set()
    uniq = []
    for c in chunks:
        k = (c[1], c[2], c[3], c[4])
        if k not in keys:
            keys.add(k)
            uniq.append(c)
    ref = oracle.ref_binary()
    a, b = str(tmp_path / "ours.dat"), str(tmp_path / "ref.dat")
    write_with(lib, a, uniq)
    write_with(ref, b, uniq)
    assert open(a, "rb").read() == open(b, "rb").read()
    read_with(lib, b, uniq)
    read_with(ref, a, uniq)
