"""GPU parity of the horizon clip path through the C-ABI: integer and row
results bit for bit against the reference's golden vectors and the CPU
oracle.

  oskar_update_horizon_mask  golden masks of the reference's CPU body,
                             including directions within ~1e-7 of the
                             horizon plane, each case and ORed over three;
  oskar_prefix_sum           numpy cumulative sums for lengths around the
                             2048-value tile and 1M values, aligned and
                             misaligned (views at an odd offset);
  oskar_sky_copy_source_data golden row copy of the reference's body;
  oskar_sky_horizon_clip     the oracle's mask over every station model,
                             its prefix sum and its order-keeping row copy,
                             for 3 models (fused), 20 models (masks ORed
                             first) and none; 1M sources carrying their own
                             index keep their order.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from gpu_util import sync_check
from helpers import load_golden
from oskar_amd import lib
from oskar_amd.mem import (OSKAR_CPU, OSKAR_DOUBLE, OSKAR_GPU, OSKAR_INT,
                           OSKAR_SINGLE, Mem, Sky, StationWork, Telescope,
                           _Borrowed, check_status)

pytestmark = pytest.mark.gpu

SCRATCH_L, SCRATCH_M, SCRATCH_N = 0x103, 0x104, 0x105
RA, DEC, I_JY, Q_JY, U_JY, V_JY = 1, 2, 3, 4, 5, 6
NFIX = 20


def pcode(prec):
    return OSKAR_DOUBLE if prec == "d" else OSKAR_SINGLE


@pytest.mark.parametrize("prec", ["f", "d"])
def test_horizon_mask_gpu_matches_golden(hip, prec):
    print("horizon mask golden", prec, flush=True)
    g = load_golden("sky_clip.npz")
    P = pcode(prec)
    l, m, n = (Mem.from_numpy(g[k], P) for k in "lmn")
    num = g["l"].size
    acc = Mem.from_numpy(np.zeros(num, np.int32), OSKAR_INT)
    for k, (ha0, dec0, lat) in enumerate(g["cases"]):
        one = Mem.from_numpy(np.zeros(num, np.int32), OSKAR_INT)
        for mask in (one, acc):
            s = C.c_int(0)
            lib.oskar_update_horizon_mask(num, l.ptr, m.ptr, n.ptr,
                                          float(ha0), float(dec0),
                                          float(lat), mask.ptr, C.byref(s))
            check_status(s, "oskar_update_horizon_mask")
        sync_check("horizon mask")
        assert np.array_equal(one.numpy(), g["mask_%s_%d" % (prec, k)])
    assert np.array_equal(acc.numpy(), g["mask_%s_or" % prec])


def gpu_prefix_sum(values, offset=0):
    src = Mem.from_numpy(np.concatenate([np.zeros(offset, np.int32),
                                         values]), OSKAR_INT)
    dst = Mem(OSKAR_INT, OSKAR_GPU, values.size + 1 + offset)
    s = C.c_int(0)
    vin = lib.oskar_mem_create_alias(src.ptr, offset, values.size,
                                     C.byref(s))
    vout = lib.oskar_mem_create_alias(dst.ptr, offset, values.size + 1,
                                      C.byref(s))
    lib.oskar_prefix_sum(values.size, vin, vout, C.byref(s))
    check_status(s, "oskar_prefix_sum")
    sync_check("prefix sum")
    out = _Borrowed(vout).numpy()
    lib.oskar_mem_free(vin, C.byref(C.c_int(0)))
    lib.oskar_mem_free(vout, C.byref(C.c_int(0)))
    return out


@pytest.mark.parametrize("num", [0, 1, 7, 2047, 2048, 2049, 6144 + 5,
                                 1_000_003])
@pytest.mark.parametrize("offset", [0, 1])
def test_prefix_sum_gpu_is_cumulative_sum(hip, num, offset):
    print("prefix sum", num, offset, flush=True)
    rng = np.random.default_rng(num + offset)
    v = rng.integers(-3, 6, num).astype(np.int32)
    got = gpu_prefix_sum(v, offset)
    assert np.array_equal(got, oracle.prefix_sum(v))


def make_sky(P, num, columns, location=OSKAR_GPU):
    sky = Sky(P, location, num)
    for ctype, values in columns:
        sky.set_column(ctype, values)
    return sky


@pytest.mark.parametrize("prec", ["f", "d"])
def test_sky_copy_gpu_matches_golden(hip, prec):
    print("sky copy golden", prec, flush=True)
    g = load_golden("sky_clip.npz")
    P = pcode(prec)
    ncol, cap_in, cap_out = (int(g[k]) for k in ("ncol", "cap_in",
                                                 "cap_out"))
    num = g["copy_mask"].size
    tin = g["table_in"].reshape(ncol, cap_in)
    types = [RA, DEC, I_JY, Q_JY, U_JY, V_JY, SCRATCH_L]
    sky_in = make_sky(P, num, [(t, tin[c]) for c, t in enumerate(types)])
    assert sky_in.attr(Sky.CAPACITY) == cap_in
    # Valid-column counts of the first 64 rows (element-wise setter).
    vin = g["valid_in"].reshape(cap_in, NFIX)
    for r in range(64):
        for t in range(NFIX):
            lib.oskar_sky_set_num_valid_columns_of_type(sky_in.ptr, t, r,
                                                        int(vin[r, t]))
    sky_out = Sky(P, OSKAR_GPU, 4000)
    assert sky_out.attr(Sky.CAPACITY) == cap_out
    mask = Mem.from_numpy(g["copy_mask"], OSKAR_INT)
    idx = Mem.from_numpy(g["copy_indices"], OSKAR_INT)
    s = C.c_int(0)
    lib.oskar_sky_copy_source_data(sky_in.ptr, mask.ptr, idx.ptr,
                                   sky_out.ptr, C.byref(s))
    check_status(s, "oskar_sky_copy_source_data")
    sync_check("sky copy")
    total = int(g["copy_indices"][-1])
    assert sky_out.attr(Sky.NUM_SOURCES) == total
    assert sky_out.attr(Sky.CAPACITY) == cap_out
    want = g["table_out_%s" % prec].reshape(ncol, cap_out)
    for c, t in enumerate(types):
        assert lib.oskar_sky_column_type(sky_out.ptr, c) == t
        assert np.array_equal(sky_out.get_column(t), want[c, :total])
    kept = np.nonzero(g["copy_mask"])[0]
    for r in range(total):
        if kept[r] >= 64:
            break
        got = [lib.oskar_sky_num_valid_columns_of_type(sky_out.ptr, t, r)
               for t in range(NFIX)]
        assert got == list(vin[kept[r]])


def clip_case(prec, num, models, gast, seed, extra_index=False):
    rng = np.random.default_rng(seed)
    P = pcode(prec)
    dt = np.float64 if prec == "d" else np.float32
    ra0, dec0 = 1.1, -0.6
    ra = rng.uniform(0, 2 * np.pi, num)
    dec = np.arcsin(rng.uniform(-1, 1, num))
    d = ra - ra0
    l = (np.cos(dec) * np.sin(d)).astype(dt)
    m = (np.cos(dec0) * np.sin(dec) -
         np.sin(dec0) * np.cos(dec) * np.cos(d)).astype(dt)
    n = (np.sin(dec0) * np.sin(dec) +
         np.cos(dec0) * np.cos(dec) * np.cos(d)).astype(dt)
    cols = [(RA, ra), (DEC, dec), (I_JY, rng.uniform(0.1, 9, num)),
            (SCRATCH_L, l), (SCRATCH_M, m), (SCRATCH_N, n)]
    if extra_index:
        cols.append((Q_JY, np.arange(num, dtype=np.float64)))
    sky_in = make_sky(P, num, cols)
    sky_in.set_reference(ra0, dec0)
    tel = Telescope(P, OSKAR_GPU, 4)
    s = C.c_int(0)
    lib.oskar_telescope_resize_station_array(tel.ptr, len(models),
                                             C.byref(s))
    check_status(s, "resize_station_array")
    for k, (lon, lat) in enumerate(models):
        st = lib.oskar_telescope_station(tel.ptr, k)
        lib.oskar_station_set_position(st, lon, lat, 0.0, 0.0, 0.0, 0.0)
    sky_out = Sky(P, OSKAR_GPU, 10)
    work = StationWork(P, OSKAR_GPU)
    lib.oskar_sky_horizon_clip(sky_out.ptr, sky_in.ptr, tel.ptr, gast,
                               work.ptr, C.byref(s))
    check_status(s, "oskar_sky_horizon_clip")
    sync_check("horizon clip")
    # Oracle: OR of every model's mask, prefix sum, row copy.
    p = "double" if prec == "d" else "single"
    want = np.zeros(num, np.int32)
    for lon, lat in models:
        want = oracle.update_horizon_mask(p, l, m, n, (gast + lon) - ra0,
                                          dec0, lat, mask=want)
    got_mask = _Borrowed(lib.oskar_station_work_horizon_mask(
        work.ptr)).numpy()[:num]
    got_idx = _Borrowed(lib.oskar_station_work_source_indices(
        work.ptr)).numpy()[:num + 1]
    return sky_in, sky_out, cols, want, got_mask, got_idx


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("nmodels", [3, 20, 0])
def test_horizon_clip_gpu_matches_oracle(hip, prec, nmodels):
    print("horizon clip", prec, nmodels, flush=True)
    rng = np.random.default_rng(nmodels)
    # Stations within ~30 deg of each other, so part of the sky is below
    # every horizon.
    models = [(rng.uniform(0.0, 0.5), rng.uniform(-0.6, -0.3))
              for _ in range(nmodels)]
    num = 50_001
    sky_in, sky_out, cols, want, got_mask, got_idx = clip_case(
        prec, num, models, 0.8, 10 + nmodels)
    assert np.array_equal(got_mask, want)
    assert np.array_equal(got_idx, oracle.prefix_sum(want))
    total = int(want.sum())
    assert sky_out.attr(Sky.NUM_SOURCES) == total
    if nmodels:
        assert 0 < total < num
    keep = want != 0
    dt = np.float64 if prec == "d" else np.float32
    for c, (t, values) in enumerate(cols):
        assert lib.oskar_sky_column_type(sky_out.ptr, c) == t
        assert np.array_equal(sky_out.get_column(t),
                              values.astype(dt)[keep])
    assert lib.oskar_sky_double(sky_out.ptr, Sky.REF_DEC_RAD) == -0.6


def test_horizon_clip_gpu_keeps_source_order_at_1m(hip):
    print("horizon clip 1M order", flush=True)
    num = 1_000_000
    sky_in, sky_out, cols, want, got_mask, got_idx = clip_case(
        "f", num, [(0.3, -0.5)], 4.4, 99, extra_index=True)
    assert np.array_equal(got_mask, want)
    total = int(want.sum())
    assert sky_out.attr(Sky.NUM_SOURCES) == total
    idx = sky_out.get_column(Q_JY)
    assert np.array_equal(idx, np.nonzero(want)[0].astype(np.float32))


def test_sky_copy_gpu_rejects_other_column_order(hip):
    P = OSKAR_SINGLE
    sky_in = make_sky(P, 10, [(RA, np.zeros(10)), (DEC, np.zeros(10))])
    sky_out = make_sky(P, 10, [(DEC, np.zeros(10))])
    mask = Mem.from_numpy(np.ones(10, np.int32), OSKAR_INT)
    idx = Mem.from_numpy(np.arange(11, dtype=np.int32), OSKAR_INT)
    s = C.c_int(0)
    lib.oskar_sky_copy_source_data(sky_in.ptr, mask.ptr, idx.ptr,
                                   sky_out.ptr, C.byref(s))
    assert s.value == -9
    sync_check("sky copy order")
