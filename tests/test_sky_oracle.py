"""Pins the oracle of the horizon clip (oracle_update_horizon_mask_*,
oracle_prefix_sum_int, oracle_sky_copy_source_data_*):

  - masks and row copies against golden vectors from the reference's own
    compiled CPU bodies (tests/golden/sky_clip.npz), bit for bit, including
    directions within ~1e-7 of the horizon plane where the rounding order
    decides;
  - the mask against the textbook elevation: a direction (l, m, n) about
    (RA0, Dec0) is converted back to (RA, Dec) and
    sin(el) = sin(lat) sin(Dec) + cos(lat) cos(Dec) cos(LST - RA) with
    LST = GAST + longitude; the mask is el > 0 (a wrong sign or a swapped
    factor in (ll, mm, nn) fails this);
  - the prefix sum against numpy's cumulative sum;
  - the row copy against numpy boolean indexing.
"""
import numpy as np
import pytest

import oracle
from helpers import load_golden


@pytest.mark.parametrize("prec", ["f", "d"])
def test_horizon_mask_oracle_matches_reference_kernels(prec):
    g = load_golden("sky_clip.npz")
    p = "double" if prec == "d" else "single"
    acc = np.zeros(g["l"].size, np.int32)
    for k, (ha0, dec0, lat) in enumerate(g["cases"]):
        one = oracle.update_horizon_mask(p, g["l"], g["m"], g["n"], ha0, dec0,
                                         lat)
        assert np.array_equal(one, g["mask_%s_%d" % (prec, k)])
        acc = oracle.update_horizon_mask(p, g["l"], g["m"], g["n"], ha0, dec0,
                                         lat, mask=acc)
    assert np.array_equal(acc, g["mask_%s_or" % prec])


def relative_directions(ra, dec, ra0, dec0):
    """Orthographic direction cosines about (ra0, dec0)."""
    d = ra - ra0
    l = np.cos(dec) * np.sin(d)
    m = np.cos(dec0) * np.sin(dec) - np.sin(dec0) * np.cos(dec) * np.cos(d)
    n = np.sin(dec0) * np.sin(dec) + np.cos(dec0) * np.cos(dec) * np.cos(d)
    return l, m, n


@pytest.mark.parametrize("case", [(0.4, -0.5, 1.1, -0.47, 2.0),
                                  (3.0, 0.9, -2.0, 0.7, 5.1),
                                  (5.5, -1.2, 0.3, -1.3, 0.2)])
def test_horizon_mask_oracle_is_textbook_elevation(case):
    ra0, dec0, lon, lat, gast = case
    rng = np.random.default_rng(1)
    ra = rng.uniform(0, 2 * np.pi, 20000)
    dec = np.arcsin(rng.uniform(-1, 1, 20000))
    l, m, n = relative_directions(ra, dec, ra0, dec0)
    lst = gast + lon
    sin_el = (np.sin(lat) * np.sin(dec) +
              np.cos(lat) * np.cos(dec) * np.cos(lst - ra))
    got = oracle.update_horizon_mask("double", l, m, n, (gast + lon) - ra0,
                                     dec0, lat)
    clear = np.abs(sin_el) > 1e-12
    assert np.array_equal(got[clear], (sin_el[clear] > 0).astype(np.int32))
    assert 0.3 < got.mean() < 0.7


def test_prefix_sum_oracle_is_cumulative_sum():
    rng = np.random.default_rng(2)
    for n in (0, 1, 5, 4097):
        v = rng.integers(-5, 6, n).astype(np.int32)
        got = oracle.prefix_sum(v)
        want = np.concatenate([[0], np.cumsum(v)]).astype(np.int32)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("prec", ["f", "d"])
def test_sky_copy_oracle_matches_reference_kernels(prec):
    g = load_golden("sky_clip.npz")
    p = "double" if prec == "d" else "single"
    ncol, cap_in, cap_out, nfix = (int(g[k]) for k in
                                   ("ncol", "cap_in", "cap_out", "nfix"))
    dt = np.float64 if prec == "d" else np.float32
    r, tout, vout = oracle.sky_copy_source_data(
        p, g["copy_mask"], g["table_in"], cap_in, cap_out, g["valid_in"],
        nfix, table_out=np.full(ncol * cap_out, -7.0, dt),
        valid_out=np.full(cap_out * nfix, -1, np.int32))
    assert r == int(g["copy_indices"][-1])
    assert np.array_equal(tout, g["table_out_%s" % prec])
    assert np.array_equal(vout, g["valid_out_%s" % prec])
    # Boolean indexing of each column gives the same rows.
    keep = g["copy_mask"] != 0
    tin = g["table_in"].astype(dt).reshape(ncol, cap_in)[:, :keep.size]
    assert np.array_equal(tout.reshape(ncol, cap_out)[:, :r], tin[:, keep])
