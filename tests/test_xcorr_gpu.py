"""GPU parity of oskar_cross_correlate (direct HIP kernel) against the CPU
oracle and the reference's golden vectors, through the C-ABI.

Tolerances (stated per precision):
  double: max |V - V_ref| / max |V_ref| < 1e-10 (the reference's own
          double-vs-double bound, Test_cross_correlate.cpp:18-38);
  single: relative RMS < 1e-5 and max rel < 1e-4 against the single-
          precision oracle on the same FP32 inputs.
"""
import numpy as np
import pytest

import oracle
from gpu_util import gpu_cross_correlate
from helpers import XCORR_CASES, load_golden, max_rel, rel_rms, xcorr_key

pytestmark = pytest.mark.gpu


def check(prec, got, ref):
    if prec == "d":
        assert max_rel(got, ref) < 1e-10
    else:
        assert rel_rms(got, ref) < 1e-5
        assert max_rel(got, ref) < 1e-4


@pytest.mark.parametrize("fixture", ["xcorr_ref.npz", "xcorr_long.npz"])
@pytest.mark.parametrize("case", XCORR_CASES,
                         ids=[xcorr_key(*c) for c in XCORR_CASES])
def test_xcorr_gpu_matches_golden(hip, fixture, case):
    prec, matrix, gauss, casa, bs, ts = case
    print("xcorr", fixture, xcorr_key(*case), flush=True)
    g = load_golden(fixture)
    got = gpu_cross_correlate(prec, casa, gauss, matrix, g, bs, ts)
    check(prec, got, g[xcorr_key(*case)])


def random_sky(rng, ns, nsrc, spread=30000.0):
    g = {}
    g["jones_m"] = rng.normal(size=ns * nsrc * 8)
    g["jones_s"] = rng.normal(size=ns * nsrc * 2)
    for k in ("u", "v", "x", "y"):
        g[k] = rng.uniform(-spread / 2, spread / 2, ns)
    g["w"] = rng.uniform(-50, 50, ns)
    g["I"] = rng.uniform(0.1, 3.0, nsrc)
    for k in ("Q", "U", "V"):
        g[k] = rng.uniform(-0.1, 0.1, nsrc)
    r = np.radians(4.0) * np.sqrt(rng.uniform(size=nsrc))
    t = rng.uniform(0, 2 * np.pi, nsrc)
    g["l"], g["m"] = r * np.cos(t), r * np.sin(t)
    g["n"] = np.sqrt(1 - g["l"] ** 2 - g["m"] ** 2)
    for k in ("a", "b", "c"):
        g[k] = rng.uniform(1e-9, 5e-8, nsrc)
    g["b"] = g["b"] * 0.1
    g.update(freq=np.array(140e6), channel_bw=np.array(50e3),
             time_avg=np.array(4.0), gast=np.array(0.9), ra0=np.array(0.3),
             dec0=np.array(-0.6))
    return g


def oracle_run(prec, casa, gauss, matrix, g, bs, ts, **kw):
    return oracle.cross_correlate(
        "double" if prec == "d" else "single", casa, gauss, matrix,
        g["jones_m"] if matrix else g["jones_s"], g["I"], g["Q"], g["U"],
        g["V"], g["l"], g["m"], g["n"], g["a"], g["b"], g["c"], g["u"],
        g["v"], g["w"], g["x"], g["y"],
        channel_bw=float(g["channel_bw"]) if bs else 0.0,
        time_avg=float(g["time_avg"]) if ts else 0.0, ra0=float(g["ra0"]),
        dec0=float(g["dec0"]), gast=float(g["gast"]), freq=float(g["freq"]),
        **kw)


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("matrix", [0, 1])
@pytest.mark.parametrize("shape", [(2, 1), (3, 5), (47, 1000), (70, 3001)])
def test_xcorr_gpu_ragged_shapes_with_split_sources(hip, prec, matrix, shape):
    """Station counts that do not fill a 32x16 block and source counts that
    are not a multiple of the source tile; (70, 3001) splits the sources over
    several workgroups, exercising the FP64 partial-sum reduction."""
    ns, nsrc = shape
    print("ragged", prec, matrix, shape, flush=True)
    g = random_sky(np.random.default_rng(ns * 7 + nsrc), ns, nsrc)
    for gauss, bs, ts in ((0, 0, 0), (1, 1, 1)):
        got = gpu_cross_correlate(prec, 0, gauss, matrix, g, bs, ts)
        want = oracle_run(prec, 0, gauss, matrix, g, bs, ts)
        check(prec, got, want)


def test_xcorr_gpu_accumulates_at_offset_and_respects_uv_filter(hip):
    """vis[offset_out + b] += V_b; filtered baselines keep their contents
    (oskar_cross_correlate_omp.cpp:92, oskar_interferometer_run_block.c:332)."""
    g = load_golden("xcorr_long.npz")
    ns = g["u"].size
    nb = ns * (ns - 1) // 2
    rng = np.random.default_rng(5)
    init = rng.normal(size=(3 * nb) * 8)
    got = gpu_cross_correlate("d", 0, 1, 1, g, 1, 1, vis_init=init,
                              offset_out=nb, uv_filter=(4000.0, 20000.0))
    want = oracle_run("d", 0, 1, 1, g, 1, 1, vis=init.copy(), offset_out=nb,
                      uv_min=4000.0, uv_max=20000.0)
    assert np.array_equal(got[:8 * nb], init[:8 * nb])
    assert np.array_equal(got[16 * nb:], init[16 * nb:])
    assert max_rel(got[8 * nb:16 * nb] - init[8 * nb:16 * nb],
                   want[8 * nb:16 * nb] - init[8 * nb:16 * nb]) < 1e-10
    # The filter leaves some baselines untouched and keeps others.
    u, v = g["u"], g["v"]
    lens = [np.hypot(u[p] - u[q], v[p] - v[q]) for q in range(ns)
            for p in range(q + 1, ns)]
    inside = [(4000.0 <= L <= 20000.0) for L in lens]
    assert any(inside) and not all(inside)
    for i, keep in enumerate(inside):
        row = got[8 * (nb + i):8 * (nb + i + 1)]
        if not keep:
            assert np.array_equal(row, init[8 * (nb + i):8 * (nb + i + 1)])


def test_xcorr_gpu_zero_sources_is_noop(hip):
    g = load_golden("xcorr_ref.npz")
    init = np.arange(171 * 8, dtype=np.float64)
    got = gpu_cross_correlate("d", 0, 0, 1, g, 0, 0, vis_init=init,
                              num_sources=0)
    assert np.array_equal(got, init)


def test_xcorr_gpu_deterministic(hip):
    g = random_sky(np.random.default_rng(77), 64, 4000)
    a = gpu_cross_correlate("f", 0, 1, 1, g, 1, 1)
    b = gpu_cross_correlate("f", 0, 1, 1, g, 1, 1)
    assert np.array_equal(a, b)
