"""The headline workload end to end at full size, checked on every
baseline:

  512 aperture-array stations (log-radial layout to 35 km), each with the
  one 256-dipole station design (elements uniform in a 38 m disc), beam
  pointed 30 deg from zenith; 1,000,000 point sources uniform over the sky
  above the horizon, Stokes I from a power law; 150 MHz, no smearing.

  exact FP32 station beam (oskar_evaluate_station_beam_aperture_array)
  -> oskar_cross_correlate_factorised (FP16x3 and BF16x3 MFMA), FP32 vis

against FP64 visibilities of the same step on all 130,816 baselines x 4
polarisations (fp64_reference: K, J = K E, the direct FP64 correlator),
from the same FP32-rounded inputs.  The FP64 reference is itself checked
against the long-double oracle of the reference's algorithm on sampled
baselines.  Tolerance: relative RMS <= 1e-5.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from fp64_reference import direct_fp64_vis, f32, rel_rms, sample_baselines
from gpu_util import sync_check
from oskar_amd import capi, lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_GPU, OSKAR_MATRIX,
                           OSKAR_SINGLE, Mem, Station, StationWork,
                           Telescope, check_status)
from workloads import make_workload

pytestmark = pytest.mark.gpu
C0 = 299792458.0
N, S, NEL, FREQ = 512, 1_000_000, 256, 150e6
CM32 = OSKAR_SINGLE | OSKAR_COMPLEX | OSKAR_MATRIX


@pytest.fixture(scope="module")
def config3(hip):
    print("config 3: building workload and exact beam", flush=True)
    rng = np.random.default_rng(2024)
    st, sky, el = make_workload(rng, N, S, NEL)
    station = Station(OSKAR_SINGLE, OSKAR_GPU, NEL)
    station.set_coords(el, el)
    lib.oskar_station_set_phase_centre(station.ptr, 4, 0.0,
                                       math.radians(60.0))
    work = StationWork(OSKAR_SINGLE, OSKAR_GPU)
    dirs = [Mem.from_numpy(sky[k], OSKAR_SINGLE) for k in ("l", "m", "n")]
    flux = [Mem.from_numpy(sky["I"], OSKAR_SINGLE)] + [
        Mem.from_numpy(np.zeros(S), OSKAR_SINGLE) for _ in range(3)]
    uvw = [Mem.from_numpy(st[k], OSKAR_SINGLE) for k in ("u", "v", "w")]
    beam = Mem(CM32, OSKAR_GPU, S)
    s = C.c_int(0)
    lib.oskar_evaluate_station_beam_aperture_array(
        station.ptr, work.ptr, S, dirs[0].ptr, dirs[1].ptr, dirs[2].ptr, 0,
        0.0, FREQ, beam.ptr, C.byref(s))
    check_status(s, "station beam")
    sync_check("station beam")
    beam_host = beam.numpy().astype(np.float64).reshape(S, 8)
    print("config 3: FP64 reference on all baselines", flush=True)
    ref = direct_fp64_vis(*(f32(st[k]) for k in ("u", "v", "w")),
                          *(f32(sky[k]) for k in ("l", "m", "n")),
                          f32(sky["I"]), FREQ, beam=beam_host)
    return dict(st=st, sky=sky, beam=beam, beam_host=beam_host, dirs=dirs,
                flux=flux, uvw=uvw, ref=ref)


def test_fp64_reference_matches_long_double_oracle(config3):
    c = config3
    p, q, bl = sample_baselines(np.random.default_rng(5), N, 6)
    want = oracle.factorised_subset(
        f32(c["sky"]["l"]), f32(c["sky"]["m"]), f32(c["sky"]["n"]),
        f32(c["sky"]["I"]), c["beam_host"].ravel(), f32(c["st"]["u"]),
        f32(c["st"]["v"]), f32(c["st"]["w"]), 2 * np.pi * FREQ / C0, p, q)
    got = c["ref"].reshape(-1, 8)[bl].ravel()
    err = rel_rms(got, want)
    print("FP64 direct kernel vs long-double oracle, 6 baselines x 1M "
          "sources: rel RMS %.3e" % err, flush=True)
    assert err < 1e-10


@pytest.mark.parametrize("mode", [0, 1], ids=["fp16x3", "bf16x3"])
def test_factorised_all_baselines_vs_fp64(config3, mode):
    c = config3
    tel = Telescope(OSKAR_SINGLE, OSKAR_GPU, N)
    vis = Mem(CM32, OSKAR_GPU, N * (N - 1) // 2)
    vis.clear()
    s = C.c_int(0)
    lib.oskar_cross_correlate_factorised(
        0, S, c["beam"].ptr, capi.MEMA(*[m.ptr for m in c["flux"]]),
        capi.MEM3(*[m.ptr for m in c["dirs"]]), tel.ptr,
        capi.MEM3(*[m.ptr for m in c["uvw"]]), FREQ, mode, 0, vis.ptr,
        C.byref(s))
    check_status(s, "oskar_cross_correlate_factorised")
    sync_check("oskar_cross_correlate_factorised")
    got = vis.numpy()
    ref = c["ref"]
    err = rel_rms(got, ref)
    mx = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    print("config 3 %s: rel RMS vs FP64 over all 130,816 baselines x 4 "
          "pols %.3e, max |dV| / max |V64| %.3e" % (
              ("fp16x3", "bf16x3")[mode], err, mx), flush=True)
    assert err <= 1e-5
