"""The batched persistent launch at its own workload: 512 stations, 8
(channel, time) units of an observation of 384 channels (50-350 MHz) x 10
time samples, with 200k, 467k, 733k and 1M sources above the horizon at the
four time samples (two channels each), every unit with its own exact FP32
station beam (the shared 256-dipole design) and fluxes; tile shape, part
size and grid from the autotuner; one hipGraph replay.  Every unit is
checked on all 130,816 baselines x 4 polarisations against FP64
visibilities of the same unit (fp64_reference: K, J = K E and the direct
FP64 correlator from the same FP32-rounded inputs); tolerance: relative
RMS <= 1e-5 per unit.  Work units as oskar_interferometer_run_block.c:
88-145 (one (time, channel) block per unit).
"""
import ctypes as C
import math

import numpy as np
import pytest

from fp64_reference import direct_fp64_vis, f32, rel_rms
from gpu_util import sync_check
from oskar_amd import lib
from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_GPU, OSKAR_MATRIX,
                           OSKAR_SINGLE, FactorisedBatch, Mem, Station,
                           StationWork, Telescope, check_status)
from workloads import elements, power_law, sky, stations

pytestmark = pytest.mark.gpu
N = 512
CM32 = OSKAR_SINGLE | OSKAR_COMPLEX | OSKAR_MATRIX


def test_config4_batch_all_baselines(hip):
    rng = np.random.default_rng(404)
    st = stations(rng, N)
    station = Station(OSKAR_SINGLE, OSKAR_GPU, 256)
    station.set_coords(elements(rng, 256))
    lib.oskar_station_set_phase_centre(station.ptr, 4, 0.0,
                                       math.radians(60.0))
    work = StationWork(OSKAR_SINGLE, OSKAR_GPU)
    freqs = 50e6 + 300e6 * np.arange(384) / 383
    NB = N * (N - 1) // 2
    units, keep = [], []
    counts = [200_000, 466_667, 733_333, 1_000_000]
    vis = Mem(CM32, OSKAR_GPU, 8 * NB)
    b = FactorisedBatch(OSKAR_SINGLE, 0, N, 8)
    s = C.c_int(0)
    for t, S in enumerate(counts):
        sk = sky(rng, S)
        ang = 2 * np.pi * 60.0 * (3 * t) / 86164.0905
        ca, sa = math.cos(ang), math.sin(ang)
        uvw = dict(u=ca * st["u"] - sa * st["v"],
                   v=sa * st["u"] + ca * st["v"], w=st["w"])
        dirs = [Mem.from_numpy(sk[k], OSKAR_SINGLE) for k in ("l", "m", "n")]
        uvw_m = [Mem.from_numpy(uvw[k], OSKAR_SINGLE) for k in "uvw"]
        zero = Mem.from_numpy(np.zeros(S), OSKAR_SINGLE)
        for c in (int(rng.integers(0, 192)), int(rng.integers(192, 384))):
            k = len(units)
            beam = Mem(CM32, OSKAR_GPU, S)
            lib.oskar_evaluate_station_beam_aperture_array(
                station.ptr, work.ptr, S, dirs[0].ptr, dirs[1].ptr,
                dirs[2].ptr, 0, 0.0, float(freqs[c]), beam.ptr, C.byref(s))
            check_status(s, "beam")
            flux = Mem.from_numpy(sk["I"] * (freqs[c] / 150e6) ** -0.7,
                                  OSKAR_SINGLE)
            b.set_unit(k, S, beam, [flux, zero, zero, zero], dirs, uvw_m,
                       float(freqs[c]), k * NB, vis)
            units.append((sk, uvw, beam, flux, float(freqs[c]), S))
            keep += [dirs, uvw_m, zero]
    sync_check("config 4 beams")
    tel = Telescope(OSKAR_SINGLE, OSKAR_GPU, N)
    cfg = b.autotune(8)
    print("config 4: autotuned %s" % cfg, flush=True)
    b.run(8, 0, tel)          # captures the graph
    vis.clear()
    sync_check("clear")
    b.run(8, 0, tel)          # a replay
    sync_check("config 4 batch")
    got = vis.numpy().reshape(8, NB * 8)
    errs = []
    for k, (sk, uvw, beam, flux, freq, S) in enumerate(units):
        ref = direct_fp64_vis(
            *(f32(uvw[x]) for x in "uvw"), *(f32(sk[x]) for x in "lmn"),
            flux.numpy().astype(np.float64), freq,
            beam=beam.numpy().astype(np.float64).reshape(-1, 8))
        errs.append(rel_rms(got[k], ref))
        print("  unit %d: %d sources at %.1f MHz, rel RMS vs FP64 over all "
              "baselines %.3e" % (k, S, freq / 1e6, errs[-1]), flush=True)
    assert max(errs) <= 1e-5
