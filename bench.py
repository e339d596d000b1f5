"""Benchmark of the visibility hot path on one MI355X.

Workload (one step = one time sample x one channel):
  512 stations (log-radial layout to 35 km), 130,816 baselines;
  1,000,000 point sources uniform over the sky above the horizon, Stokes I
  from a power law (index -1.8, 0.01-10 Jy);
  one 256-dipole aperture-array station design (random positions in a 38 m
  disc) shared by all stations, beam pointed 30 deg from zenith;
  150 MHz, no smearing, complex64 (FP32) visibilities.

Measured:
  value    source*baseline*channel evaluations per second of the factorised
           pipeline: exact station beam (oskar_evaluate_station_beam_aperture_array,
           FP32) + oskar_cross_correlate_factorised (split-FP16 MFMA);
  plain    the same step through the plain FP32 pipeline of the reference:
           station beam -> oskar_evaluate_jones_K -> oskar_jones_join ->
           oskar_cross_correlate (direct FP32 kernel);
  error    relative RMS of both against an FP64 / long-double CPU oracle on
           a random sample of baselines at the full source count;
  roofline the MFMA product kernel, timed with HIP events on its stream;
  clip     oskar_sky_horizon_clip of a 1M-source all-sky catalogue (18
           columns) at 10 sidereal times, against the 8 TB/s HBM peak;
  auto     oskar_auto_correlate on the plain pipeline's joined Jones
           (512 x 1M FP32 2x2 matrices, 16.4 GB), an HBM-bound sum,
           against the 8 TB/s HBM peak;
  cpu      the reference's own CPU correlator (compiled from its sources into
           oracle/_ref) on a bounded sample of the workload.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
# CPU baseline threads pinned one per core, set before any OpenMP runtime
# is loaded.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

from workloads import make_workload, power_law  # noqa: E402

C0 = 299792458.0
METRIC = ("source·baseline·channel evals/s at 512 stations vs plain "
          "FP32; rel RMS err vs FP64")
PEAK_MFMA_TFLOPS = 2500.0   # MI355X dense FP16/BF16 MFMA
PEAK_HBM_GBPS = 8000.0      # MI355X HBM3E


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_info():
    """Host CPU model, sockets and NUMA nodes (lscpu), for the CPU
    baseline record."""
    import subprocess
    out = {"cpu_model": None, "sockets": None, "numa_nodes": None,
           "host_cpus": os.cpu_count()}
    try:
        text = subprocess.run(["lscpu"], capture_output=True, text=True,
                              timeout=10).stdout
    except Exception:
        return out
    for line in text.splitlines():
        k, _, v = line.partition(":")
        k, v = k.strip(), v.strip()
        if k == "Model name":
            out["cpu_model"] = v
        elif k == "Socket(s)":
            out["sockets"] = int(v) if v.isdigit() else v
        elif k == "NUMA node(s)":
            out["numa_nodes"] = int(v) if v.isdigit() else v
    return out


def measured_traffic(num_stations, num_sources):
    """HBM bytes per launch of the MFMA product kernel at this workload,
    from the committed PMC summary (FETCH_SIZE + WRITE_SIZE of the kernel,
    each scaled by the same counter's reading for a copy of known size);
    None if there is no summary for this size."""
    path = os.path.join(ROOT, "profiles",
                        "fact_xcorr_kernel_512st_1Msrc_traffic.md")
    if num_stations != 512 or num_sources != 1_000_000 or \
            not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def copy_calibration():
    """Five elementwise copies of 1 GiB (y = x + 0, one kernel reading 1 GiB
    and writing 1 GiB each), for scaling the FETCH_SIZE / WRITE_SIZE
    counters to bytes."""
    import torch
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        torch.add(x, 0.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n,
                      "copies": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--stations", type=int, default=512)
    ap.add_argument("--sources", type=int, default=1_000_000)
    ap.add_argument("--elements", type=int, default=256)
    ap.add_argument("--freq", type=float, default=150e6)
    ap.add_argument("--mode", choices=["fp16x3", "bf16x3"], default="fp16x3")
    ap.add_argument("--tile-shape", type=int, default=-1,
                    help="tile shape of the one-call factorised path "
                         "(0-4; default: the library's)")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--clip-steps", type=int, default=10,
                    help="horizon clips of a 1M-source catalogue (0 = skip)")
    ap.add_argument("--dftw-steps", type=int, default=5,
                    help="oskar_dftw calls at 4096 elements x 1M directions "
                         "(0 = skip)")
    ap.add_argument("--tile-steps", type=int, default=5,
                    help="beams of a 16-tile x 16-dipole hierarchical "
                         "station at 1M directions (0 = skip)")
    ap.add_argument("--direct-channels", type=int, default=16,
                    help="channels of the 512 x 100k Gaussian + smearing "
                         "direct-correlator leg (0 = skip)")
    ap.add_argument("--sky-steps", type=int, default=5,
                    help="relative directions and Gaussian parameters of a "
                         "1M-source sky (0 = skip)")
    ap.add_argument("--auto-steps", type=int, default=5,
                    help="oskar_auto_correlate steps on the plain "
                         "pipeline's Jones (0 = skip)")
    ap.add_argument("--jones-steps", type=int, default=3,
                    help="oskar_evaluate_jones_K and oskar_jones_join steps "
                         "at 512 x 1M (0 = skip; needs the plain leg)")
    ap.add_argument("--vis-blocks", type=int, default=3,
                    help="512-station vis blocks written and read back "
                         "(0 = skip)")
    ap.add_argument("--station-steps", type=int, default=1,
                    help="distinct-layout leg: 512 station beams + the "
                         "per-station factorised product (0 = skip)")
    ap.add_argument("--station-part", type=int, default=0,
                    help="sources per part of the per-station factorised "
                         "product (0 = the library's default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--batch-units", type=int, default=16,
                    help="(channel, time) units of the batched run; 0 = off")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-batch-tune", action="store_true")
    ap.add_argument("--surrogate-steps", type=int, default=20000,
                    help="training steps of the beam surrogate; 0 = off")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the copy used to scale PMC byte counters")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X "
              "(no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from oskar_amd import capi, lib
    from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_GPU, OSKAR_MATRIX,
                               OSKAR_SINGLE, Jones, Mem, Station,
                               StationWork, Telescope, check_status)
    assert lib.oskar_device_count_hip() > 0, "no HIP device visible"

    N, S, f = args.stations, args.sources, args.freq
    NB = N * (N - 1) // 2
    evals = float(NB) * S
    rng = np.random.default_rng(2024)
    st, sky, el = make_workload(rng, N, S, args.elements)
    log("workload: %d stations, %d sources, %d elements" % (N, S,
                                                           args.elements))
    F32 = OSKAR_SINGLE
    CM = F32 | OSKAR_COMPLEX | OSKAR_MATRIX
    mode = 0 if args.mode == "fp16x3" else 1

    # Station model (shared design), beam 30 deg from zenith.
    station = Station(F32, OSKAR_GPU, args.elements)
    station.set_coords(el, el)
    lib.oskar_station_set_phase_centre(station.ptr, 4, np.radians(0.0),
                                       np.radians(60.0))
    work = StationWork(F32, OSKAR_GPU)
    dirs = [Mem.from_numpy(sky[k], F32) for k in ("l", "m", "n")]
    flux = [Mem.from_numpy(sky["I"], F32)] + [
        Mem.from_numpy(np.zeros(S), F32) for _ in range(3)]
    uvw = [Mem.from_numpy(st[k], F32) for k in ("u", "v", "w")]
    tel = Telescope(F32, OSKAR_GPU, N)
    beam = Mem(CM, OSKAR_GPU, S)
    vis = Mem(CM, OSKAR_GPU, NB)
    status = C.c_int(0)

    def step_beam():
        lib.oskar_evaluate_station_beam_aperture_array(
            station.ptr, work.ptr, S, dirs[0].ptr, dirs[1].ptr, dirs[2].ptr,
            0, 0.0, f, beam.ptr, C.byref(status))

    def step_fact():
        step_beam()
        lib.oskar_cross_correlate_factorised(
            0, S, beam.ptr, capi.MEMA(*[m.ptr for m in flux]),
            capi.MEM3(*[m.ptr for m in dirs]), tel.ptr,
            capi.MEM3(*[m.ptr for m in uvw]), f, mode, 0, vis.ptr,
            C.byref(status))

    timer = lib.oskar_timer_create(1)

    def timed(fn, steps, warmup, label):
        for i in range(warmup):
            fn()
        lib.oskar_device_synchronize(OSKAR_GPU)
        check_status(status, label)
        log("%s: warmup done" % label)
        lib.oskar_timer_start(timer)
        for i in range(steps):
            fn()
        lib.oskar_device_synchronize(OSKAR_GPU)
        t = lib.oskar_timer_elapsed(timer)
        s = C.c_int(0)
        lib.oskar_device_check_error_hip(C.byref(s))
        check_status(status, label)
        check_status(s, label)
        log("%s: %d steps in %.3f s" % (label, steps, t))
        return t

    # --- factorised pipeline -------------------------------------------
    if args.tile_shape >= 0 and \
            lib.oskar_cross_correlate_factorised_set_tile_shape(
                args.tile_shape) < 0:
        raise SystemExit("bad --tile-shape %d" % args.tile_shape)
    vis.clear()
    t_fact = timed(step_fact, args.steps, args.warmup, "factorised")
    fact_rate = evals * args.steps / t_fact
    ms_per_step = 1e3 * t_fact / args.steps

    # Dominant kernel: MFMA product, HIP events on the launch stream.
    lib.oskar_cross_correlate_factorised_set_profiling(1)
    stage = np.zeros(3)
    nprof = max(3, min(args.steps, 10))
    for i in range(nprof):
        step_fact()
        ms = (C.c_double * 3)()
        assert lib.oskar_cross_correlate_factorised_last_times(ms) == 0
        stage += np.array(ms[:])
    lib.oskar_cross_correlate_factorised_set_profiling(0)
    stage /= nprof
    # Beam time alone.
    t_beam = timed(step_beam, args.steps, 1, "beam")
    log("stage ms: prepare %.3f product %.3f reduce %.3f beam %.3f" % (
        stage[0], stage[1], stage[2], 1e3 * t_beam / args.steps))
    shape = lib.oskar_cross_correlate_factorised_set_tile_shape(-1)
    tp, tq = {0: (128, 64), 1: (64, 64), 2: (128, 32), 3: (64, 64),
              4: (128, 64)}[shape]
    num_tiles = sum(min((N + tq - 1) // tq,
                        (min(tp * i + tp - 1, N - 1) - 1) // tq + 1)
                    for i in range((N + tp - 1) // tp))
    part = 8192
    s_pad = ((S + part - 1) // part) * part
    # Executed MFMA work: TP x 4 TQ outputs, Re and Im, 2 reals per source,
    # 2 flops per multiply-add, 3 passes (hi*hi, hi*lo, lo*hi).
    exec_flops = num_tiles * float(s_pad) * tp * 4 * tq * 2 * 2 * 2 * 3
    t_prod = stage[1] * 1e-3
    achieved = 64.0 * evals / t_prod / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 2),
                "peak": PEAK_MFMA_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_MFMA_TFLOPS, 4),
                "traffic": measured_traffic(N, S),
                "kernel": "fact_xcorr_kernel (tile %d x %d, %s)" % (
                    tp, tq, "f16" if mode == 0 else "bf16"),
                "tile_shape": shape,
                "kernel_ms": round(stage[1], 4),
                "algorithmic_flops_per_eval": 64,
                "mfma_executed_tflops": round(exec_flops / t_prod / 1e12, 2),
                "mfma_executed_frac": round(exec_flops / t_prod / 1e12 /
                                            PEAK_MFMA_TFLOPS, 4)}

    # Keep one factorised result for the error check.
    vis.clear()
    step_fact()
    lib.oskar_device_synchronize(OSKAR_GPU)
    vis_fact = vis.numpy()
    beam_host = beam.numpy().astype(np.float64)

    # --- plain FP32 pipeline ------------------------------------------
    plain_rate = None
    vis_plain = None
    auto = None
    jones = None
    if not args.no_plain:
        E = Jones(CM, OSKAR_GPU, N, S)
        K = Jones(F32 | OSKAR_COMPLEX, OSKAR_GPU, N, S)
        ext = [Mem(F32, OSKAR_GPU, 1) for _ in range(3)]
        tel_p = Telescope(F32, OSKAR_GPU, N)
        Emem = lib.oskar_jones_mem(E.ptr)

        def step_plain():
            step_beam()
            for a in range(N):      # identical stations: duplicate the beam
                lib.oskar_mem_copy_contents(Emem, beam.ptr, a * S, 0, S,
                                            C.byref(status))
            lib.oskar_evaluate_jones_K(K.ptr, S, dirs[0].ptr, dirs[1].ptr,
                                       dirs[2].ptr, uvw[0].ptr, uvw[1].ptr,
                                       uvw[2].ptr, f, flux[0].ptr, -1e30,
                                       1e30, 0, C.byref(status))
            lib.oskar_jones_join(E.ptr, K.ptr, E.ptr, C.byref(status))
            lib.oskar_cross_correlate(
                0, 0, S, E.ptr, capi.MEMA(*[m.ptr for m in flux]),
                capi.MEM3(*[m.ptr for m in dirs]),
                capi.MEM3(*[m.ptr for m in ext]), tel_p.ptr,
                capi.MEM3(*[m.ptr for m in uvw]), 0.0, f, 0, vis.ptr,
                C.byref(status))

        vis.clear()
        t_plain = timed(step_plain, args.plain_steps, 1, "plain")
        plain_rate = evals * args.plain_steps / t_plain
        vis.clear()
        step_plain()
        lib.oskar_device_synchronize(OSKAR_GPU)
        vis_plain = vis.numpy()

        # Auto-correlations of the joined Jones: every 2x2 FP32 value is
        # read once (32 B per station and source), so the sum is HBM-bound.
        if args.auto_steps > 0:
            vis_auto = Mem(CM, OSKAR_GPU, N)

            def step_auto():
                lib.oskar_auto_correlate(
                    S, E.ptr, capi.MEMA(*[m.ptr for m in flux]), 0,
                    vis_auto.ptr, C.byref(status))
            t_auto = timed(step_auto, args.auto_steps, 1, "auto")
            ms_auto = 1e3 * t_auto / args.auto_steps
            auto_bytes = 32 * N * S + 4 * 4 * S + 2 * 32 * N
            gbps = auto_bytes / (ms_auto * 1e-3) / 1e9
            auto = {"workload": "%d stations x %d sources, FP32 2x2 Jones"
                                % (N, S),
                    "ms": round(ms_auto, 4),
                    "algorithmic_bytes": auto_bytes,
                    "achieved_GBps": round(gbps, 1),
                    "hbm_peak_GBps": PEAK_HBM_GBPS,
                    "frac": round(gbps / PEAK_HBM_GBPS, 4)}
            log("auto-correlation: %.3f ms, %.0f GB/s" % (ms_auto, gbps))
            del vis_auto

        # Jones K and the join alone.  K: 8 B written per (station,
        # source) plus the source and station inputs once; join: a 2x2
        # FP32 matrix read and written and a complex scalar read (72 B).
        if args.jones_steps > 0:
            def step_K():
                lib.oskar_evaluate_jones_K(
                    K.ptr, S, dirs[0].ptr, dirs[1].ptr, dirs[2].ptr,
                    uvw[0].ptr, uvw[1].ptr, uvw[2].ptr, f, flux[0].ptr,
                    -1e30, 1e30, 0, C.byref(status))

            def step_join():
                lib.oskar_jones_join(E.ptr, K.ptr, E.ptr, C.byref(status))
            ms_k = 1e3 * timed(step_K, args.jones_steps, 1, "jones K") / \
                args.jones_steps
            ms_j = 1e3 * timed(step_join, args.jones_steps, 1, "join") / \
                args.jones_steps
            k_bytes = 8 * N * S + 16 * S + 12 * N
            j_bytes = 72 * N * S
            jones = {"workload": "%d stations x %d sources, FP32" % (N, S),
                     "K_ms": round(ms_k, 4),
                     "K_GBps": round(k_bytes / ms_k / 1e6, 1),
                     "K_frac": round(k_bytes / ms_k / 1e6 / PEAK_HBM_GBPS,
                                     4),
                     "join_ms": round(ms_j, 4),
                     "join_GBps": round(j_bytes / ms_j / 1e6, 1),
                     "join_frac": round(j_bytes / ms_j / 1e6 /
                                        PEAK_HBM_GBPS, 4)}
            log("jones K %.3f ms (%.0f GB/s), join %.3f ms (%.0f GB/s)" % (
                ms_k, k_bytes / ms_k / 1e6, ms_j, j_bytes / ms_j / 1e6))
        del E, K

    # --- error against FP64 on every baseline ----------------------------
    # FP64 visibilities of the same step (K, J = K E, the direct FP64
    # correlator) from the same FP32-rounded inputs; the FP64 kernel is
    # checked against the long-double CPU oracle on a few baselines.
    err_fact = err_plain = maxerr_fact = ref_check = None
    ref = None
    if not args.no_error:
        import oracle
        from fp64_reference import (direct_fp64_vis, f32, rel_rms,
                                    sample_baselines)
        t0 = time.time()
        ref = direct_fp64_vis(*(f32(st[k]) for k in ("u", "v", "w")),
                              *(f32(sky[k]) for k in ("l", "m", "n")),
                              f32(sky["I"]), f,
                              beam=beam_host.reshape(S, 8))
        log("FP64 reference on all %d baselines: %.1f s" % (
            NB, time.time() - t0))
        p, q, bl = sample_baselines(rng, N, 6)
        t0 = time.time()
        want = oracle.factorised_subset(
            f32(sky["l"]), f32(sky["m"]), f32(sky["n"]), f32(sky["I"]),
            beam_host, f32(st["u"]), f32(st["v"]), f32(st["w"]),
            2 * np.pi * f / C0, p, q)
        ref_check = rel_rms(ref.reshape(-1, 8)[bl].ravel(), want)
        log("FP64 reference vs long-double oracle on 6 baselines: %.3e "
            "(%.1f s)" % (ref_check, time.time() - t0))
        err_fact = rel_rms(vis_fact, ref)
        maxerr_fact = float(np.max(np.abs(vis_fact - ref)) /
                            np.max(np.abs(ref)))
        if vis_plain is not None:
            err_plain = rel_rms(vis_plain, ref)
        log("rel RMS vs FP64 (all baselines x 4 pols): factorised %.3e "
            "(max %.3e) plain %s" % (
                err_fact, maxerr_fact,
                "%.3e" % err_plain if err_plain is not None else "-"))

    # --- learned beam surrogate feeding the factorised correlator -------
    surrogate = None
    if args.surrogate_steps > 0:
        surrogate = bench_surrogate(args, st, sky, station, work, mode, tel,
                                    flux, dirs, uvw, beam, ref, timed,
                                    evals)

    # --- horizon clip of an all-sky catalogue --------------------------
    vis = None
    if args.vis_blocks > 0:
        vis = bench_vis(args, rng)
    clip = None
    if args.clip_steps > 0:
        clip = bench_clip(args, rng)
    sky_params = None
    if args.sky_steps > 0:
        sky_params = bench_sky_params(args, rng)
    direct = None
    if args.direct_channels > 0:
        direct = bench_direct_smearing(args, rng)
    dftw = None
    if args.dftw_steps > 0:
        dftw = bench_dftw(args, rng)
    tiles = None
    if args.tile_steps > 0:
        tiles = bench_tiled_station(args, rng)

    # --- batched (channel, time) units in one persistent launch ---------
    batch = None
    if args.batch_units > 0:
        batch = bench_batch(args, st, station, work, mode)

    # --- distinct station layouts: per-station factorised form ----------
    stations = None
    if args.station_steps > 0:
        stations = bench_stations(args, st, sky, dirs, flux, uvw, mode,
                                  timed, evals)

    # --- CPU baseline: the reference's own correlator ------------------
    cpu = None
    if not args.no_cpu:
        import oracle
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        os.environ.setdefault("OMP_NUM_THREADS", str(cores))
        ref_lib = oracle.ref()
        kind = "reference" if ref_lib is not None else "port"

        def cpu_run(ns):
            # J = K E for the first ns sources, FP64, on the host.
            k = 2 * np.pi * f / C0
            ph = k * (np.outer(st["u"], sky["l"][:ns]) +
                      np.outer(st["v"], sky["m"][:ns]) +
                      np.outer(st["w"], sky["n"][:ns] - 1.0))
            Kc = np.exp(1j * ph)
            Ec = (beam_host[:8 * ns:2] + 1j * beam_host[1:8 * ns:2]).reshape(
                ns, 4)
            J = Kc[:, :, None] * Ec[None]
            jr = np.ascontiguousarray(
                np.stack([J.real, J.imag], -1).reshape(-1))
            z = np.zeros(N)
            zs = np.zeros(ns)
            out = np.zeros(NB * 8)
            arrs = [np.ascontiguousarray(a, np.float64) for a in (
                sky["I"][:ns], zs, zs, zs, sky["l"][:ns], sky["m"][:ns],
                sky["n"][:ns], st["u"], st["v"], st["w"], z, z)]
            t0 = time.time()
            if ref_lib is not None:
                P = lambda a: a.ctypes.data_as(C.c_void_p)
                D = C.c_double
                ref_lib.oskar_cross_correlate_point_omp_d(
                    C.c_int(0), C.c_int(ns), C.c_int(N), C.c_int(0), P(jr),
                    *[P(a) for a in arrs], D(0.0), D(3.4e38), D(f / C0),
                    D(0.0), D(0.0), D(0.0), D(0.0), P(out))
            else:
                oracle.cross_correlate("double", 0, 0, 1, jr, *arrs[:7],
                                       None, None, None, *arrs[7:], freq=f)
            return time.time() - t0

        t1 = cpu_run(500)
        ns = int(min(max(500, 500 * 3.0 / max(t1, 1e-3)), 50000))
        cpu_run(ns)                               # warm-up
        times = [cpu_run(ns) for _ in range(5)]
        tc = float(np.median(times))
        host = cpu_info()
        cpu = {"value": NB * ns / tc, "unit": "evals/s", "cores": cores,
               "kind": kind,
               "sample": "%d stations x %d sources (of %d), FP64 matrix "
                         "point-source correlator, median of 5 after 1 "
                         "warm-up" % (N, ns, S),
               "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
               "omp_places": os.environ.get("OMP_PLACES"),
               **host}
        log("CPU baseline (%s): %.3e evals/s on %d cores" % (
            kind, cpu["value"], cores))

    line = {
        "metric": METRIC,
        "value": fact_rate,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16x3" if mode == 0 else "bf16x3",
        "data": "synthetic",
        "config": {
            "workload": "512 stations x 1M point sources, 1 channel, "
                        "shared 256-dipole station beam (exact FP32 DFT), "
                        "no smearing (factorised MFMA pipeline)",
            "stations": N, "baselines": NB, "sources": S,
            "elements": args.elements, "freq_hz": f, "mode": args.mode,
            "parallelism": "single GPU"},
        "plain_fp32_evals_per_s": plain_rate,
        "speedup_vs_plain_fp32": (fact_rate / plain_rate) if plain_rate
        else None,
        "rel_rms_vs_fp64": err_fact,
        "rel_rms_vs_fp64_scope": "all 130,816 baselines x 4 pols; FP64 "
                                 "direct kernel on the same FP32-rounded "
                                 "inputs",
        "max_abs_err_over_max_vs_fp64": maxerr_fact,
        "fp64_reference_vs_long_double_oracle": ref_check,
        "rel_rms_plain_fp32_vs_fp64": err_plain,
        "stage_ms": {"prepare": round(stage[0], 4),
                     "product": round(stage[1], 4),
                     "reduce": round(stage[2], 4),
                     "beam": round(1e3 * t_beam / args.steps, 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batch": batch,
        "distinct_stations": stations,
        "surrogate": surrogate,
        "auto_correlate": auto,
        "jones_K_join": jones,
        "horizon_clip": clip,
        "sky_parameters": sky_params,
        "direct_gaussian_smearing": direct,
        "dftw": dftw,
        "tiled_station": tiles,
        "vis_block": vis,
    }
    print(json.dumps(line), flush=True)


def bench_clip(args, rng):
    """oskar_sky_horizon_clip of a 1M-source all-sky catalogue (RA, Dec,
    I, Q, U, V, reference frequency, spectral index, rotation measure,
    major, minor, PA, l, m, n and the three Gaussian parameters: 18 FP32
    columns) about a phase centre at Dec -30 deg, for one station model at
    latitude -26.7 deg, at `clip_steps` sidereal times over a day.  Each
    clip ends with the 4-byte read of the kept count (the host needs it).
    Bytes per clip: l, m, n read and the mask written (16 B per source),
    mask read and index written (8 B), and for each kept source its row
    read and written (2 x 18 x 4 B) with its valid-column counts
    (2 x 80 B)."""
    from oskar_amd import lib
    from oskar_amd.mem import (OSKAR_GPU, OSKAR_SINGLE, Sky, StationWork,
                               Telescope, check_status)
    S = 1_000_000
    ra0, dec0 = 0.0, np.radians(-30.0)
    ra = rng.uniform(0, 2 * np.pi, S)
    dec = np.arcsin(rng.uniform(-1, 1, S))
    d = ra - ra0
    l = np.cos(dec) * np.sin(d)
    m = np.cos(dec0) * np.sin(dec) - np.sin(dec0) * np.cos(dec) * np.cos(d)
    n = np.sin(dec0) * np.sin(dec) + np.cos(dec0) * np.cos(dec) * np.cos(d)
    cols = [(1, ra), (2, dec), (3, rng.uniform(0.01, 10, S))]
    cols += [(t, rng.uniform(-0.1, 0.1, S)) for t in (4, 5, 6)]
    cols += [(7, np.full(S, 1e8)), (8, np.full(S, -0.7)), (9, np.zeros(S))]
    cols += [(t, np.zeros(S)) for t in (10, 11, 12)]
    cols += [(0x103, l), (0x104, m), (0x105, n)]
    cols += [(t, np.zeros(S)) for t in (0x100, 0x101, 0x102)]
    sky_in = Sky(OSKAR_SINGLE, OSKAR_GPU, S)
    for t, v in cols:
        sky_in.set_column(t, v)
    sky_in.set_reference(ra0, dec0)
    sky_out = Sky(OSKAR_SINGLE, OSKAR_GPU, S)
    tel = Telescope(OSKAR_SINGLE, OSKAR_GPU, 4)
    work = StationWork(OSKAR_SINGLE, OSKAR_GPU)
    status = C.c_int(0)
    lib.oskar_telescope_resize_station_array(tel.ptr, 1, C.byref(status))
    lib.oskar_station_set_position(lib.oskar_telescope_station(tel.ptr, 0),
                                   np.radians(116.7), np.radians(-26.7), 0.0,
                                   0.0, 0.0, 0.0)
    gast = np.linspace(0, 2 * np.pi, args.clip_steps, endpoint=False)

    def clip_once(g):
        lib.oskar_sky_horizon_clip(sky_out.ptr, sky_in.ptr, tel.ptr, float(g),
                                   work.ptr, C.byref(status))

    clip_once(0.0)                          # warm-up (sizes the outputs)
    check_status(status, "horizon clip")
    kept = []
    timer = lib.oskar_timer_create(1)
    lib.oskar_timer_start(timer)
    for g in gast:
        clip_once(g)
        kept.append(lib.oskar_sky_int(sky_out.ptr, Sky.NUM_SOURCES))
    t = lib.oskar_timer_elapsed(timer)
    lib.oskar_timer_free(timer)
    check_status(status, "horizon clip")
    ms = 1e3 * t / len(gast)
    ncol = len(cols)
    nbytes = [S * 24 + k * 2 * (ncol * 4 + 80) for k in kept]
    gbps = float(np.mean(nbytes)) / (ms * 1e-3) / 1e9
    log("horizon clip: %.3f ms per clip, kept %d-%d of %d, %.0f GB/s" % (
        ms, min(kept), max(kept), S, gbps))
    return {"workload": "1M-source all-sky catalogue, 18 FP32 columns, one "
                        "station model, %d sidereal times" % len(gast),
            "ms_per_clip": round(ms, 4),
            "kept_min": int(min(kept)), "kept_max": int(max(kept)),
            "algorithmic_bytes_mean": float(np.mean(nbytes)),
            "achieved_GBps": round(gbps, 1),
            "hbm_peak_GBps": PEAK_HBM_GBPS,
            "frac": round(gbps / PEAK_HBM_GBPS, 4)}


def bench_tiled_station(args, rng):
    """oskar_evaluate_station_beam_aperture_array on a hierarchical station:
    16 alike tiles (child stations) of 16 half-wave dipoles on a 1.25 m grid,
    tile centres within 18 m, FP32 2x2 beam at 1M directions over the upper
    hemisphere, 150 MHz, pointed at az 30 deg, el 60 deg.  The alike tiles'
    beam is evaluated once and read by every tile through an all-zero data
    index of the station's DFT.  Unit: one direction of the station beam.
    The oracle composition (tile beam, then the station's weighted DFT) runs
    a 20k-direction sample on the host cores and is the accuracy reference
    for the first 20k directions.  Roofline: the VALU-issue bound of the two
    weighted DFTs (16 + 16 element terms per direction at the dftw leg's 40
    cycles per wave64 term: 256 CUs x 4 SIMDs x 64 lanes x 2.4 GHz / 40 =
    3.93e12 terms/s)."""
    import oracle
    from oskar_amd import lib
    from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_GPU, OSKAR_MATRIX,
                               OSKAR_SINGLE, Mem, Station, StationView,
                               StationWork, check_status)
    nt, nd, f = 16, 1_000_000, 150e6
    g = (np.arange(4) - 1.5) * 1.25
    tx = np.stack([np.repeat(g, 4), np.tile(g, 4), np.zeros(16)])
    r = 18.0 * np.sqrt(rng.uniform(size=nt))
    th = rng.uniform(0, 2 * np.pi, nt)
    sx = np.stack([r * np.cos(th), r * np.sin(th), np.zeros(nt)])
    az, el = np.radians(30.0), np.radians(60.0)
    status = C.c_int(0)
    st = Station(OSKAR_SINGLE, OSKAR_GPU, nt)
    st.set_coords(sx)
    lib.oskar_station_create_child_stations(st.ptr, C.byref(status))
    for i in range(nt):
        ch = StationView(lib.oskar_station_child(st.ptr, i), OSKAR_SINGLE)
        lib.oskar_station_resize(ch.ptr, 16, C.byref(status))
        ch.set_coords(tx)
        lib.oskar_station_set_phase_centre(ch.ptr, 4, az, el)
    lib.oskar_station_set_phase_centre(st.ptr, 4, az, el)
    lib.oskar_station_analyse(st.ptr, C.byref(C.c_int(0)), C.byref(status))
    check_status(status, "tiled station set-up")
    nz = rng.uniform(0, 1, nd)
    ph = rng.uniform(0, 2 * np.pi, nd)
    x = np.sqrt(1 - nz * nz) * np.sin(ph)
    y = np.sqrt(1 - nz * nz) * np.cos(ph)
    mx, my, mz = (Mem.from_numpy(a, OSKAR_SINGLE) for a in (x, y, nz))
    out = Mem(OSKAR_SINGLE | OSKAR_COMPLEX | OSKAR_MATRIX, OSKAR_GPU, nd)
    work = StationWork(OSKAR_SINGLE, OSKAR_GPU)

    def one():
        lib.oskar_evaluate_station_beam_aperture_array(
            st.ptr, work.ptr, nd, mx.ptr, my.ptr, mz.ptr, 0, 0.0, f,
            out.ptr, C.byref(status))

    one()
    lib.oskar_device_synchronize(OSKAR_GPU)
    timer = lib.oskar_timer_create(1)
    lib.oskar_timer_start(timer)
    for _ in range(args.tile_steps):
        one()
    lib.oskar_device_synchronize(OSKAR_GPU)
    t = lib.oskar_timer_elapsed(timer) / args.tile_steps
    lib.oskar_timer_free(timer)
    check_status(status, "tiled station beam")
    res = {"workload": "16 alike tiles x 16 dipoles, 1M directions, FP32 "
                       "2x2 beam, 150 MHz",
           "identical_children": lib.oskar_station_identical_children(
               st.ptr),
           "ms": round(1e3 * t, 4), "directions_per_s": nd / t}
    log("tiled station 16x16 at 1M directions: %.3f ms, %.3e directions/s"
        % (1e3 * t, nd / t))
    ns = 20000
    k = 2 * np.pi * f / C0
    d = np.array([np.cos(el) * np.sin(az), np.cos(el) * np.cos(az),
                  np.sin(el)])
    f4 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    t0 = time.time()
    tile = oracle.station_beam_aa(
        1, *[f4(v) for v in tx], *[f4(v) for v in tx], np.zeros(16),
        np.zeros(16, np.int32), np.array([0], np.int32),
        np.array([0.5 * C0 / f]), d, f4(x[:ns]), f4(y[:ns]), f4(nz[:ns]), f)
    w = np.exp(-1j * k * (d @ f4(sx)))
    want = oracle.dftw("double", 1, 0, k,
                       np.stack([w.real, w.imag], 1).ravel(),
                       *[f4(v) for v in sx], f4(x[:ns]), f4(y[:ns]),
                       f4(nz[:ns]), None, np.tile(tile, nt))
    tc = time.time() - t0
    got = out.numpy()[: ns * 8].astype(np.float64)
    res["rel_rms_vs_oracle"] = float(
        np.sqrt(np.sum((got - want) ** 2) / np.sum(want * want)))
    terms = nd * (nt + 16)
    res["roofline"] = {"bound": "valu", "achieved": terms / t,
                       "peak": 3.93e12, "unit": "terms/s",
                       "frac": round(terms / t / 3.93e12, 4),
                       "traffic": None}
    log("tiled station vs oracle (20k directions): rel RMS %.3e"
        % res["rel_rms_vs_oracle"])
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    res["cpu_baseline"] = {"directions_per_s": ns / tc, "cores": cores,
                           "kind": "port",
                           "sample": "the same station, 20k directions"}
    log("tiled station CPU (oracle, %d cores): %.3e directions/s"
        % (cores, ns / tc))
    return res


def bench_dftw(args, rng):
    """oskar_dftw alone: 4096 elements (76 m disc, planar) x 1M directions
    over the upper hemisphere, FP32, one 2x2 element pattern per direction
    (one element type), 150 MHz.  Unit: one (element, direction) term.  Its
    phases stay within 54 revolutions, so the kernel's inner loop takes the
    path without range reduction: 4 VALU instructions (two fused
    multiply-adds for the phase with the weight's phase, two into the array
    factor with the weight's amplitude) and 2 transcendental ones (sine and
    cosine) per term; at 4 cycles per wave64 instruction that bounds it at
    256 CUs x 4 SIMDs x 64 lanes x 2.4 GHz / 24 = 6.55e12 terms/s, the
    roofline reported.  The oracle's OpenMP body (the reference's CPU
    arithmetic) runs a 20k-direction sample on the host cores."""
    import oracle
    from oskar_amd import lib
    from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_GPU, OSKAR_INT,
                               OSKAR_MATRIX, OSKAR_SINGLE, Mem, check_status)
    ne, nd = 4096, 1_000_000
    f = 150e6
    k = 2 * np.pi * f / C0
    r = 76.0 * np.sqrt(rng.uniform(size=ne))
    th = rng.uniform(0, 2 * np.pi, ne)
    x_in, y_in = r * np.cos(th), r * np.sin(th)
    w = np.exp(-1j * rng.uniform(0, 2 * np.pi, ne))
    wr = np.stack([w.real, w.imag], -1).ravel()
    nz = rng.uniform(0, 1, nd)
    ph = rng.uniform(0, 2 * np.pi, nd)
    xo = np.sqrt(1 - nz * nz) * np.sin(ph)
    yo = np.sqrt(1 - nz * nz) * np.cos(ph)
    data = rng.normal(size=nd * 8).astype(np.float32)
    CT = OSKAR_SINGLE | OSKAR_COMPLEX
    mw = Mem.from_numpy(wr, CT)
    mx, my = (Mem.from_numpy(a, OSKAR_SINGLE) for a in (x_in, y_in))
    ox, oy = (Mem.from_numpy(a, OSKAR_SINGLE) for a in (xo, yo))
    midx = Mem.from_numpy(np.zeros(ne, np.int32), OSKAR_INT)
    md = Mem.from_numpy(data, CT | OSKAR_MATRIX)
    out = Mem.from_numpy(np.zeros(nd * 8, np.float32), CT | OSKAR_MATRIX)
    status = C.c_int(0)

    def one():
        lib.oskar_dftw(1, ne, k, mw.ptr, mx.ptr, my.ptr, None, 0, nd, ox.ptr,
                       oy.ptr, None, midx.ptr, md.ptr, 1, 1, 0, out.ptr,
                       C.byref(status))

    one()
    lib.oskar_device_synchronize(OSKAR_GPU)
    timer = lib.oskar_timer_create(1)
    lib.oskar_timer_start(timer)
    for _ in range(args.dftw_steps):
        one()
    lib.oskar_device_synchronize(OSKAR_GPU)
    t = lib.oskar_timer_elapsed(timer) / args.dftw_steps
    lib.oskar_timer_free(timer)
    check_status(status, "oskar_dftw")
    rate = ne * nd / t
    bound = 256 * 4 * 64 * 2.4e9 / 24.0
    res = {"workload": "4096 elements x 1M directions, FP32, one element "
                       "pattern per direction, 150 MHz",
           "ms": round(1e3 * t, 4), "terms_per_s": rate,
           "roofline": {"bound": "valu issue (4 VALU + 2 transcendental "
                                 "instructions per term, 4 cycles each)",
                        "achieved": rate, "peak": bound,
                        "unit": "terms/s", "frac": round(rate / bound, 3)}}
    log("dftw 4096 x 1M: %.3f ms, %.3e terms/s (%.2f of the issue bound)"
        % (1e3 * t, rate, rate / bound))
    ns = 20000
    t0 = time.time()
    oracle.dftw("single", 1, 1, k, wr, x_in, y_in, None, xo[:ns], yo[:ns],
                None, np.zeros(ne, np.int32), data[:ns * 8])
    tc = time.time() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    res["cpu_baseline"] = {"terms_per_s": ne * ns / tc, "cores": cores,
                           "kind": "port",
                           "sample": "4096 elements x 20k directions"}
    log("dftw CPU (oracle, %d cores): %.3e terms/s" % (cores,
                                                        ne * ns / tc))
    return res


def bench_direct_smearing(args, rng):
    """oskar_cross_correlate (the direct kernel) on 512 stations x 100,000
    sources, 30 % Gaussian (FWHM 5-60 arcsec, a, b, c from the library's
    oskar_sky_evaluate_gaussian_source_parameters), full Stokes, bandwidth
    smearing (10 kHz) and time smearing (1 s), RADEC phase centre, in FP32
    and FP64, one call per channel over `direct_channels` channels in
    50-350 MHz (log-radial layout to 35 km; Jones = CN(0, 1) + I per
    station and source).  Work per evaluation: 133 flops + 2 sines + 2
    divisions + 1 exponential.  The reference's own
    oskar_cross_correlate_gaussian_omp_d (oracle/_ref) runs the same
    configuration on the host for a sample of the sources."""
    import oracle
    from oskar_amd import capi, lib
    from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_DOUBLE, OSKAR_GPU,
                               OSKAR_MATRIX, OSKAR_SINGLE, Jones, Mem, Sky,
                               Telescope, check_status)
    N, S = 512, 100_000
    NB = N * (N - 1) // 2
    ra0, dec0, gast = 0.4, -0.5, 1.3
    bw, tavg = 10e3, 1.0
    freqs = np.linspace(50e6, 350e6, args.direct_channels)
    r = np.exp(rng.uniform(np.log(100.0), np.log(35000.0), N))
    th = rng.uniform(0, 2 * np.pi, N)
    uvw = [r * np.cos(th), r * np.sin(th), rng.uniform(-50, 50, N)]
    rho = np.radians(5.0) * np.sqrt(rng.uniform(0, 1, S))
    phi = rng.uniform(0, 2 * np.pi, S)
    l, m = np.sin(rho) * np.cos(phi), np.sin(rho) * np.sin(phi)
    nn = np.sqrt(1 - l * l - m * m)
    dec = np.arcsin(nn * np.sin(dec0) + m * np.cos(dec0))
    ra = ra0 + np.arctan2(l, nn * np.cos(dec0) - m * np.sin(dec0))
    I_ = power_law(rng, S)
    pol = rng.uniform(0, 0.2, S) * I_
    d = rng.normal(size=(3, S))
    d /= np.linalg.norm(d, axis=0)
    gauss = rng.uniform(0, 1, S) < 0.3
    maj = np.where(gauss, np.radians(rng.uniform(5, 60, S) / 3600), 0.0)
    mn = maj * rng.uniform(0.2, 1.0, S)
    pa = rng.uniform(0, np.pi, S)
    cols = {1: ra, 2: dec, 3: I_, 4: pol * d[0], 5: pol * d[1],
            6: pol * d[2], 10: maj, 11: mn, 12: pa}
    J = rng.standard_normal(N * S * 8, dtype=np.float32)
    J *= np.float32(1 / np.sqrt(2))
    J = J.reshape(N, S, 8)
    J[:, :, 0] += 1
    J[:, :, 6] += 1
    status = C.c_int(0)
    timer = lib.oskar_timer_create(1)
    out = {"workload": "512 stations x 100k sources (30 %% Gaussian, full "
                       "Stokes), bandwidth (10 kHz) and time (1 s) "
                       "smearing, %d channels in 50-350 MHz" % len(freqs)}
    evals = float(NB) * S * len(freqs)
    for prec, code in (("f32", OSKAR_SINGLE), ("f64", OSKAR_DOUBLE)):
        sky = Sky(code, OSKAR_GPU, S)
        for t, v in cols.items():
            sky.set_column(t, v)
        nf = C.c_int(0)
        lib.oskar_sky_evaluate_relative_directions(sky.ptr, ra0, dec0,
                                                   C.byref(status))
        lib.oskar_sky_evaluate_gaussian_source_parameters(
            sky.ptr, 0, ra0, dec0, C.byref(nf), C.byref(status))
        jt = code | OSKAR_COMPLEX | OSKAR_MATRIX
        jones = Jones(jt, OSKAR_GPU, N, S)
        jones.set(J if code == OSKAR_SINGLE else J.astype(np.float64))
        tel = Telescope(code, OSKAR_GPU, N)
        tel.set_station_ecef(uvw[0], uvw[1], np.zeros(N))
        lib.oskar_telescope_set_channel_bandwidth(tel.ptr, bw)
        lib.oskar_telescope_set_time_average(tel.ptr, tavg)
        lib.oskar_telescope_set_phase_centre(tel.ptr, 2, ra0, dec0)
        um = [Mem.from_numpy(x, code) for x in uvw]
        vis = Mem.from_numpy(np.zeros(NB * 8), jt)
        fl = capi.MEMA(*[sky.column(t).ptr for t in (3, 4, 5, 6)])
        dr = capi.MEM3(*[sky.column(t).ptr for t in (0x103, 0x104, 0x105)])
        ex = capi.MEM3(*[sky.column(t).ptr for t in (0x100, 0x101, 0x102)])
        uv = capi.MEM3(*[x.ptr for x in um])

        def one(f):
            lib.oskar_cross_correlate(0, 1, S, jones.ptr, fl, dr, ex, tel.ptr,
                                      uv, gast, float(f), 0, vis.ptr,
                                      C.byref(status))

        one(freqs[0])                       # warm-up
        lib.oskar_device_synchronize(OSKAR_GPU)
        lib.oskar_timer_start(timer)
        for f in freqs:
            one(f)
        lib.oskar_device_synchronize(OSKAR_GPU)
        t = lib.oskar_timer_elapsed(timer)
        check_status(status, "direct correlator (%s)" % prec)
        rate = evals / t
        out[prec] = {"ms_per_channel": round(1e3 * t / len(freqs), 3),
                     "evals_per_s": rate,
                     "achieved_TFLOPs_133_per_eval": round(rate * 133e-12,
                                                           2)}
        log("direct Gaussian + smearing %s: %.1f ms per channel, %.3e "
            "evals/s" % (prec, 1e3 * t / len(freqs), rate))
        del jones, sky, vis
    lib.oskar_timer_free(timer)
    out["fp32_vector_peak_TFLOPs"] = 157.3
    ref = oracle.ref()
    if ref is not None:
        dirs = oracle.relative_directions("d", ra, dec, ra0, dec0)
        abc = oracle.gaussian_parameters("d", ra, dec, maj, mn, pa, ra0,
                                         dec0)[:3]
        ns = 2000
        f = float(freqs[-1])
        ha0 = gast - ra0
        P = lambda a: np.ascontiguousarray(a, np.float64).ctypes.data_as(
            C.c_void_p)
        keep = [np.ascontiguousarray(J[:, :ns].astype(np.float64))]
        arrs = [cols[k][:ns] for k in (3, 4, 5, 6)] + \
            [x[:ns] for x in dirs] + [x[:ns] for x in abc] + uvw + \
            [uvw[0], uvw[1]]
        arrs = [np.ascontiguousarray(a, np.float64) for a in arrs]
        vout = np.zeros(NB * 8)
        D = C.c_double
        t0 = time.time()
        ref.oskar_cross_correlate_gaussian_omp_d(
            C.c_int(0), C.c_int(ns), C.c_int(N), C.c_int(0),
            keep[0].ctypes.data_as(C.c_void_p),
            *[a.ctypes.data_as(C.c_void_p) for a in arrs],
            D(0.0), D(3.4e38), D(f / C0), D(bw / f), D(tavg), D(ha0),
            D(dec0), vout.ctypes.data_as(C.c_void_p))
        tc = time.time() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["reference_cpu"] = {
            "evals_per_s": NB * ns / tc, "cores": cores,
            "kind": "reference",
            "sample": "512 stations x %d sources, one channel, FP64" % ns}
        out["speedup_f32_vs_reference_cpu"] = round(
            out["f32"]["evals_per_s"] / (NB * ns / tc), 1)
        log("reference direct Gaussian + smearing (CPU, %d cores): %.3e "
            "evals/s" % (cores, NB * ns / tc))
    return out


def bench_sky_params(args, rng):
    """oskar_sky_evaluate_relative_directions and
    oskar_sky_evaluate_gaussian_source_parameters on a 1M-source sky in
    device memory (sources within 40 deg of the phase centre, 30 %
    Gaussian with FWHM 5-60 arcsec, axis ratio 0.2-1), FP64 and FP32.  The
    Gaussian call includes its 4-byte read of the failure count.  The
    reference's own function (oracle/_ref, serial on the host, as the
    reference runs it) is timed on the first 100k sources of the FP64 sky.
    Bytes per source: relative directions 2 read + 3 written values;
    Gaussian parameters 5 read + 3 written (the 70 % point sources read 2
    and write nothing)."""
    import oracle
    from oskar_amd import lib
    from oskar_amd.mem import (OSKAR_DOUBLE, OSKAR_GPU, OSKAR_SINGLE, Sky,
                               check_status)
    S = 1_000_000
    ra0, dec0 = 0.5, -0.45
    r = np.radians(40.0) * np.sqrt(rng.uniform(0, 1, S))
    th = rng.uniform(0, 2 * np.pi, S)
    l, m = np.sin(r) * np.sin(th), np.sin(r) * np.cos(th)
    nn = np.sqrt(1 - l * l - m * m)
    dec = np.arcsin(nn * np.sin(dec0) + m * np.cos(dec0))
    ra = ra0 + np.arctan2(l, nn * np.cos(dec0) - m * np.sin(dec0))
    gauss = rng.uniform(0, 1, S) < 0.3
    maj = np.where(gauss, np.radians(rng.uniform(5, 60, S) / 3600), 0.0)
    mn = maj * rng.uniform(0.2, 1.0, S)
    pa = rng.uniform(0, np.pi, S)
    cols = {1: ra, 2: dec, 10: maj, 11: mn, 12: pa}
    out = {"workload": "1M-source sky within 40 deg of the phase centre, "
                       "30 % Gaussian (FWHM 5-60 arcsec)"}
    status = C.c_int(0)
    timer = lib.oskar_timer_create(1)
    for prec, code in (("f64", OSKAR_DOUBLE), ("f32", OSKAR_SINGLE)):
        sky = Sky(code, OSKAR_GPU, S)
        for t, v in cols.items():
            sky.set_column(t, v)
        nf = C.c_int(0)

        def rel():
            lib.oskar_sky_evaluate_relative_directions(
                sky.ptr, ra0, dec0, C.byref(status))

        def gau():
            lib.oskar_sky_evaluate_gaussian_source_parameters(
                sky.ptr, 0, ra0, dec0, C.byref(nf), C.byref(status))

        res = {}
        for name, fn in (("relative_directions", rel),
                         ("gaussian_parameters", gau)):
            fn()
            lib.oskar_device_synchronize(OSKAR_GPU)
            lib.oskar_timer_start(timer)
            for _ in range(args.sky_steps):
                fn()
            lib.oskar_device_synchronize(OSKAR_GPU)
            res[name] = 1e3 * lib.oskar_timer_elapsed(timer) / args.sky_steps
        check_status(status, "sky parameters")
        out[prec] = {"relative_directions_ms": round(
                         res["relative_directions"], 4),
                     "gaussian_parameters_ms": round(
                         res["gaussian_parameters"], 4),
                     "failed": nf.value // (args.sky_steps + 1)}
        log("sky parameters %s: relative directions %.3f ms, Gaussian "
            "parameters %.3f ms per 1M sources" % (
                prec, res["relative_directions"],
                res["gaussian_parameters"]))
        del sky
    lib.oskar_timer_free(timer)
    ref = oracle.ref_sky()
    if ref is not None:
        ns = 100_000
        st = C.c_int(0)
        rsky = ref.oskar_sky_create(8, 0, ns, C.byref(st))
        for t, v in cols.items():
            p = ref.oskar_sky_column(rsky, t, 0, C.byref(st))
            dst = np.ctypeslib.as_array(
                C.cast(ref.oskar_mem_void(p), C.POINTER(C.c_double)), (ns,))
            dst[:] = v[:ns]
        nf = C.c_int(0)
        t0 = time.time()
        ref.oskar_sky_evaluate_gaussian_source_parameters(
            rsky, 0, ra0, dec0, C.byref(nf), C.byref(st))
        tr = time.time() - t0
        ref.oskar_sky_free(rsky, C.byref(st))
        ms_1m = 1e3 * tr * S / ns
        out["reference_cpu"] = {
            "gaussian_parameters_ms_per_1M": round(ms_1m, 2),
            "cores": 1, "kind": "reference",
            "sample": "first 100k sources of the FP64 sky, serial"}
        out["speedup_gaussian_f64"] = round(
            ms_1m / out["f64"]["gaussian_parameters_ms"], 1)
        log("reference Gaussian parameters (CPU, serial): %.1f ms per 1M "
            "sources" % ms_1m)
    return out


def bench_vis(args, rng):
    """The correlator's output container at 512 stations: a device vis
    block of 10 times x 8 channels of FP32 2x2 cross-correlations
    (130 816 baselines, 335 MB) with auto-correlations and station
    (u, v, w), written `vis_blocks` times to a .vis file under /tmp
    (device-to-host through pinned staging, CRC-32C, file write), read back
    into a device block (file read, CRC-32C, host-to-device, baseline
    (u, v, w) from the station (u, v, w) by the HIP kernel), and the
    station-to-baseline kernel alone, timed with HIP events on the library
    stream.  Kernel bytes: 3 x 4 B read per station and 3 x 4 B written per
    baseline, per time."""
    import tempfile
    from oskar_amd import lib
    from oskar_amd.mem import OSKAR_GPU, OSKAR_SINGLE, Mem, check_status
    from oskar_amd.vis import Binary, VisBlock, VisHeader
    ns, nt, nc = 512, 10, 8
    nb = ns * (ns - 1) // 2
    amp = OSKAR_SINGLE | 0x20 | 0x40
    hdr = VisHeader(amp, OSKAR_SINGLE, nt, nt * args.vis_blocks, nc, nc, ns,
                    1, 1)
    blk = VisBlock(OSKAR_GPU, header=hdr)
    st = C.c_int(0)
    for m, n in ((blk.cross(), nt * nc * nb * 8), (blk.auto(), nt * nc * ns
                                                     * 8)):
        src = Mem.from_numpy(rng.normal(size=n).astype(np.float32), amp)
        lib.oskar_mem_copy(m.ptr, src.ptr, C.byref(st))
        src.free()
    for k in range(3):
        src = Mem.from_numpy(rng.uniform(-5e4, 5e4, nt * ns).astype(
            np.float32), OSKAR_SINGLE)
        lib.oskar_mem_copy(blk.station_uvw(k).ptr, src.ptr, C.byref(st))
        src.free()
    check_status(st, "vis block fill")
    block_bytes = (nt * nc * (nb + ns) * 32) + 3 * nt * ns * 4
    fd, path = tempfile.mkstemp(suffix=".vis", dir="/tmp")
    os.close(fd)
    try:
        h = hdr.write(path)
        t0 = time.perf_counter()
        for i in range(args.vis_blocks):
            lib.oskar_vis_block_set_start_time_index(blk.ptr, nt * i)
            blk.write(h, i)
        h.close()
        t_write = (time.perf_counter() - t0) / args.vis_blocks
        b = Binary(path, "r")
        got_hdr = VisHeader.read(b)
        got = VisBlock(OSKAR_GPU, header=got_hdr)
        got.read(got_hdr, b, 0)                 # warm-up (sizes arrays)
        t0 = time.perf_counter()
        for i in range(args.vis_blocks):
            got.read(got_hdr, b, i)
        lib.oskar_device_synchronize(OSKAR_GPU)
        t_read = (time.perf_counter() - t0) / args.vis_blocks
        b.close()
        same = bool(np.array_equal(got.cross().numpy()[:4096],
                                   blk.cross().numpy()[:4096]))
    finally:
        os.remove(path)
    # The conversion kernel alone, HIP events on the library stream.
    reps = 50
    timer = lib.oskar_timer_create(1)
    lib.oskar_vis_block_station_to_baseline_coords(blk.ptr, hdr.ptr,
                                                   C.byref(st))
    lib.oskar_timer_start(timer)
    for _ in range(reps):
        lib.oskar_vis_block_station_to_baseline_coords(blk.ptr, hdr.ptr,
                                                       C.byref(st))
    t_k = lib.oskar_timer_elapsed(timer) / reps
    lib.oskar_timer_free(timer)
    check_status(st, "station to baseline")
    ref_gbps = None
    if not args.no_cpu:
        # The reference's binary writer (oskar/binary/src, built into
        # oracle/_ref) on the same block bytes already in host memory.
        import oracle
        ref = oracle.ref_binary()
        if ref is not None:
            host = blk.cross().numpy()
            fd, path = tempfile.mkstemp(suffix=".dat", dir="/tmp")
            os.close(fd)
            try:
                rs = C.c_int(0)
                t0 = time.perf_counter()
                for i in range(args.vis_blocks):
                    rh = ref.oskar_binary_create(path.encode(), b"w",
                                                 C.byref(rs))
                    ref.oskar_binary_write(rh, amp, 12, 3, i, host.nbytes,
                                           host.ctypes.data, C.byref(rs))
                    ref.oskar_binary_free(rh)
                t_ref = (time.perf_counter() - t0) / args.vis_blocks
                ref_gbps = round(host.nbytes / t_ref / 1e9, 2)
            finally:
                os.remove(path)
            log("vis block: reference writer from host memory %.2f GB/s"
                % ref_gbps)
    k_bytes = nt * (3 * ns * 4 + 3 * nb * 4)
    k_gbps = k_bytes / t_k / 1e9
    w_gbps = block_bytes / t_write / 1e9
    r_gbps = block_bytes / t_read / 1e9
    log("vis block: write %.1f ms (%.2f GB/s), read %.1f ms (%.2f GB/s), "
        "station->baseline %.1f us (%.0f GB/s)" % (
            1e3 * t_write, w_gbps, 1e3 * t_read, r_gbps, 1e6 * t_k, k_gbps))
    return {"workload": "512 stations, 10 times x 8 channels, FP32 2x2 "
                        "cross- and auto-correlations, station uvw",
            "block_bytes": block_bytes, "blocks": args.vis_blocks,
            "write_ms": round(1e3 * t_write, 2),
            "write_GBps": round(w_gbps, 2),
            "read_ms": round(1e3 * t_read, 2),
            "read_GBps": round(r_gbps, 2),
            "read_back_identical": same,
            "reference_host_write_GBps": ref_gbps,
            "station_to_baseline_us": round(1e6 * t_k, 2),
            "station_to_baseline_GBps": round(k_gbps, 1),
            "station_to_baseline_frac": round(k_gbps / PEAK_HBM_GBPS, 4)}


def bench_batch(args, st, station, work, mode):
    """A representative subset of the 384-channel (50-350 MHz) x 10-sample
    observation at 512 stations: args.batch_units units spread over 4 time
    samples and evenly spaced channels.  Each time sample has its own
    above-horizon source count (200k at the first sample to 1M at the last),
    its own sky draw and station (u, v, w) (rotated by the Earth over 60 s
    per sample); each unit has its own station beam (exact FP32 DFT at its
    frequency) and fluxes (spectral index -0.7).  All units run in one
    persistent launch replayed from a hipGraph, with the tile shape, part
    size and grid chosen by the autotuner.  Error: relative RMS against FP64
    visibilities of every unit on all baselines."""
    from oskar_amd import lib
    from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_GPU, OSKAR_MATRIX,
                               OSKAR_SINGLE, FactorisedBatch, Mem, Telescope,
                               check_status)
    F32 = OSKAR_SINGLE
    CM = F32 | OSKAR_COMPLEX | OSKAR_MATRIX
    N = st["u"].size
    NB = N * (N - 1) // 2
    T, NCH = 10, 384
    nu = args.batch_units
    nt = min(4, nu)
    times = np.linspace(0, T - 1, nt).round().astype(int)
    per_t = [nu // nt + (1 if i < nu % nt else 0) for i in range(nt)]
    counts = {t: int(200_000 + 800_000 * t / (T - 1)) for t in range(T)}
    freqs = 50e6 + 300e6 * np.arange(NCH) / (NCH - 1)
    rng = np.random.default_rng(77)
    samples = {}
    units = []
    for i, t in enumerate(times):
        S = counts[t]
        z = rng.uniform(0.0, 1.0, S)
        z[z == 0.0] = 1e-7
        ph = rng.uniform(0, 2 * np.pi, S)
        s = np.sqrt(1.0 - z * z)
        sky = dict(l=s * np.sin(ph), m=s * np.cos(ph), n=z,
                   I=power_law(rng, S))
        ang = 2 * np.pi * 60.0 * t / 86164.0905
        ca, sa = np.cos(ang), np.sin(ang)
        uvw = dict(u=ca * st["u"] - sa * st["v"], v=sa * st["u"] + ca * st["v"],
                   w=st["w"])
        samples[t] = dict(
            sky=sky, uvw=uvw,
            dirs=[Mem.from_numpy(sky[k], F32) for k in ("l", "m", "n")],
            zero=Mem.from_numpy(np.zeros(S), F32),
            uvw_m=[Mem.from_numpy(uvw[k], F32) for k in ("u", "v", "w")])
        for c in np.linspace(0, NCH - 1, per_t[i]).round().astype(int):
            units.append((t, int(c)))
    log("batch: %d units, sources per sample %s" % (
        len(units), [counts[t] for t in times]))
    vis = Mem(CM, OSKAR_GPU, len(units) * NB)
    tel = Telescope(F32, OSKAR_GPU, N)
    b = FactorisedBatch(F32, mode, N, len(units))
    beams, fluxes = [], []
    status = C.c_int(0)
    total_evals = 0.0

    def eval_beams():
        for k, (t, c) in enumerate(units):
            d = samples[t]["dirs"]
            lib.oskar_evaluate_station_beam_aperture_array(
                station.ptr, work.ptr, counts[t], d[0].ptr, d[1].ptr,
                d[2].ptr, 0, 0.0, float(freqs[c]), beams[k].ptr,
                C.byref(status))

    for k, (t, c) in enumerate(units):
        smp = samples[t]
        S = counts[t]
        beams.append(Mem(CM, OSKAR_GPU, S))
        fluxes.append(Mem.from_numpy(
            smp["sky"]["I"] * (freqs[c] / 150e6) ** -0.7, F32))
        flux = [fluxes[k], smp["zero"], smp["zero"], smp["zero"]]
        b.set_unit(k, S, beams[k], flux, smp["dirs"], smp["uvw_m"],
                   float(freqs[c]), k * NB, vis)
        total_evals += float(NB) * S
    eval_beams()
    lib.oskar_device_synchronize(OSKAR_GPU)
    check_status(status, "batch beams")
    t0 = time.time()
    if not args.no_batch_tune:
        cfg = b.autotune(len(units))
        log("batch autotune: %s in %.1f s" % (cfg, time.time() - t0))
    cfg = b.config()
    # Warm-up (captures the graph), then timed replays.
    b.run(len(units), 0, tel)
    lib.oskar_device_synchronize(OSKAR_GPU)
    log("batch: graph captured")
    kernel_ms = []
    t0 = time.time()
    for i in range(args.batch_steps):
        b.run(len(units), 0, tel)
        kernel_ms.append(b.last_run_ms())
    lib.oskar_device_synchronize(OSKAR_GPU)
    wall = (time.time() - t0) / args.batch_steps
    s = C.c_int(0)
    lib.oskar_device_check_error_hip(C.byref(s))
    check_status(s, "batch")
    kms = float(np.median(kernel_ms))
    log("batch: %d runs, kernels %.2f ms (median), wall %.2f ms per run" % (
        args.batch_steps, kms, 1e3 * wall))
    # Beams + batch, as one pipeline step over all units.
    t0 = time.time()
    for i in range(args.batch_steps):
        eval_beams()
        b.run(len(units), 0, tel)
    lib.oskar_device_synchronize(OSKAR_GPU)
    wall_pipe = (time.time() - t0) / args.batch_steps
    check_status(status, "batch beams")
    log("batch: beams + batch %.2f ms per run" % (1e3 * wall_pipe))

    # Error of every unit on all baselines against FP64 (the direct FP64
    # kernel on the unit's FP32-rounded inputs and beam).
    err = []
    if not args.no_error:
        from fp64_reference import direct_fp64_vis, f32, rel_rms
        vis.clear()
        b.run(len(units), 0, tel)
        lib.oskar_device_synchronize(OSKAR_GPU)
        v_all = vis.numpy().reshape(len(units), NB * 8)
        t0 = time.time()
        for k, (t, c) in enumerate(units):
            smp = samples[t]
            sky, uvw = smp["sky"], smp["uvw"]
            ref = direct_fp64_vis(
                *(f32(uvw[x]) for x in ("u", "v", "w")),
                *(f32(sky[x]) for x in ("l", "m", "n")),
                fluxes[k].numpy().astype(np.float64), float(freqs[c]),
                beam=beams[k].numpy().astype(np.float64).reshape(-1, 8))
            err.append(rel_rms(v_all[k], ref))
        log("batch: FP64 reference on all baselines of %d units: %.1f s; "
            "max rel RMS %.3e" % (len(units), time.time() - t0, max(err)))
    return {
        "workload": "%d of 3840 (channel, time) units of 384 ch (50-350 MHz) "
                    "x 10 samples, 512 stations, 200k-1M sources per sample"
                    % len(units),
        "units": [{"time": int(t), "channel": int(c),
                   "freq_hz": float(freqs[c]), "sources": counts[t]}
                  for t, c in units],
        "evals": total_evals,
        "kernel_ms": round(kms, 3),
        "evals_per_s": total_evals / (kms * 1e-3),
        "wall_ms_per_run": round(1e3 * wall, 3),
        "pipeline_ms_per_run_with_beams": round(1e3 * wall_pipe, 3),
        "pipeline_evals_per_s_with_beams": total_evals / wall_pipe,
        "config": cfg,
        "autotuned": not args.no_batch_tune,
        "rel_rms_vs_fp64_per_unit": [float("%.3e" % e) for e in err],
        "rel_rms_vs_fp64_max": max(err) if err else None,
    }



def bench_stations(args, st, sky, dirs, flux, uvw, mode, timed, evals):
    """The 512-station workload read with distinct station layouts: every
    station has its own 256-dipole layout (uniform in a 38 m disc), all
    pointed 30 deg from zenith and normalised there.  One step:
    oskar_evaluate_jones_E (512 exact FP32 beams of the 1M sources) +
    oskar_cross_correlate_factorised_jones (per-station factorised MFMA
    product).  Beside it: the plain FP32 pipeline on the same E
    (oskar_evaluate_jones_K, oskar_jones_join, oskar_cross_correlate) and
    the error of both against FP64 on every baseline."""
    from oskar_amd import capi, lib
    from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_DOUBLE, OSKAR_GPU,
                               OSKAR_MATRIX, OSKAR_SINGLE, Jones, Mem,
                               StationView, StationWork, Telescope,
                               check_status)
    from workloads import elements
    F32 = OSKAR_SINGLE
    CM = F32 | OSKAR_COMPLEX | OSKAR_MATRIX
    N, S, f = st["u"].size, sky["l"].size, args.freq
    NB = N * (N - 1) // 2
    ENU_DIR, AZEL = 1, 4
    status = C.c_int(0)
    tel = Telescope(F32, OSKAR_GPU, N)
    lib.oskar_telescope_resize_station_array(tel.ptr, N, C.byref(status))
    lib.oskar_telescope_set_phase_centre(tel.ptr, AZEL, 0.0, np.radians(60))
    rng = np.random.default_rng(31)
    for k in range(N):
        sv = StationView(lib.oskar_telescope_station(tel.ptr, k), F32)
        lib.oskar_station_resize(sv.ptr, args.elements, C.byref(status))
        sv.set_coords(elements(rng, args.elements))
        lib.oskar_station_set_phase_centre(sv.ptr, AZEL, 0.0,
                                           np.radians(60.0))
    check_status(status, "station models")
    work = StationWork(F32, OSKAR_GPU)
    E = Jones(CM, OSKAR_GPU, N, S)
    vis = Mem(CM, OSKAR_GPU, NB)
    mjd = 60000.0

    def step_E():
        lib.oskar_evaluate_jones_E(
            E.ptr, ENU_DIR, S, capi.MEM3(*[m.ptr for m in dirs]), 0.0, 0.0,
            tel.ptr, 0, mjd, mjd, f, work.ptr, C.byref(status))

    def step_fs():
        lib.oskar_cross_correlate_factorised_jones(
            0, S, E.ptr, capi.MEMA(*[m.ptr for m in flux]),
            capi.MEM3(*[m.ptr for m in dirs]), tel.ptr,
            capi.MEM3(*[m.ptr for m in uvw]), f, mode, 0, vis.ptr,
            C.byref(status))

    def step():
        step_E()
        step_fs()

    if args.station_part > 0:
        lib.oskar_cross_correlate_factorised_jones_set_part_size(
            args.station_part)
    t_E = timed(step_E, 2, 1, "jones E (512 distinct beams)") / 2
    t_fs = timed(step_fs, args.steps, 1, "factorised jones") / args.steps
    lib.oskar_cross_correlate_factorised_set_profiling(1)
    stage = np.zeros(3)
    for i in range(3):
        step_fs()
        ms = (C.c_double * 3)()
        assert lib.oskar_cross_correlate_factorised_jones_last_times(ms) == 0
        stage += np.array(ms[:]) / 3
    lib.oskar_cross_correlate_factorised_set_profiling(0)
    t_step = timed(step, 2, 0, "jones E + factorised jones") / 2
    vis.clear()
    step_fs()
    lib.oskar_device_synchronize(OSKAR_GPU)
    vis_fs = vis.numpy()
    check_status(status, "stations leg")
    # Plain FP32 pipeline on the same station beams.
    K = Jones(F32 | OSKAR_COMPLEX, OSKAR_GPU, N, S)
    J = Jones(CM, OSKAR_GPU, N, S)
    ext = [Mem(F32, OSKAR_GPU, 1) for _ in range(3)]

    def step_plain_corr():
        lib.oskar_evaluate_jones_K(K.ptr, S, dirs[0].ptr, dirs[1].ptr,
                                   dirs[2].ptr, uvw[0].ptr, uvw[1].ptr,
                                   uvw[2].ptr, f, flux[0].ptr, -1e30, 1e30,
                                   0, C.byref(status))
        lib.oskar_jones_join(J.ptr, E.ptr, K.ptr, C.byref(status))
        lib.oskar_cross_correlate(
            0, 0, S, J.ptr, capi.MEMA(*[m.ptr for m in flux]),
            capi.MEM3(*[m.ptr for m in dirs]),
            capi.MEM3(*[m.ptr for m in ext]), tel.ptr,
            capi.MEM3(*[m.ptr for m in uvw]), 0.0, f, 0, vis.ptr,
            C.byref(status))

    vis.clear()
    t_pc = timed(step_plain_corr, 1, 0, "plain FP32 on distinct beams")
    vis_plain = vis.numpy()
    del K, J
    err = err_plain = None
    if not args.no_error:
        from fp64_reference import direct_fp64_vis, f32, rel_rms
        e_host = E.mem.numpy().reshape(N, S, 8)
        CM64 = OSKAR_DOUBLE | OSKAR_COMPLEX | OSKAR_MATRIX

        def e_chunk(s0, s1, mem):
            up = Mem.from_numpy(e_host[:, s0:s1].astype(np.float64).ravel(),
                                CM64, 0)
            s = C.c_int(0)
            lib.oskar_mem_copy_contents(mem, up.ptr, 0, 0, N * (s1 - s0),
                                        C.byref(s))
            check_status(s, "E chunk")

        t0 = time.time()
        ref = direct_fp64_vis(*(f32(st[k]) for k in ("u", "v", "w")),
                              *(f32(sky[k]) for k in ("l", "m", "n")),
                              f32(sky["I"]), f, station_beams=e_chunk)
        del e_host
        err = rel_rms(vis_fs, ref)
        err_plain = rel_rms(vis_plain, ref)
        log("distinct stations: FP64 reference %.1f s; rel RMS factorised "
            "%.3e, plain FP32 %.3e" % (time.time() - t0, err, err_plain))
    rate = evals / t_step
    plain_rate = evals / (t_E + t_pc)
    log("distinct stations: beams %.2f ms, factorised %.2f ms (operands "
        "%.2f, product %.2f, reduce %.2f), step %.2f ms = %.3e evals/s; "
        "plain FP32 %.3e evals/s" % (1e3 * t_E, 1e3 * t_fs, *stage,
                                     1e3 * t_step, rate, plain_rate))
    achieved = 64.0 * evals / (stage[1] * 1e-3) / 1e12
    return {
        "workload": "512 stations x 1M point sources, 1 channel, a distinct "
                    "256-dipole layout per station (exact FP32 beams, "
                    "oskar_evaluate_jones_E) -> "
                    "oskar_cross_correlate_factorised_jones",
        "evals_per_s": rate,
        "ms_per_step": round(1e3 * t_step, 3),
        "beams_ms": round(1e3 * t_E, 3),
        "factorised_ms": round(1e3 * t_fs, 3),
        "stage_ms": {"operands": round(stage[0], 4),
                     "product": round(stage[1], 4),
                     "reduce": round(stage[2], 4)},
        "product_algorithmic_tflops": round(achieved, 2),
        "product_frac_of_2500": round(achieved / PEAK_MFMA_TFLOPS, 4),
        "plain_fp32_evals_per_s": plain_rate,
        "speedup_vs_plain_fp32": rate / plain_rate,
        "rel_rms_vs_fp64": err,
        "rel_rms_plain_fp32_vs_fp64": err_plain,
    }


def bench_surrogate(args, st, sky, station, work, mode, tel, flux, dirs, uvw,
                    beam_exact, ref, timed, evals):
    """The learned station-beam surrogate of the 256-dipole station over
    50-350 MHz: trained on the GPU on 10M samples of the exact beam (200
    frequencies x 50,000 directions), then used in place of the exact beam
    for the 1M-source workload at 150 MHz (with the exact beam where the
    surrogate is below 1e-3 of the peak), feeding the factorised correlator.
    Reports training time, the beam error over the sources, the fused
    inference time against the exact DFT at 256 and at 4096 elements (the
    surrogate's cost does not depend on the element count), and the
    visibility error of the surrogate pipeline against FP64 with the exact
    beam."""
    from oskar_amd import capi, lib
    from oskar_amd.mem import (OSKAR_COMPLEX, OSKAR_GPU, OSKAR_MATRIX,
                               OSKAR_SINGLE, BeamSurrogate, Mem, Station,
                               StationWork, beam_training_set, check_status)
    F32 = OSKAR_SINGLE
    CM = F32 | OSKAR_COMPLEX | OSKAR_MATRIX
    N = st["u"].size
    NB = N * (N - 1) // 2
    S = sky["l"].size
    f = args.freq
    l0, m0 = 0.0, 0.5            # beam at azimuth 0, elevation 60 deg
    t0 = time.time()
    tl, tm, tf, tb = beam_training_set(station, work, 200, 50000, 50e6,
                                       350e6, l0, m0, seed=11)
    lib.oskar_device_synchronize(OSKAR_GPU)
    t_set = time.time() - t0
    log("surrogate: %d exact-beam samples in %.2f s" % (len(tl), t_set))
    sur = BeamSurrogate(50e6, 350e6, l0, m0, 19.0, seed=5)
    t0 = time.time()
    hist = sur.train(len(tl), tl, tm, tf, tb, 8192, args.surrogate_steps,
                     2e-3, 2e-5, seed=6)
    lib.oskar_device_synchronize(OSKAR_GPU)
    t_train = time.time() - t0
    log("surrogate: %d steps in %.2f s, loss %.3e -> %.3e" % (
        args.surrogate_steps, t_train, hist[:10].mean(), hist[-50:].mean()))
    del tl, tm, tf, tb
    status = C.c_int(0)
    bs = Mem(CM, OSKAR_GPU, S)

    def step_sur():
        sur.evaluate(S, dirs[0], dirs[1], f, bs)

    def step_exact():
        lib.oskar_evaluate_station_beam_aperture_array(
            station.ptr, work.ptr, S, dirs[0].ptr, dirs[1].ptr, dirs[2].ptr,
            0, 0.0, f, beam_exact.ptr, C.byref(status))

    reps = 20
    t_sur = timed(step_sur, reps, 2, "surrogate inference") / reps
    t_ex = timed(step_exact, reps, 2, "exact beam 256") / reps
    # A 4096-element station of the same density (radius 76 m).
    rng = np.random.default_rng(9)
    ne = 4096
    re = 76.0 * np.sqrt(rng.uniform(size=ne))
    te = rng.uniform(0, 2 * np.pi, ne)
    el = np.stack([re * np.cos(te), re * np.sin(te), np.zeros(ne)])
    big = Station(F32, OSKAR_GPU, ne)
    big.set_coords(el, el)
    lib.oskar_station_set_phase_centre(big.ptr, 4, 0.0, np.radians(60.0))
    wbig = StationWork(F32, OSKAR_GPU)
    bbig = Mem(CM, OSKAR_GPU, S)

    def step_big():
        lib.oskar_evaluate_station_beam_aperture_array(
            big.ptr, wbig.ptr, S, dirs[0].ptr, dirs[1].ptr, dirs[2].ptr, 0,
            0.0, f, bbig.ptr, C.byref(status))

    t_big = timed(step_big, 5, 1, "exact beam 4096") / 5
    del bbig, big, wbig
    # Beam error over the sources (exact beam at 150 MHz as truth).
    step_exact()
    step_sur()
    lib.oskar_device_synchronize(OSKAR_GPU)
    be = beam_exact.numpy().astype(np.float64)
    bsur = bs.numpy().astype(np.float64)
    beam_err = float(np.sqrt(np.sum((bsur - be) ** 2) / np.sum(be ** 2)))
    # With the exact-beam fallback below 1e-3 of the peak.
    nfb = sur.evaluate_with_fallback(station, work, S, dirs[0], dirs[1],
                                     dirs[2], f, 1e-3, bs)
    lib.oskar_device_synchronize(OSKAR_GPU)
    log("surrogate: inference %.3f ms, exact 256 el %.3f ms, exact 4096 el "
        "%.3f ms; beam rel RMS %.3e; %d exact fallbacks" % (
            1e3 * t_sur, 1e3 * t_ex, 1e3 * t_big, beam_err, nfb))
    # Surrogate beam -> factorised correlator.
    vis = Mem(CM, OSKAR_GPU, NB)

    def step_pipe():
        step_sur()
        lib.oskar_cross_correlate_factorised(
            0, S, bs.ptr, capi.MEMA(*[x.ptr for x in flux]),
            capi.MEM3(*[x.ptr for x in dirs]), tel.ptr,
            capi.MEM3(*[x.ptr for x in uvw]), f, mode, 0, vis.ptr,
            C.byref(status))

    t_pipe = timed(step_pipe, args.steps, 1, "surrogate pipeline") / \
        args.steps
    vis_err = None
    if ref is not None:
        vis.clear()
        sur.evaluate_with_fallback(station, work, S, dirs[0], dirs[1],
                                   dirs[2], f, 1e-3, bs)
        step = lambda: lib.oskar_cross_correlate_factorised(
            0, S, bs.ptr, capi.MEMA(*[x.ptr for x in flux]),
            capi.MEM3(*[x.ptr for x in dirs]), tel.ptr,
            capi.MEM3(*[x.ptr for x in uvw]), f, mode, 0, vis.ptr,
            C.byref(status))
        step()
        lib.oskar_device_synchronize(OSKAR_GPU)
        got = vis.numpy()
        vis_err = float(np.sqrt(np.sum((got - ref) ** 2) / np.sum(ref ** 2)))
        log("surrogate pipeline: vis rel RMS vs FP64 (exact beam) %.3e" %
            vis_err)
    check_status(status, "surrogate")
    return {
        "workload": "256-dipole station, 50-350 MHz; 10M exact-beam "
                    "training samples (200 freq x 50k directions); "
                    "inference for the 1M sources at 150 MHz",
        "training_samples": 200 * 50000,
        "training_set_s": round(t_set, 3),
        "train_steps": args.surrogate_steps, "batch": 8192,
        "train_s": round(t_train, 3),
        "loss_first": float(hist[:10].mean()),
        "loss_last": float(hist[-50:].mean()),
        "beam_rel_rms_vs_exact": beam_err,
        "inference_ms_1M": round(1e3 * t_sur, 4),
        "exact_beam_ms_1M_256el": round(1e3 * t_ex, 4),
        "exact_beam_ms_1M_4096el": round(1e3 * t_big, 4),
        "exact_fallbacks_below_1e-3_peak": nfb,
        "pipeline_ms": round(1e3 * t_pipe, 3),
        "pipeline_evals_per_s": evals / t_pipe,
        "vis_rel_rms_vs_fp64_exact_beam": vis_err,
    }


if __name__ == "__main__":
    main()
