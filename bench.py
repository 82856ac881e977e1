#!/usr/bin/env python3
"""Benchmark: photons propagated per second on one MI355X.

Workload (one step = one launch of the propagation kernel over the batch):
  Cherenkov cascade — a 1e4 GeV electron cascade (ExtendedCherenkovEmitter,
  longitudinal Γ profile a = 2.01849 + 1.45469·log10 E, b = 0.63207,
  X0 = 0.3561 m) emitting 1e7 photons (300-800 nm Frank-Tamm table, 30
  knots), propagated through Cascadia water (HG g = 0.95) to a string of 20
  MultiPMTDetector modules (r = 0.21 m, 16 PMTs) at 50 m vertical spacing,
  20 m from the cascade axis.  Synthetic inputs, all device-resident before
  the timed region; hits go to device-resident columns.
  `--workload cluster`: the same cascade inside the serving cluster (a centre
  string and a hexagon of strings at 50 m, 7 × 20 modules), for the kernel's
  rate with several strings; `--workload isotropic`: an isotropic point source
  and one sphere at 10 m.

Prints one JSON line.  `roofline` is for the propagation kernel (FP32 VALU
bound): algorithmic FLOP per launch from the kernel's own step/test counters
÷ the kernel time measured with HIP events on its stream; `traffic` is the
HBM bytes per launch from the committed rocprofv3 summary for this workload.
`cpu_baseline`
times the CPU oracle (the reference algorithm restated in C, OpenMP) on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP32_PEAK_TFLOPS = 157.3        # MI355X published FP32 vector peak
PMT_AREA = (75e-3 / 2) ** 2 * math.pi

# Algorithmic work (counting rule: add/mul/cmp/select 1, fma 2, div 1,
# transcendental 1): per scattering step 61 FLOP + 4 transcendentals besides
# the sphere tests, 18 FLOP per sphere test, 27 FLOP per string-bound test
# (the kernel tests a step against each vertical string's bounding cylinder
# before its spheres); per photon initialisation 49 FLOP
# + 9 transcendentals (isotropic) and 95 FLOP + 20 transcendentals more for an
# extended Cherenkov emitter.
FLOP_STEP = 61 + 4
FLOP_TEST = 18
FLOP_STRING_BOUND = 27   # z-range + segment/axis distance test of one string
FLOP_INIT = {"isotropic": 49 + 9, "isotropic_r021": 49 + 9, "cascade": 49 + 9 + 95 + 20,
             "cluster": 49 + 9 + 95 + 20}
WORKLOAD_NAMES = {"cascade": "cherenkov_cascade_1e4GeV_1e7ph_20x16pmt_string",
                  "cluster": "cherenkov_cascade_1e4GeV_in_7x20x16pmt_cluster",
                  "isotropic": "isotropic_point_source_1_sphere_10m"}


PROFILE = ROOT / "profiles" / "propagate_kernel_cascade_1e7ph.md"
_T0 = time.perf_counter()


def progress(msg: str) -> None:
    """One line on stderr when a part of the run ends (stdout carries only the JSON result)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def profiled_traffic(args):
    """HBM bytes per launch of the propagation kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (scaled by the copy calibration),
    for the configuration that summary was taken on; None otherwise."""
    if (args.workload != "cascade" or args.photons != 10_000_000 or args.no_culling
            or not PROFILE.exists()):
        return None
    for line in PROFILE.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":", 1)[1])
    return None


def build_workload(name: str, photons: int, direction_arithmetic: str = "fp64"):
    import neutrinotelescopes_amd as nt
    med = nt.make_cascadia_medium_properties(0.95, direction_arithmetic=direction_arithmetic)
    if name == "cascade":
        spec = nt.CherenkovSpectrum((300.0, 800.0), 30, med)
        lp = nt.LongitudinalParameterisation.from_energy(1e4, med, nt.PEMinus)
        src = nt.ExtendedCherenkovEmitter(np.zeros(3, np.float32), np.array([1, 0, 0], np.float32),
                                          0.0, photons, lp)
        coords = nt.make_pom_pmt_coordinates(np.float32)
        tgts = [nt.MultiPMTDetector(np.array([20.0, 0.0, -475.0 + 50.0 * k], np.float32), 0.21,
                                    PMT_AREA, coords, k + 1) for k in range(20)]
    elif name == "cluster":
        # the serving cluster: a centre string and a hexagon of strings at
        # 50 m, 20 modules each; the cascade inside it, pointing obliquely up
        from neutrinotelescopes_amd.serving import cluster_targets
        spec = nt.CherenkovSpectrum((300.0, 800.0), 30, med)
        lp = nt.LongitudinalParameterisation.from_energy(1e4, med, nt.PEMinus)
        src = nt.ExtendedCherenkovEmitter(np.array([10.0, 20.0, 5.0], np.float32),
                                          np.array([0.6, 0.0, 0.8], np.float32), 0.0, photons, lp)
        tgts = cluster_targets()
    elif name in ("isotropic", "isotropic_r021"):
        spec = nt.CherenkovSpectrum((300.0, 600.0), 30, med)
        src = nt.PointlikeIsotropicEmitter(np.zeros(3, np.float32), 0.0, photons)
        tgts = [nt.DetectionSphere(np.array([0, 0, 10.0], np.float32), 0.21, 16, PMT_AREA, 1)]
    else:
        raise ValueError(name)
    return nt.PhotonPropSetup([src], tgts, med, spec)


def host_cpu_info() -> dict:
    """The host CPUs this process may use: nproc, its affinity mask, the
    cgroup quota and OMP_NUM_THREADS (the box grants one GPU a share of the
    host), and the CPU model from lscpu."""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        quota, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if quota != "max":
            info["cgroup_cpus"] = int(quota) / int(period)
    except (OSError, ValueError):
        pass
    model = None
    try:
        import subprocess
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, ValueError):
        pass
    info["cpu_model"] = model or _cpu_model()
    share = info["affinity_cpus"]
    if info["omp_num_threads"]:
        share = min(share, int(info["omp_num_threads"]))
    if "cgroup_cpus" in info:
        share = min(share, max(1, int(info["cgroup_cpus"])))
    info["cpu_share"] = share
    return info


def cpu_baseline(workload: str, target_s: float = 12.0):
    """Oracle photons/s on the host cores for a bounded sample: the C
    restatement of the reference kernel built here with -march=native (the
    shipped x86-64-v3 build if the compiler is missing), OpenMP over all CPUs
    of this process's share, and on one core."""
    import tempfile
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle
    info = host_cpu_info()
    try:
        lib = _oracle.load_native(tempfile.mkdtemp(prefix="ntoracle_"))
        build = "gcc -O3 -march=native -fopenmp -ffp-contract=off"
    except Exception:                       # no compiler: the shipped build
        lib = _oracle.load()
        build = "gcc -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off (shipped)"
    threads = info["cpu_share"]
    n = 20000
    for _ in range(6):
        s = build_workload(workload, n)
        t0 = time.perf_counter()
        _oracle.propagate(s, seed=1, threads=threads, capacity=1 << 20, lib=lib)
        dt = time.perf_counter() - t0
        if dt > 1.0:
            break
        n = int(n * min(50.0, 1.5 / max(dt, 1e-3)))
    # the sizing runs above are the warm-up; then the median of five timed
    # runs of a fifth of the target time each
    n_final = int(n * max(1.0, target_s / 5.0 / max(dt, 1e-6)))
    s = build_workload(workload, n_final)
    times = []
    for rep in range(5):
        t0 = time.perf_counter()
        _oracle.propagate(s, seed=2 + rep, threads=threads, capacity=1 << 22, lib=lib)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    rate = n_final / dt
    # the same workload on one core, a ~4 s sample
    n1 = max(10000, int(rate / threads * 4.0))
    s1 = build_workload(workload, n1)
    t0 = time.perf_counter()
    _oracle.propagate(s1, seed=7, threads=1, capacity=1 << 22, lib=lib)
    dt1 = time.perf_counter() - t0
    return {"value": rate, "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": f"{n_final} photons of the {workload} workload (FP64 direction arithmetic) on {threads} threads, "
                      f"median of 5 runs ({min(times):.2f}-{max(times):.2f} s) after warm-up (CPU oracle, {build}, "
                      f"OpenMP)",
            "one_core_value": n1 / dt1, "one_core_sample": f"{n1} photons, {dt1:.1f} s on 1 thread",
            "host": info}


def measure_kernel(workload: str, photons: int, arithmetic: str = "fp64", steps: int = 20, warmup: int = 3,
                   culling: bool = True, epilogue: bool = False, config_mode: bool = False) -> dict:
    """One workload on nt_launch: `warmup` untimed launches, then `steps`
    timed launches (HIP events on the launch stream bracket each one; the
    wall clock brackets all of them), inputs and hit columns resident in HBM.
    FLOP per launch from the kernel's own step and test counters."""
    import torch
    from neutrinotelescopes_amd import _abi
    from neutrinotelescopes_amd.propagation import PropagationEngine
    dev = torch.device("cuda:0")
    setup = build_workload(workload, photons, arithmetic)
    eng = PropagationEngine(0)
    try:
        eng.setup(setup)
        if config_mode:
            # the configuration's physics where it departs from the reference:
            # emission at 41.2°, a tabulated angular acceptance
            eng.set_cherenkov_angle(math.radians(41.2))
        if epilogue:
            import neutrinotelescopes_amd as nt
            eng.set_epilogue(setup.medium, source_position=setup.sources[0].position,
                             angular_acceptance=nt.lambertian_angular_acceptance() if config_mode else None)
        _abi.check(eng.lib.nt_set_option(eng.handle, _abi.NT_OPT_TARGET_CULLING, int(culling)), "nt_set_option")
        cap = 1 << 22
        cols = dict(position=torch.empty(cap, 3, device=dev), direction=torch.empty(cap, 3, device=dev),
                    initial_direction=torch.empty(cap, 3, device=dev), wavelength=torch.empty(cap, device=dev),
                    time=torch.empty(cap, device=dev), dist_travelled=torch.empty(cap, device=dev),
                    module_id=torch.empty(cap, dtype=torch.int16, device=dev),
                    source_index=torch.empty(cap, dtype=torch.int32, device=dev),
                    photon_id=torch.empty(cap, dtype=torch.int64, device=dev))
        if epilogue:
            cols.update(pmt_id=torch.empty(cap, dtype=torch.uint8, device=dev),
                        total_weight=torch.empty(cap, device=dev), tres=torch.empty(cap, device=dev))
        counters = torch.zeros(4, dtype=torch.int64, device=dev)
        h = _abi.nt_hits()
        for k, v in cols.items():
            setattr(h, k, v.data_ptr())
        h.capacity = cap
        stream = torch.cuda.current_stream()

        def step(seed):
            _abi.check(eng.lib.nt_launch(eng.handle, C.c_uint64(seed), C.byref(h),
                                         C.c_void_p(counters.data_ptr()), C.c_void_p(stream.cuda_stream)),
                       "nt_launch")

        for i in range(warmup):
            step(1000 + i)
        torch.cuda.synchronize()
        c = counters.cpu().numpy()
        hits_per_step, tests_per_step, bound_tests_per_step = int(c[0]), int(c[2]), int(c[3])
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            ev[i][0].record(stream)
            step(i)
            ev[i][1].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kernel_ms = [a.elapsed_time(b) for a, b in ev]
        c = counters.cpu().numpy()
        assert int(c[1]) == photons, c
    finally:
        eng.close()
    n_targets = len(setup.targets)
    n_strings = len({(float(t.position[0]), float(t.position[1])) for t in setup.targets})
    # every step tests every string bound (culling on) or every sphere (off)
    steps_total = (bound_tests_per_step / n_strings if bound_tests_per_step
                   else tests_per_step / max(n_targets, 1))
    flop = (photons * FLOP_INIT[workload] + steps_total * FLOP_STEP
            + tests_per_step * FLOP_TEST + bound_tests_per_step * FLOP_STRING_BOUND)
    k_ms = float(np.mean(kernel_ms))
    achieved = flop / (k_ms * 1e-3) / 1e12
    return {"value": photons * steps / wall, "ms_per_step": wall / steps * 1e3, "kernel_ms": k_ms,
            "config": {"photons_per_step": photons, "targets": n_targets, "hits_per_step": hits_per_step,
                       "direction_arithmetic": arithmetic, "string_culling": culling, "hit_epilogue": epilogue,
                       "config_mode": config_mode,
                       "scattering_steps_per_photon": steps_total / photons,
                       "sphere_tests_per_scattering_step": tests_per_step / max(steps_total, 1),
                       "string_bound_tests_per_scattering_step": bound_tests_per_step / max(steps_total, 1)},
            "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP32_PEAK_TFLOPS, "kernel": "propagate_kernel", "kernel_ms": k_ms,
                         "flop_per_launch": flop}}


def other_legs(args) -> dict:
    """Beside the headline: the same cascade in the reference's FP32 direction
    arithmetic, the cascade with the device hit epilogue (PMT id, weights,
    time residual: the acceptance the C2 configuration names) run after each
    launch, and the plain configuration's throughput size (isotropic source,
    one r = 0.21 m sphere at 10 m, 1e9 photons per launch)."""
    legs = {}
    r = measure_kernel("cascade", args.photons, "fp32", args.steps, args.warmup)
    legs["cascade_fp32_reference_arithmetic"] = {"value": r["value"], "kernel_ms": r["kernel_ms"],
                                                 "frac": r["roofline"]["frac"]}
    progress("leg cascade_fp32_reference_arithmetic")
    r = measure_kernel("cascade", args.photons, args.direction_arithmetic, args.steps, args.warmup, epilogue=True)
    legs["cascade_with_hit_epilogue"] = {"value": r["value"], "ms_per_step": r["ms_per_step"],
                                         "hits_per_step": r["config"]["hits_per_step"]}
    progress("leg cascade_with_hit_epilogue")
    r = measure_kernel("cascade", args.photons, args.direction_arithmetic, args.steps, args.warmup, epilogue=True,
                       config_mode=True)
    legs["cascade_config_mode_41deg_angular_acceptance"] = {"value": r["value"], "ms_per_step": r["ms_per_step"],
                                                            "hits_per_step": r["config"]["hits_per_step"]}
    for arith in (args.direction_arithmetic, "fp32"):
        r = measure_kernel("isotropic_r021", 1_000_000_000, arith, 3, 1)
        legs[f"isotropic_1e9ph_r021_{arith}"] = {"value": r["value"], "kernel_ms": r["kernel_ms"],
                                                "frac": r["roofline"]["frac"],
                                                "hits_per_step": r["config"]["hits_per_step"]}
        progress(f"leg isotropic_1e9ph_r021_{arith}")
    return legs


def serving_benchmark(mode: str = "worklist", rate_hz: float | None = None, duration_s: float = 15.0,
                      seed: int = 11, round_photons: int = 10_000_000, max_inflight_rounds: int = 2):
    """Serving configuration: Poisson stream of cascade requests (position
    uniform in a 300 m cylinder, isotropic direction, E log-uniform 1e2-1e6 GeV,
    1e4·E/100 GeV photons) against a 7-string × 20-module cluster, arrival
    rate `rate_hz`, by default 0.7 × the server's measured capacity.  `mode`:
    "worklist" (one persistent kernel claims photon tickets of published
    requests) or "rounds" (requests packed into launches of `round_photons`)."""
    import neutrinotelescopes_amd as nt
    from neutrinotelescopes_amd.propagation import PropagationEngine
    from neutrinotelescopes_amd.serving import PhotonServer, cascade_requests, cluster_targets, run_poisson_stream
    med = nt.make_cascadia_medium_properties(0.95)
    eng = PropagationEngine(0)
    eng.set_medium(med)
    eng.set_spectrum(nt.CherenkovSpectrum((300.0, 800.0), 30, med))
    eng.set_targets(cluster_targets())
    rng = np.random.default_rng(seed)
    mean_photons = 1e4 * (1e6 - 1e2) / math.log(1e4) / 100.0
    with PhotonServer(eng, device_budget=32 << 30, host_budget=8 << 30, round_photons=round_photons,
                      max_inflight_rounds=max_inflight_rounds, mode=mode) as srv:
        # capacity: a burst submitted at once (256 requests, ~3e9 photons, so
        # that the partly filled first and last rounds weigh little)
        burst = cascade_requests(256, med, rng)
        t0 = time.perf_counter()
        ids = [srv.submit(r.sources, r.seed, r.capacity) for r in burst]
        for i in ids:
            srv.wait(i, raise_on_overflow=False)
        cap_rate = sum(s.photons for r in burst for s in r.sources) / (time.perf_counter() - t0)
        progress(f"serving ({mode}) capacity burst: {cap_rate:.4g} photons/s")
        rate = rate_hz if rate_hz is not None else 0.7 * cap_rate / mean_photons
        n = max(50, int(rate * duration_s))
        reqs = cascade_requests(n, med, rng)
        res = run_poisson_stream(srv, reqs, rate_hz=rate, rng=rng)
        st = srv.stats()
    eng.close()
    return {"config": "poisson_cascades_1e2-1e6GeV_7x20_cluster", "mode": mode, "requests": res["requests"],
            "arrival_rate_hz": rate, "p50_ms": res["p50_ms"], "p99_ms": res["p99_ms"],
            "photons_per_s": res["photons_per_s"], "capacity_photons_per_s": cap_rate,
            "launches": st["rounds"], "device_busy_fraction": st["busy_ms"] / 1e3 / res["wall_s"],
            **({"round_photons": round_photons, "max_inflight_rounds": max_inflight_rounds}
               if mode == "rounds" else {})}


def surrogate_benchmark():
    """Flow surrogate: 128 engine photon-table launches (cascades log10 E
    uniform 2-6, 2e8 photons each, 192 modules at 1-300 m), 192000 Adam steps
    (lr 1e-3, ×0.1 over the last third, weights averaged over that third) of
    batch 16384 on the bf16-MFMA training kernels (count rows weighted ×2),
    per-module expected-count
    error and time-density KL against direct propagation (128 cascades; pairs
    selected with >= 400 hits in one 2e8-photon run, truth from an independent
    6e8-photon run), and event
    likelihood throughput for 1024 cascades (log-uniform 1e3-1e6 GeV) on the
    10-string x 20-module x 16-PMT array against direct propagation of their
    full Cherenkov light yield."""
    from neutrinotelescopes_amd.surrogate import surrogate_study
    r = surrogate_study()
    keep = ("config", "table_photons", "table_seconds", "train_steps", "train_step_ms", "holdout_loss_initial",
            "holdout_loss_final", "error_pairs", "count_err_median", "count_err_mean", "count_bias",
            "error_pairs_50hits", "count_err_median_50hits", "count_err_mean_50hits", "count_bias_50hits",
            "count_err_median_untrained", "count_err_median_scale_only", "count_err_median_50hits_scale_only",
            "truth_rel_stat_err_median", "truth_rel_stat_err_median_50hits", "truth_rel_stat_err_weighted_median",
            "truth_rel_stat_err_weighted_median_50hits", "count_err_rms", "count_err_rms_intrinsic",
            "count_err_rms_50hits", "count_err_rms_intrinsic_50hits", "time_kl_median", "time_kl_median_untrained",
            "events",
            "llh_seconds", "llh_flow_seconds", "llh_fallback_seconds", "llh_pairs", "llh_hits", "fallback_pairs",
            "surrogate_events_per_s", "direct_llh_events", "direct_llh_photons_per_event", "direct_llh_seconds",
            "direct_events_per_s", "direct_llh_finite", "speedup", "direct_photons_per_s")
    return {k: r[k] for k in keep if k in r}


def flow_kernels_benchmark(steps: int = 2000, llh_reps: int = 10):
    """Flow kernels on synthetic data: `steps` Adam steps of the conditional
    RQ-spline flow on batches of 16384 rows (80 % time rows, 20 % count rows,
    random network weights) and the event likelihood of 1024 cascades on the
    10-string x 20-module x 16-PMT array (3 276 800 pairs, Poisson(2)
    observed hits per PMT), `llh_reps` times after 3 warm-up evaluations."""
    import torch

    from neutrinotelescopes_amd import surrogate as S
    from neutrinotelescopes_amd.flow import FlowModel, FlowRows
    rng = np.random.default_rng(0)
    n = 1 << 20
    x = rng.uniform(-1, 1, (n, 5)).astype(np.float32)
    t = rng.normal(10 + 20 * x[:, 0], 3 + 2 * (x[:, 1] + 1)).astype(np.float32)
    cnt = rng.poisson(np.exp(1 + x[:, 2])).astype(np.float32)
    flags = np.where(rng.random(n) < 0.8, 1, 2).astype(np.uint8)
    m = FlowModel(0, seed=1)
    rows = FlowRows.from_numpy(x, t, cnt, flags)
    batch = 16384
    gen = torch.Generator(device=m.dev).manual_seed(0)
    batches = [rows.slice(torch.randint(0, n, (batch,), generator=gen, device=m.dev)) for _ in range(16)]
    for k in range(20):
        m.train_step(batches[k % 16], 1e-3)
    torch.cuda.synchronize(m.dev)
    t0 = time.perf_counter()
    for k in range(steps):
        m.train_step(batches[k % 16], 1e-3)
    torch.cuda.synchronize(m.dev)
    step_ms = (time.perf_counter() - t0) / steps * 1e3
    modules = S.array_detector()
    casc = S.random_cascades(1024, rng)
    counts = rng.poisson(2.0, (1024, len(modules), 16))
    k = np.repeat(np.arange(counts.size), counts.reshape(-1))
    hits = np.column_stack([k // (len(modules) * 16), (k // 16) % len(modules), k % 16,
                            rng.uniform(-10, 300, k.size)])
    f = S.pair_features(S.Cascades(casc.pos[:4], casc.dir[:4], casc.energy[:4]), modules).reshape(-1, 5)
    sur = S.Surrogate(m, S.FeatureNorm(f.mean(0), f.std(0)),
                      S.CountScale(np.linspace(0, 6, 9), np.linspace(-10, -20, 9)))
    d = S.to_device(casc, modules, S.Observation(counts, hits), str(m.dev))
    for _ in range(3):
        S.surrogate_llh(sur, d)
    torch.cuda.synchronize(m.dev)
    t0 = time.perf_counter()
    for _ in range(llh_reps):
        S.surrogate_llh(sur, d)
    torch.cuda.synchronize(m.dev)
    llh_ms = (time.perf_counter() - t0) / llh_reps * 1e3
    m.close()
    return {"train_steps": steps, "batch": batch, "train_step_ms": step_ms, "llh_events": 1024,
            "llh_pairs": int(1024 * len(modules) * 16), "llh_hits": int(d.hit_t.shape[0]), "llh_ms": llh_ms}


def pmt_event_hits(events: int, rng):
    """Detected hits of `events` events on the 10-string × 20-module × 16-PMT
    array (3200 channels each; module ids 200·e + m so every event's channels
    are distinct): per channel Poisson(λ) hits, λ log-uniform in [0.1, 300]
    (near and far modules), times a prompt peak (N(t_p, 1.5 ns), 60 %) plus a
    scattered tail (t_p + Exp(25 ns)), t_p uniform in [20, 200] ns."""
    n_ch = events * 200 * 16
    lam = np.exp(rng.uniform(np.log(0.1), np.log(300.0), n_ch))
    k = rng.poisson(lam)
    ch = np.repeat(np.arange(n_ch), k)
    tp = rng.uniform(20, 200, n_ch)[ch]
    prompt = rng.random(ch.size) < 0.6
    t = np.where(prompt, tp + rng.normal(0, 1.5, ch.size), tp + rng.exponential(25.0, ch.size))
    return {"module_id": (ch // 16).astype(np.uint16), "pmt_id": (ch % 16 + 1).astype(np.uint8),
            "time": t.astype(np.float32), "source_index": np.zeros(ch.size, np.uint32),
            "photon_id": np.arange(ch.size, dtype=np.uint64)}


def pmt_frontend_benchmark(events: int = 16, reps: int = 10, cpu_target_s: float = 10.0, cpu: bool = True):
    """PMT front end (make_reco_pulses for every channel in one call): `events`
    events of pmt_event_hits on the device, `reps` timed calls after 2 warm-up
    calls (HIP events on the launch stream), and the CPU oracle (hit draws and
    the per-channel chain, FP64, one channel per thread over all host threads)
    on a bounded sample of the same channels."""
    import torch
    from concurrent.futures import ThreadPoolExecutor

    from neutrinotelescopes_amd import pmt_frontend as pf
    sys.path.insert(0, str(ROOT / "tests"))
    import _oracle
    rng = np.random.default_rng(0)
    h = pmt_event_hits(events, rng)
    n = h["time"].size
    fe = pf.PMTFrontEnd(pf.std_pmt_config(), 0)
    dev = torch.device("cuda:0")
    up = lambda a: torch.from_numpy(a.view({np.uint16: np.int16, np.uint32: np.int32, np.uint64: np.int64}.get(
        a.dtype.type, a.dtype))).to(dev)
    cols = {k: up(v) for k, v in h.items()}
    cap = events * 200 * 17
    for _ in range(2):
        out, n_ch, _ = fe.make_reco_pulses_device(cols, n, seed=1, capacity=cap)
    torch.cuda.synchronize(dev)
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for r in range(reps):
        out, n_ch, _ = fe.make_reco_pulses_device(cols, n, seed=1 + r, capacity=cap)
    e1.record(stream)
    torch.cuda.synchronize(dev)
    ms = e0.elapsed_time(e1) / reps
    channels = int(n_ch.item())
    pulses = int(out["n_pulses"][:channels].sum().item())
    st = int((out["status"][:channels] != 0).sum().item())
    fe.close()
    res = {"config": f"{events} events x 3200 channels (10 strings x 20 modules x 16 PMTs), Poisson hits "
                     f"log-uniform 0.1-300 per channel; STD_PMT_CONFIG (601-sample waveform, 76 ADC samples, "
                     f"3001-column FNNLS)",
           "hits": n, "channels": channels, "pulses": pulses, "channels_with_status": st, "ms_per_call": ms,
           "channels_per_s": channels / (ms * 1e-3), "events_per_s": events / (ms * 1e-3)}
    if not cpu:
        return res
    # CPU oracle on a sample of the same channels (all host threads, one channel each)
    cfg = pf.std_pmt_config()
    ca, keep = cfg.to_abi()
    ch_all = h["module_id"].astype(np.int64) * 17 + h["pmt_id"]
    order = np.argsort(ch_all, kind="stable")
    bounds = np.r_[0, np.flatnonzero(np.diff(ch_all[order])) + 1, order.size]
    threads = host_cpu_info()["cpu_share"]

    def one(c):
        idx = order[bounds[c]:bounds[c + 1]]
        tp, q, _ = _oracle.pmt_hit_draws(ca, h["time"][idx], h["source_index"][idx], h["photon_id"][idx], 1)
        _oracle.pmt_channel(ca, tp, q, int(ch_all[idx[0]]), 1)

    sample = rng.permutation(len(bounds) - 1)
    done, t0 = 0, time.perf_counter()
    with ThreadPoolExecutor(threads) as pool:
        while time.perf_counter() - t0 < cpu_target_s and done < sample.size:
            chunk = sample[done:done + 4 * threads]
            list(pool.map(one, chunk))
            done += chunk.size
    cpu_dt = time.perf_counter() - t0
    cpu_rate = done / cpu_dt
    return dict(res, cpu_oracle_channels_per_s=cpu_rate, cpu_threads=threads,
                cpu_sample=f"{done} channels in {cpu_dt:.1f} s on {threads} threads (FP64 C oracle)",
                gpu_vs_cpu=channels / (ms * 1e-3) / cpu_rate)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cascade", choices=["cascade", "cluster", "isotropic"])
    ap.add_argument("--photons", type=int, default=10_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the FP32-arithmetic, hit-epilogue and 1e9-photon isotropic legs")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving (p50/p99) measurement")
    ap.add_argument("--serving-only", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the flow-surrogate study (tables, training, error, likelihood speed-up)")
    ap.add_argument("--surrogate-only", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device-to-device copy 5 times (reference for "
                         "FETCH_SIZE/WRITE_SIZE counter scaling)")
    ap.add_argument("--direction-arithmetic", default="fp64", choices=["fp64", "fp32"],
                    help="direction sampling and rotations in FP64 (default) or in the reference's FP32 arithmetic")
    ap.add_argument("--no-culling", action="store_true",
                    help="test every sphere on every step instead of string bounds first")
    ap.add_argument("--flow-kernels", action="store_true",
                    help="only time the flow kernels on synthetic data (training steps, event likelihood)")
    ap.add_argument("--no-pmt-frontend", action="store_true", help="skip the PMT front-end measurement")
    ap.add_argument("--pmt-frontend-only", action="store_true",
                    help="only time the PMT front end (waveforms and unfolding of 16 array events)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU per process; multi-GPU runs are not "
                                   "part of this build (--gpus must be 1)"}))
        sys.exit(2)

    import torch
    if args.serving_only:
        srv = serving_benchmark()
        print(json.dumps({"serving": srv, "serving_rounds": serving_benchmark("rounds", srv["arrival_rate_hz"])}),
              flush=True)
        return
    if args.surrogate_only:
        print(json.dumps({"surrogate": surrogate_benchmark()}), flush=True)
        return
    if args.flow_kernels:
        print(json.dumps({"flow_kernels": flow_kernels_benchmark()}), flush=True)
        return
    if args.pmt_frontend_only:
        print(json.dumps({"pmt_frontend": pmt_frontend_benchmark(cpu=not args.no_cpu_baseline)}), flush=True)
        return
    if args.copy_calibration:
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": a.numel() * 4, "dispatches": 5}))
        return
    leg = measure_kernel(args.workload, args.photons, args.direction_arithmetic, args.steps, args.warmup,
                         culling=not args.no_culling)
    value = leg["value"]
    progress(f"headline {WORKLOAD_NAMES[args.workload]}: {value:.4g} photons/s")
    out = {
        "metric": "photons propagated/s (1 GPU)",
        "value": value,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": leg["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.direction_arithmetic == "fp64" else "f32",
        "data": "synthetic",
        "config": dict(leg["config"], workload=WORKLOAD_NAMES[args.workload]),
        "roofline": dict(leg["roofline"], traffic=profiled_traffic(args)),
    }
    if args.workload == "cascade" and not args.no_legs:
        out["legs"] = other_legs(args)
        progress("legs done")
    if not args.no_serving and args.workload == "cascade":
        out["serving"] = serving_benchmark()
        out["serving_rounds"] = serving_benchmark("rounds", out["serving"]["arrival_rate_hz"])
        progress("serving done")
    if not args.no_surrogate and args.workload == "cascade":
        out["surrogate"] = surrogate_benchmark()
        progress("surrogate done")
    if not args.no_pmt_frontend and args.workload == "cascade":
        out["pmt_frontend"] = pmt_frontend_benchmark()
        progress("pmt front end done")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload)
        out["gpu_vs_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
