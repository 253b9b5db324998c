"""Benchmark: scattering events/s of the fused forward kernel on one MI355X.

Default workload: the largest single-GPU configuration, a 3-D acoustic
medium (exponential ACF, lc = 0.05, f = 40, cv = 0.1, Q = 300), cartesian
frame, point source, 1e9 particles, 121^3 cells of 0.1 x 8 direction bins x
501 times (t = 0:0.02:10): a 28.4 GB uint32 histogram resident in HBM, FP64
flight (the reference's precision).  One step = one full forward solve
(histogram cleared, every particle born, propagated through all observation
times and deposited, the time-major scratch added into the histogram),
material tables already in HBM.

Prints one JSON line (the repository's benchmark contract) with:
  roofline      the tracing kernel against HBM (8 algorithmic bytes per
                histogram deposit, a 4-byte read and write of the counter),
                kernel-only; `step_level` adds the zero-fill and final write
                of the histogram (+2 H) over the whole step; `traffic` is the
                kernel's measured HBM bytes per launch (rocprofv3 FETCH_SIZE /
                WRITE_SIZE, committed under profiles/)
  cpu_baseline  the oracle (the reference's sub-step algorithm restated in C,
                FP64) on a bounded sample of the same workload, best of 3, on
                this host's CPU share, with the reference-compatible
                two-variable (x, y) marginal histogram
  fp32_flight   the same workload with FP32 flight arithmetic (FP64 time)
  config2_elastic2d, batched_fwd_adjoint_vs_plain_fd   the 2-D elastic
                workload and the 64-material forward + adjoint comparison

`--config N` times another workload alone (1: the README case with its LDS
histogram and an FP64-VALU roofline; 2: 2-D elastic; 3: the elastic
half-space, FP64-VALU roofline; 4: the 64-material batch).
"""
import argparse
import copy
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
BYTES_PER_DEPOSIT = 8
FLOP_PER_EVENT = {(3, True): 175, (2, True): 117, (3, False): 177, (2, False): 119}
FLOP_PER_DEPOSIT = {"radial": 24, "cartesian": 16}
FLOP_PER_BOUNDARY = 30
# committed rocprofv3 FETCH_SIZE / WRITE_SIZE summaries (bytes per launch of the tracing kernel)
TRAFFIC = {(5, "fp64"): "k_forward_acoustic3d_1e9_fp64_traffic.json",
           (5, "fp64_lds"): "k_forward_acoustic3d_1e9_fp64_lds_traffic.json",
           (2, "fp64"): "k_forward_elastic2d_1e7_traffic.json",
           (3, "fp64"): "k_forward_elastic3d_halfspace_5e7_traffic.json"}
FP32_PER_BIN = ("FP32 flight vs FP64 on a 31^3 x 8 x 101 version of this workload, 1e7 particles each: "
                "0.27 % of 875 279 populated bins beyond 3 sigma, max |z| 5.0, mean z^2 1.002 "
                "(tests/test_gpu_forward.py::test_fp32_flight_per_bin_error_against_fp64)")


LDS_TOL = 1e-5
LDS_PER_BIN = ("compressed tables (tol 1e-5 rad) vs full tables on a 31^3 x 8 x 101 version of this workload, 1e7 "
               "particles each (tests/test_gpu_lds_table.py::test_lds_table_per_bin_error_against_full_table)")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, choices=[1, 2, 3, 4, 5], default=5)
    ap.add_argument("--chunks", type=int, default=0,
                    help="forward calls per step, each on its own particle range (0 = one call)")
    ap.add_argument("--batch-scale", type=float, default=1.0,
                    help="config 4: multiply every material's particle count (1 = the configuration's own)")
    ap.add_argument("--autotune", action="store_true", help="config 4: tune the batched launch first")
    ap.add_argument("--mask-counts", type=int, default=16,
                    help="config 4: bins enter the timed loss with E_target > 0 and at least this many counts "
                         "(log E must stay defined in the perturbed solves of the finite differences)")
    ap.add_argument("--dump-gradients", default="", metavar="NPZ",
                    help="config 4: write the per-material LR and FD gradients to this .npz")
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload only: skip FP32 flight, the 2-D elastic and the 64-material runs")
    ap.add_argument("--particles", type=int, default=0, help="0 = the configuration's own count")
    ap.add_argument("--cpu-seconds", type=float, default=8.0,
                    help="target length of one CPU baseline run (three runs are timed)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--ppt", type=int, default=0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--atomic-probe", action="store_true",
                    help="only measure the global-atomic request rate (rtmc_probe_atomics) over several ranges")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                    help="flight arithmetic of the fused kernel (fp32: FP32 flight, FP64 time)")
    ap.add_argument("--lds-tol", type=float, default=0.0,
                    help="> 0: draw scattering angles from compressed tables held in LDS, within this many "
                         "radians of the full tables (0: the full tables in global memory)")
    ap.add_argument("--lds-max-knots", type=int, default=1024)
    return ap.parse_args()


def traffic_from_profiles(config, precision, key="bytes_per_launch"):
    """A figure of the tracing kernel's committed counter summary (HBM bytes per
    launch by default), if any."""
    name = TRAFFIC.get((config, precision))
    p = os.path.join(ROOT, "profiles", name) if name else None
    if p and os.path.exists(p):
        with open(p) as f:
            return json.load(f).get(key)
    return None


def probe_atomics(E):
    """The request rate of no-return global atomics (rtmc_probe_atomics) at random
    words and random aligned word pairs over the histogram's own buffer (its
    contents are overwritten): the ceiling of the time-major deposits."""
    import ctypes as C
    import torch
    from radiativetransfermontecarlo_amd import _abi
    lib = _abi.load()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {"bytes": int(E.numel()) * 4}
    for kind, name in ((0, "u32_random_per_s"), (1, "u64_random_pairs_per_s")):
        r = C.c_double()
        _abi.check(lib.rtmc_probe_atomics(C.c_void_p(E.data_ptr()), int(E.numel()), kind, 2000, C.byref(r), stream),
                   "rtmc_probe_atomics")
        out[name] = r.value
    return out


WORKLOADS = {
    1: "3-D acoustic, spherical, 1e6 particles, exp ACF lc=0.1, f=10, x=0:0.05:6, t=0:0.02:5 "
       "(config 1, the reference's README case)",
    2: "2-D elastic P-SV, cartesian, 1e7 particles, Q=500, x,y=-50:0.5:50, 36 direction bins, t=0:0.05:20 "
       "(config 2)",
    3: "3-D elastic half-space with free surface at z=0, cylindrical, 5e7 particles, von Karman nu=0.5 lc=0.2 f=20, "
       "r=0:0.1:30, 90 direction bins, t=0:0.05:15 (config 3)",
    5: "3-D acoustic, cartesian, 1e9 particles, exp ACF lc=0.05 f=40 cv=0.1 Q=300, 121^3 cells of 0.1, "
       "8 direction bins, t=0:0.02:10 (config 5)",
}
DEFAULT_N = {1: 1_000_000, 2: 10_000_000, 3: 50_000_000, 5: 1_000_000_000}


def measure(config, precision, steps, warmup, particles=0, chunks=1, block=0, ppt=0, quiet=False, lds_tol=0.0,
            lds_max_knots=1024):
    """Times `steps` full forward solves of a workload; returns the figures and
    the problem (for the CPU baseline)."""
    import ctypes as C
    import numpy as np
    import torch
    from radiativetransfermontecarlo_amd import _abi, solver
    from radiativetransfermontecarlo_amd.material import MaterialClass
    from radiativetransfermontecarlo_amd.observation import initializeObservation
    from tests import cases

    dev = torch.device("cuda", 0)
    N = particles or DEFAULT_N[config]
    geometry, source, material, observation = getattr(cases, f"config{config}")(n=N)
    workload = WORKLOADS[config]
    if N != DEFAULT_N[config]:
        workload += f" [{N} particles]"
    if chunks > 1:
        workload += f", {chunks} calls of {N // chunks} particles"
    d = geometry['dimension']
    material = MaterialClass.prepareSigma(material, d)
    O = initializeObservation(geometry, bool(material.acoustics), observation, N, allow_nd=True)
    dmat = solver.DeviceMaterial(material, d, dev, with_zoeppritz=len(geometry.get('bnd', [])) > 0, lds_tol=lds_tol,
                                 lds_max_knots=lds_max_knots)
    if lds_tol > 0:
        workload += (f", angles from compressed tables in LDS ({dmat.c.lds[0].m} knots, "
                     f"within {dmat.c.lds[0].tol:g} rad of the full table)")
    dbins = solver.DeviceBins(O, dev)
    bounds = [N * k // chunks for k in range(chunks + 1)]
    E = torch.zeros(O.size, dtype=torch.int32, device=dev)
    cnt = torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)
    lib = _abi.load()

    def step(seed):
        E.zero_()
        cnt.zero_()
        for k in range(chunks):
            solver.forward_histogram(geometry, source, material, O, bounds[k + 1] - bounds[k], seed=seed,
                                     device=dev, E=E, counters=cnt, dmat=dmat, block_size=block,
                                     particles_per_thread=ppt, precision=precision, particle_offset=bounds[k],
                                     bins=dbins)

    tag = f"config {config} {precision}"
    for w in range(warmup):
        step(1000 + w)
        torch.cuda.synchronize()
        print(f"{tag}: warmup step {w} done", file=sys.stderr, flush=True)
    _abi.check(lib.rtmc_set_trace_timing(1), "rtmc_set_trace_timing")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        step(s)
        if config in (3, 5):  # long steps: a progress line each
            torch.cuda.synchronize()
            print(f"{tag}: step {s} done at {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the tracing kernel alone: HIP events recorded by the library around its launch on the launch stream
    tk = (C.c_double * 1024)()
    nk = C.c_int32()
    _abi.check(lib.rtmc_trace_times(tk, 1024, C.byref(nk)), "rtmc_trace_times")
    _abi.check(lib.rtmc_set_trace_timing(0), "rtmc_set_trace_timing")
    c = cnt.cpu().numpy()
    out = {"config": config, "precision": precision, "workload": workload, "particles": N,
           "histogram_bytes": int(O.size) * 4, "ms_per_step": 1000.0 * wall / steps,
           "kernel_ms": float(np.sum([tk[i] for i in range(nk.value)])) / steps,
           "scatter_events": int(c[_abi.CNT_SCATTER]), "deposits": int(c[_abi.CNT_DEPOSIT]),
           "boundary_hits": int(c[_abi.CNT_BOUNDARY]), "removed": int(c[_abi.CNT_KILLED]),
           "dim": d, "acoustics": bool(material.acoustics), "frame": geometry.get('frame', 'spherical')}
    out["value"] = out["scatter_events"] / (out["ms_per_step"] * 1e-3)
    prob = (geometry, source, material, observation, O)
    if config in (2, 3, 5):
        out["probe"] = probe_atomics(E)
    del E
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out, prob


def atomic_view(m, config, key):
    """The histogram atomics' request rate against the probed ceiling (None
    without a committed WRITE_SIZE summary or a probe)."""
    wb = traffic_from_profiles(config, key, "write_bytes")
    if not wb or not m.get("probe"):
        return None
    req = wb / 32.0 / (m["kernel_ms"] * 1e-3)
    pr = m["probe"]
    return {"requests_per_launch": wb / 32.0, "requests_per_s": req,
            "probe_u32_random_per_s": pr["u32_random_per_s"],
            "probe_u64_random_pairs_per_s": pr["u64_random_pairs_per_s"], "probe_bytes": pr["bytes"],
            "frac_of_u32_probe": req / pr["u32_random_per_s"],
            "definition": "32-byte write requests of the histogram atomics (committed WRITE_SIZE summary) over the "
                          "kernel's time, against rtmc_probe_atomics over the histogram's buffer in the same run"}


def roofline(m, config, precision, variant=""):
    """Dominant kernel against its bound; HBM (histogram deposits) or FP64 VALU.
    variant: suffix of the traffic summary's key ("_lds": compressed angle tables)."""
    if config in (2, 5):
        alg = BYTES_PER_DEPOSIT * m["deposits"]
        k_gbs = alg / (m["kernel_ms"] * 1e-3) / 1e9
        s_bytes = alg + 2 * m["histogram_bytes"]
        s_gbs = s_bytes / (m["ms_per_step"] * 1e-3) / 1e9
        atomic = atomic_view(m, config, precision + variant)
        return {"bound": "hbm", "achieved": k_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": k_gbs / HBM_PEAK_GBS, "traffic": traffic_from_profiles(config, precision + variant),
                "atomic_view": atomic,
                "kernel": "k_trace", "kernel_ms": m["kernel_ms"],
                "algorithmic_unit": "kernel-only: 8 B per histogram deposit (4 B read + 4 B write of the counter) "
                                    "x deposits per launch, over the kernel's duration",
                "step_level": {"achieved": s_gbs, "frac": s_gbs / HBM_PEAK_GBS, "bytes": s_bytes,
                               "unit": "GB/s",
                               "definition": "(8 B x deposits + 2 x histogram bytes: zero-fill and final write) "
                                             "over the whole step"}}
    key = (m["dim"], m["acoustics"])
    fobs = FLOP_PER_DEPOSIT["radial" if m["frame"] != "cartesian" else "cartesian"]
    flops = FLOP_PER_EVENT[key] * m["scatter_events"] + fobs * m["deposits"] + FLOP_PER_BOUNDARY * m["boundary_hits"]
    tf = flops / (m["kernel_ms"] * 1e-3) / 1e12
    peak = FP64_PEAK_TFLOPS if precision == "fp64" else FP32_PEAK_TFLOPS
    out = {"bound": "fp64-valu" if precision == "fp64" else "fp32-valu", "achieved": tf, "peak": peak,
           "unit": "TFLOP/s", "frac": tf / peak, "traffic": traffic_from_profiles(config, precision + variant),
           "kernel": "k_trace", "kernel_ms": m["kernel_ms"],
           "algorithmic_unit": f"{FLOP_PER_EVENT[key]} flop-eq per scattering event + {fobs} per deposit + "
                               f"{FLOP_PER_BOUNDARY} per boundary hit"}
    if m["histogram_bytes"] > (64 << 20) or config == 3:
        alg = BYTES_PER_DEPOSIT * m["deposits"]
        out["hbm_view"] = {"achieved": alg / (m["kernel_ms"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": alg / (m["kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS}
        out["atomic_view"] = atomic_view(m, config, precision + variant)
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, m, prob):
    """The oracle on a bounded sample of the same workload, best of three runs,
    with the reference-compatible (x, y) marginal when the workload bins more
    than two variables (initializeObservation.m:93-95)."""
    import numpy as np
    from oracle import oracle as orc
    from radiativetransfermontecarlo_amd.material import MaterialClass
    from radiativetransfermontecarlo_amd.observation import initializeObservation
    geometry, source, material, observation, O = prob
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS") or avail)
    obs = dict(observation)
    marginal = ""
    if len(O.ibins) > 2:
        for k in ('z', 'directions'):
            obs.pop(k, None)
        marginal = ", histogram = the (x, y) marginal (the reference bins at most two variables)"
    O2 = initializeObservation(geometry, bool(material.acoustics), obs, 1)
    # preparation as the reference does it on every call (MaterialClass.prepareSigma)
    t0 = time.perf_counter()
    MaterialClass.prepareSigma(copy.deepcopy(material), geometry['dimension'])
    prep_s = time.perf_counter() - t0
    pk = 2000
    n_cal = pk * threads
    t0 = time.perf_counter()
    orc.forward(geometry, source, material, O2, n_cal, seed=99, packet=pk, threads=threads)
    t_cal = time.perf_counter() - t0
    n = int(max(n_cal, n_cal * args.cpu_seconds / max(t_cal, 1e-3)))
    n = max(n_cal, (n // n_cal) * n_cal)
    best = None
    for r in range(3):
        t0 = time.perf_counter()
        _E, dep, nscat = orc.forward(geometry, source, material, O2, n, seed=98 + r, packet=pk, threads=threads,
                                     return_counters=True)
        t = time.perf_counter() - t0
        if best is None or t < best[0]:
            best = (t, nscat, dep)
        print(f"cpu baseline run {r} done: {t:.1f} s", file=sys.stderr, flush=True)
    t, nscat, dep = best
    full_t = t * m["particles"] / n
    return {"value": nscat / t, "unit": "scattering events/s", "cores": threads, "kind": "port",
            "sample": f"{n} particles of the same workload (all {O.Nt} observation times, best of 3 runs of "
                      f"{t:.1f}-s class), the sub-step loop of propagateParticleSmallDt restated in C (FP64, "
                      f"OpenMP over packets of {pk}){marginal}; {dep / t:.3e} deposits/s",
            "value_prep_included": nscat / t * full_t / (full_t + prep_s),
            "prep_s": prep_s, "cpu_model": _cpu_model(), "host_cpus_visible": os.cpu_count(),
            "cpus_allowed": avail, "threads_note": "OMP_NUM_THREADS (the host's CPU share for one GPU) threads"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU only: sharding particles across GPUs (by "
                                   "particle_offset, histograms summed) is not implemented", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    if args.config == 4:
        return run_batch_adjoint(args)
    if args.atomic_probe:
        return run_atomic_probe()
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4}))
        return
    m, prob = measure(args.config, args.precision, args.steps, args.warmup, args.particles, max(1, args.chunks),
                      args.block, args.ppt, lds_tol=args.lds_tol, lds_max_knots=args.lds_max_knots)
    out = {
        "metric": "scattering events/s on 1 MI355X",
        "value": m["value"],
        "unit": "scattering events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 flight / f64 time",
        "data": "synthetic (Philox4x32-10 particles; material tables from the reference's formulas)",
        "config": {"workload": m["workload"], "particles": m["particles"], "histogram_bytes": m["histogram_bytes"],
                   "deposits_per_step": m["deposits"], "scatter_events_per_step": m["scatter_events"],
                   "deposits_per_s": m["deposits"] / (m["ms_per_step"] * 1e-3)},
        "roofline": roofline(m, args.config, args.precision, "_lds" if args.lds_tol else ""),
    }
    extras = args.config == 5 and not args.no_extras and not args.particles
    if extras and args.precision == "fp64":
        f, _ = measure(5, "fp32", args.steps, args.warmup)
        out["fp32_flight"] = {"value": f["value"], "unit": "scattering events/s", "ms_per_step": f["ms_per_step"],
                              "kernel_ms": f["kernel_ms"], "deposits_per_step": f["deposits"],
                              "scatter_events_per_step": f["scatter_events"], "roofline": roofline(f, 5, "fp32"),
                              "per_bin_error_vs_fp64": FP32_PER_BIN}
    if extras and args.precision == "fp64" and not args.lds_tol:
        t, _ = measure(5, "fp64", args.steps, args.warmup, lds_tol=LDS_TOL)
        out["lds_angle_table"] = {"workload": t["workload"], "value": t["value"], "unit": "scattering events/s",
                                  "ms_per_step": t["ms_per_step"], "kernel_ms": t["kernel_ms"],
                                  "deposits_per_step": t["deposits"], "scatter_events_per_step": t["scatter_events"],
                                  "roofline": roofline(t, 5, "fp64", "_lds"),
                                  "per_bin_error_vs_full_tables": LDS_PER_BIN}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, m, prob)
    if extras:
        c2, _ = measure(2, "fp64", max(args.steps, 5), 1)
        out["config2_elastic2d"] = {"workload": c2["workload"], "value": c2["value"], "unit": "scattering events/s",
                                    "ms_per_step": c2["ms_per_step"], "kernel_ms": c2["kernel_ms"],
                                    "deposits_per_step": c2["deposits"], "roofline": roofline(c2, 2, "fp64")}
        b = run_batch_adjoint(args, emit=False)
        out["batched_fwd_adjoint_vs_plain_fd"] = {
            "workload": b["config"]["workload"], "speedup_gpu_only": b["value"],
            "speedup_kernels_only": b["kernel_only_speedup"], "speedup_wall": b["wall_speedup"],
            "speedup_wall_excluding_preparation": b["prep_excluded_speedup"], "loss": b["config"]["loss"],
            "batched": b["batched"], "plain_fd": b["plain_fd"], "gradient_agreement": b["gradient_agreement"]}
    print(json.dumps(out), flush=True)


def run_atomic_probe():
    """Request rate of no-return global atomics (the ceiling of global-memory deposits)."""
    import ctypes as C
    import torch
    from radiativetransfermontecarlo_amd import _abi
    lib = _abi.load()
    words_max = 1_155_000_000  # the 4.6 GB histogram of the 2-D elastic workload
    buf = torch.zeros(words_max, dtype=torch.int32, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rows = []
    for kind, name in enumerate(("u32 random", "u64 random pairs", "u32 runs of 2", "u32 runs of 8",
                                 "u64 random pairs, returning")):
        for words in (1 << 20, 1 << 25, 1 << 28, words_max):
            r = C.c_double()
            _abi.check(lib.rtmc_probe_atomics(C.c_void_p(buf.data_ptr()), words, kind, 2000, C.byref(r), stream),
                       "rtmc_probe_atomics")
            rows.append({"kind": name, "bytes": words * 4, "requests_per_s": r.value})
    print(json.dumps({"atomic_probe": rows}), flush=True)


def run_batch_adjoint(args, emit=True):
    """64 acoustic materials: batched forward + adjoint against 64 plain forward
    solves with central finite differences (7 solves per material).  Both
    sides prepare their materials natively in one batched call (the plain side
    all 448 perturbed materials at once); the reported value is the ratio of
    GPU time."""
    import ctypes as C
    import numpy as np
    import torch
    from radiativetransfermontecarlo_amd import _abi
    from radiativetransfermontecarlo_amd import batch as rb
    from radiativetransfermontecarlo_amd import native_material as nm
    from radiativetransfermontecarlo_amd import solver
    from radiativetransfermontecarlo_amd.observation import initializeObservation
    from tests import cases

    dev = torch.device("cuda", 0)
    geometry, source, base, observation, params, Ns = cases.config4()
    Ns = np.maximum(1000, np.round(Ns * args.batch_scale)).astype(np.int64)
    n = len(Ns)
    O = initializeObservation(geometry, True, observation, 1)
    KT = 0.0  # tables that follow sigma: the score differentiates sigma
    H = np.array([0.02, 0.02, 0.05])  # central-difference steps (relative)
    scales = rb.count_scales(O, [int(N) for N in Ns], dev)
    lib = _abi.load()

    def sync():
        torch.cuda.synchronize()

    # target: perturbed parameters, independent particles (not timed)
    pt = params * np.array([1.05, 0.95, 1.10])
    LT = dict(lds_tol=args.lds_tol, lds_max_knots=args.lds_max_knots)  # compressed angle tables (0: none)
    Bt = rb.Batch.from_params(geometry, O, base, pt, source, list(Ns), [7_000 + i for i in range(n)], device=dev,
                              knot_threshold=KT, **LT)
    Bt.forward()
    target = Bt.energy_density()

    def loss_and_weights(E, counts=args.mask_counts):
        """loss per item and dL/dcount (the backward of the loss through the normalisation)"""
        dens = E.to(torch.float64) * scales
        mask = (target > 0) & (E >= counts)
        lg = torch.where(mask, torch.log(torch.where(mask, dens, 1.0)) - torch.log(torch.where(mask, target, 1.0)),
                         0.0)
        return (lg * lg).sum(dim=1), torch.where(mask, 2.0 * lg / torch.where(mask, dens, 1.0), 0.0) * scales, mask

    loss_and_weights(Bt.E)  # loads the elementwise kernels outside the timed regions
    sync()
    del Bt
    seeds = [1_000 + i for i in range(n)]

    # ---- batched forward + adjoint -------------------------------------
    sync()
    t0 = time.perf_counter()
    B = rb.Batch.from_params(geometry, O, base, params, source, list(Ns), seeds, device=dev, knot_threshold=KT,
                             **LT)
    tabs = B.derivative_tables()
    t_prep_a = time.perf_counter() - t0
    tuning = B.autotune(0.02) if (args.autotune or not emit) else None
    sync()
    _abi.check(lib.rtmc_set_trace_timing(1), "rtmc_set_trace_timing")
    t1 = time.perf_counter()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record()
    B.forward()
    e1.record()
    loss, w, mask = loss_and_weights(B.E)
    grad = B.adjoint(w, tables=tabs)
    e2.record()
    sync()
    t_a = time.perf_counter() - t1
    g_lr = grad.cpu().numpy()
    ms_fwd, ms_adj = e0.elapsed_time(e1), e1.elapsed_time(e2)
    tk = (C.c_double * 4)()
    nk = C.c_int32()
    _abi.check(lib.rtmc_trace_times(tk, 4, C.byref(nk)), "rtmc_trace_times")
    _abi.check(lib.rtmc_set_trace_timing(0), "rtmc_set_trace_timing")
    k_fwd, k_adj = (tk[0], tk[1]) if nk.value >= 2 else (float("nan"), float("nan"))
    nscat = int(B.counters[:, 0].sum())
    ndep = int(B.counters[:, 1].sum())
    # the gradient at other loss thresholds (untimed): the same forward, one adjoint each
    thresholds = sorted({4, args.mask_counts, 100})
    g_lr = {args.mask_counts: g_lr}
    masks = {args.mask_counts: mask}
    for c in thresholds:
        if c not in g_lr:
            _l, wc, masks[c] = loss_and_weights(B.E, c)
            g_lr[c] = B.adjoint(wc, tables=tabs).cpu().numpy()
    del B
    torch.cuda.empty_cache()

    # ---- plain forward solves + central differences ---------------------
    # the same native batched preparation: all 7 x 64 perturbed materials in one call
    sync()
    tp = time.perf_counter()
    plist = []
    for i in range(n):
        plist.append(nm.params_for(base, 3, *params[i], knot_threshold=KT, **LT))
        for k in range(3):
            for s in (1.0, -1.0):
                p = params[i].copy()
                p[k] += s * H[k] * params[i, k]
                plist.append(nm.params_for(base, 3, *p, knot_threshold=KT, **LT))
    nat = nm.NativeMaterials(plist)
    t_prep_b = time.perf_counter() - tp
    dbins = solver.DeviceBins(O, dev)
    _abi.check(lib.rtmc_set_trace_timing(1), "rtmc_set_trace_timing")
    g_fd = {c: np.zeros((n, 3)) for c in thresholds}
    gpu_ms_b = 0.0
    sync()
    t2 = time.perf_counter()
    for i in range(n):
        tgt = target[i]
        msk = mask[i]
        sc = scales[i]

        def solve(j):
            nonlocal gpu_ms_b
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            E, _c, _k = solver.forward_histogram(geometry, source, base, O, int(Ns[i]), seed=seeds[i], device=dev,
                                                 dmat=nat.items[7 * i + j], bins=dbins)
            b.record()
            de = E.to(torch.float64) * sc
            r = torch.where(msk, torch.log(torch.where(msk, de, 1.0)) - torch.log(torch.where(msk, tgt, 1.0)), 0.0)
            L = float((r * r).sum())
            gpu_ms_b += a.elapsed_time(b)
            return L, de

        _L0, _d0 = solve(0)
        for k in range(3):
            Lp, dp = solve(1 + 2 * k)
            Lm, dm = solve(2 + 2 * k)
            for c in thresholds:
                if c == args.mask_counts:
                    g_fd[c][i, k] = (Lp - Lm) / (2 * H[k] * params[i, k])
                    continue
                mc = masks[c][i]
                # a bin empty in a perturbed solve has no log: the loss at this threshold is undefined
                if bool(((dp == 0) & mc).any()) or bool(((dm == 0) & mc).any()):
                    g_fd[c][i, k] = math.nan
                    continue
                lp = torch.where(mc, torch.log(torch.where(mc, dp, 1.0)) - torch.log(torch.where(mc, tgt, 1.0)), 0.0)
                lm = torch.where(mc, torch.log(torch.where(mc, dm, 1.0)) - torch.log(torch.where(mc, tgt, 1.0)), 0.0)
                g_fd[c][i, k] = float((lp * lp).sum() - (lm * lm).sum()) / (2 * H[k] * params[i, k])
    sync()
    t_b = time.perf_counter() - t2
    tkb = (C.c_double * 1024)()
    _abi.check(lib.rtmc_trace_times(tkb, 1024, C.byref(nk)), "rtmc_trace_times")
    _abi.check(lib.rtmc_set_trace_timing(0), "rtmc_set_trace_timing")
    kernel_ms_b = float(sum(tkb[i] for i in range(nk.value)))
    nat.close()

    if args.dump_gradients:
        np.savez(args.dump_gradients, params=params, Ns=Ns,
                 **{f"lr_{c}": g_lr[c] for c in thresholds}, **{f"fd_{c}": g_fd[c] for c in thresholds})

    def agreement(lr, fd):
        ok = np.all(np.isfinite(fd), axis=1)
        rel = np.abs(lr[ok] - fd[ok]) / np.maximum(np.abs(fd[ok]), 1e-300)
        out = {"materials_with_defined_fd": int(ok.sum())}
        for k, name in enumerate(("c", "lc", "Q")):
            out[name] = {"median_rel_diff": float(np.median(rel[:, k])) if ok.any() else None,
                         "within_2pct": int(np.sum(rel[:, k] < 0.02)), "within_10pct": int(np.sum(rel[:, k] < 0.10)),
                         "sum_rel_diff": float(abs(lr[ok, k].sum() - fd[ok, k].sum()) / abs(fd[ok, k].sum()))
                         if ok.any() else None}
        return out
    agree = {f"min_counts_{c}": agreement(g_lr[c], g_fd[c]) for c in thresholds}
    gpu_a = ms_fwd + ms_adj
    out = {
        "metric": "batched fwd+adjoint vs plain+FD GPU time on 1 MI355X",
        "value": gpu_ms_b / gpu_a,
        "unit": "x (GPU time of 64 plain solves with central differences / GPU time of the batched forward + "
                "adjoint)",
        "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": gpu_a,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Philox4x32-10 particles; 64 materials from numpy default_rng(20240604))",
        "config": {"workload": "64 3-D acoustic materials, cv=[c c] c~U[.02,.15], lc~U[.05,.5], Q~U[50,1000], "
                               "exp ACF, f=10, README geometry/observation; particles log-uniform 1e5..1e7 "
                               "(config 4)", "particles_total": int(Ns.sum()), "batch_scale": args.batch_scale,
                   "loss": f"sum over bins with E_target > 0 and E >= {args.mask_counts} counts of "
                           "(log E - log E_target)^2 (without a count threshold a bin counted in the forward can be "
                           "empty in a perturbed solve and the finite difference is undefined); target at "
                           "(+5% c, -5% lc, +10% Q), independent seed",
                   "fd_steps_rel": H.tolist(),
                   "tables": "knot_threshold=0" + (f", compressed in LDS within {args.lds_tol:g} rad" if args.lds_tol
                                                   else "")},
        "batched": {"wall_s": t_a + t_prep_a, "prep_s": t_prep_a,
                    "preparation": "rtmc_prepare_materials + rtmc_lc_derivatives (64 materials at once)",
                    "solve_s": t_a, "forward_ms": ms_fwd, "adjoint_ms": ms_adj, "forward_kernel_ms": k_fwd,
                    "adjoint_kernel_ms": k_adj, "scatter_events": nscat, "deposits": ndep,
                    "events_per_s_forward": nscat / (ms_fwd * 1e-3), "tuning": tuning},
        "plain_fd": {"wall_s": t_b + t_prep_b, "prep_s": t_prep_b, "solves": 7 * n, "gpu_ms": gpu_ms_b,
                     "kernel_ms": kernel_ms_b,
                     "preparation": "rtmc_prepare_materials (448 materials at once)"},
        "kernel_only_speedup": kernel_ms_b / (k_fwd + k_adj),
        "wall_speedup": (t_b + t_prep_b) / (t_a + t_prep_a),
        "prep_excluded_speedup": t_b / t_a,
        "gradient_agreement": agree,
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
