"""Benchmark of the radiation step on one MI355X.

Metric: LW field evaluations per second = (particle-step x pixel) pairs / s,
N_pairs = sum_screens n_pix x sum_particles (NP - 1).  One "step" is one full
tg_integrate of the workload (all particles x all pixels x all bins) with the
trajectories already resident in HBM (generated on the device by the Vay
pusher before timing).

Default workload: c2, the configuration the metric is quoted on (1e4
electrons, 4000 steps, 64 x 64 pixel flat screen at 10 m, 1024 bins).  The
mixed-precision kernel gives `value`; the FP64 kernel is timed on the same
workload (fp64_value, mixed_over_fp64) and the mixed result's per-pixel
energy error against it is reported (eps_W_mixed_vs_fp64, bound 1e-5).  c3
(1e5 electrons, 128 x 128 pixels, 2048 bins: ~2.5 min per mixed step) and c4
(32 uneven screens in one launch) run with --workload.
The CPU baseline is the reference's own compiled integrateFieldTrace
(oracle/_ref) driven by an OpenMP pixel loop on the host cores (the oracle
restatement where the reference build is absent), timed on a bounded sample
of the same workload and scaled by pairs.

roofline: the radiation kernel's algorithmic FP operations (106 per pair) per
launch / its mean launch duration from HIP events on the launch stream,
against the vector peak of the precision mix (VALU-bound: no MFMA shape).
traffic: HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE, read
from profiles/traffic_radiation_kernel_<precision>_<workload>.txt when present.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2]
       python bench.py --calibrate-copy BYTES   (a plain copy of BYTES, for
                                                 scaling the PMC byte counters)
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import teufel_amd  # noqa: E402
from teufel_amd import TG_FP64, TG_MIXED  # noqa: E402
from teufel_amd import workloads as W  # noqa: E402
from teufel_amd.metrics import energy_rel_error, trace_rel_error  # noqa: E402

FLOP_PER_PAIR = 106.0          # algorithmic FP ops per (point, pixel): 72 LW + 34 deposit
PEAK_FP64 = 78.6e12            # MI355X FP64 vector
PEAK_FP32 = 157.3e12           # MI355X FP32 vector (published, packed)
FP64_OPS_MIXED = 23.0          # of the 106, done in FP64 by the mixed kernel
METRIC = "LW field evals/s (particle-step\u00d7pixel), mixed vs FP64 GPU at \u22641e-5 rel energy"
KERNEL_NAME = {TG_MIXED: "radiation_kernel<1, true>", TG_FP64: "radiation_kernel<0, true>"}
PREC_NAME = {TG_MIXED: "mixed", TG_FP64: "fp64"}


def mixed_peak():
    return FLOP_PER_PAIR / (FP64_OPS_MIXED / PEAK_FP64 + (FLOP_PER_PAIR - FP64_OPS_MIXED) / PEAK_FP32)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def build_workload(ctx, name, n_particles=None):
    w = W.make(name, n_particles=n_particles)
    u = w.und
    t = time.time()
    ctx.track_undulator(w.bunch.x0, w.bunch.u0, w.dt, w.n_steps, u.bpeak, u.lambda_u, u.n_periods, u.origin,
                        charge=w.bunch.charge, mass=w.bunch.mass)
    log("workload %s: %d particles x %d points tracked on device in %.2f s"
        % (name, w.n_particles, w.n_points, time.time() - t))
    ids = [ctx.add_screen(s.obs, s.t0, s.dt, s.n_bins) for s in w.screens]
    return w, ids


def time_integrate(ctx, prec, steps, warmup, tuning=None):
    for i in range(warmup):
        ctx.integrate(prec, tuning)
        log("  warmup %d/%d: %.3f s" % (i + 1, warmup, ctx.last_timing()[1] / 1e3))
    ctx.synchronize()
    kms = []
    t0 = time.perf_counter()
    for i in range(steps):
        ctx.integrate(prec, tuning)
        kms.append(ctx.last_timing())
        log("  step %d/%d: %.3f s" % (i + 1, steps, kms[-1][1] / 1e3))
    ctx.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return wall, float(np.mean([k for k, _ in kms])) / 1e3, float(np.mean([t for _, t in kms])) / 1e3


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), for the baseline's record."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ctx, w, target_s=15.0, max_particles=64):
    """Reference integrateFieldTrace (compiled from the reference sources) on the
    host cores; falls back to the oracle restatement if the reference build is
    absent.  Returns dict for the JSON line."""
    import oracle
    n_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    s = w.screens[0]
    kind = "reference" if oracle.ref_lib() is not None else "port"
    npart = min(w.n_particles, max_particles)
    d = ctx.get_trajectories(0, npart)
    tr = oracle.Trajectories(**d)
    # calibrate on a small pixel sample, then scale the sample to ~target_s
    stride = max(1, s.n_pix // 64)
    pix = np.arange(0, s.n_pix, stride)
    def run(pix, npart):
        sub = tr.subset(0, npart)
        t = time.perf_counter()
        if kind == "reference":
            rb = oracle.RefBeam(sub)
            rb.integrate(s.obs[pix], s.t0[pix], s.dt, s.n_bins, n_threads=n_threads)
            rb.close()
        else:
            oracle.integrate_screen(sub, s.obs[pix], s.t0[pix], s.dt, s.n_bins, n_threads=n_threads)
        return time.perf_counter() - t
    npc = max(1, min(npart, 4))
    dt0 = run(pix, npc)
    rate0 = len(pix) * npc * (w.n_points - 1) / dt0
    want_pairs = rate0 * target_s
    npc2 = int(min(npart, max(1, want_pairs / (len(pix) * (w.n_points - 1)))))
    if npc2 == npart and want_pairs > len(pix) * npart * (w.n_points - 1):
        stride2 = max(1, int(s.n_pix * npart * (w.n_points - 1) / want_pairs))
        pix = np.arange(0, s.n_pix, stride2)
    dt1 = run(pix, npc2)
    pairs = len(pix) * npc2 * (w.n_points - 1)
    return {"value": pairs / dt1, "unit": "LW field evals/s", "cores": n_threads, "kind": kind,
            "cpu": cpu_model(),
            "sample": "%d particles x %d steps x %d of %d pixels (every %d-th), %d bins, %.1f s"
                      % (npc2, w.n_points - 1, len(pix), s.n_pix, max(1, s.n_pix // max(1, len(pix))),
                         s.n_bins, dt1)}


def calibrate_copy(nbytes):
    """One elementwise copy kernel reading and writing nbytes each (the known
    size the FETCH_SIZE / WRITE_SIZE counters are scaled by)."""
    import torch
    n = nbytes // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    torch.mul(x, 2.0, out=y)
    torch.cuda.synchronize()
    log("copy kernel: read %d B, wrote %d B" % (4 * n, 4 * n))


def read_traffic(prec_name, workload):
    """HBM bytes per launch of the radiation kernel measured by rocprofv3
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), from the
    committed summary profiles/traffic_radiation_kernel_<prec>_<workload>.txt."""
    path = os.path.join(ROOT, "profiles", "traffic_radiation_kernel_%s_%s.txt" % (prec_name, workload))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--particles", type=int, default=None, help="shrink the workload (testing only)")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--window", type=int, default=0)
    ap.add_argument("--graph", action="store_true", help="replay the launch sequence from a hipGraph")
    ap.add_argument("--groups", type=int, default=0, help="window groups per (tile, chunk); 0 = automatic")
    ap.add_argument("--chunk", type=int, default=0, help="particles per chunk; 0 = automatic")
    ap.add_argument("--walk", type=int, default=0, help="1 = wave-uniform walk, 2 = per-lane walk, 0 = default")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the radiation step shards by pixel tiles "
              "across GPUs by design, but no multi-GPU path is built); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    ctx = teufel_amd.RadiationContext(0)
    dev, n_cu = ctx.device_info()
    log("device %s, %d CUs" % (dev.strip(), n_cu))
    w, ids = build_workload(ctx, args.workload, args.particles)
    n_pairs = w.n_pairs()
    n_inwin = ctx.count_in_window(-1)   # pairs passing the window test (particle.cpp:793)
    log("pairs: %d, passing the window test: %d (%.4f)" % (n_pairs, n_inwin, n_inwin / n_pairs))
    prec = TG_MIXED if args.precision == "mixed" else TG_FP64
    tuning = None
    if args.window or args.graph or args.groups or args.chunk or args.walk:
        tuning = {"bins_per_window": args.window, "use_graph": 1 if args.graph else 0,
                  "window_groups": args.groups, "particles_per_chunk": args.chunk, "walk": args.walk}
    tuned = None
    if not args.no_autotune and not args.window:
        t = time.time()
        tuned = ctx.autotune(prec)
        log("autotune %s: %s in %.1f s" % (args.precision, tuned, time.time() - t))
        if args.graph:
            tuning = {"bins_per_window": 0, "use_graph": 1}

    wall, kern, total = time_integrate(ctx, prec, args.steps, args.warmup, tuning)
    value = n_pairs / wall
    log("%s: %.3f s/step (kernel %.3f s, kernel+reduce %.3f s) -> %.4g pairs/s"
        % (args.precision, wall, kern, total, value))
    st = ctx.last_stats()
    if st["work_items"]:
        ideal = n_pairs / 64.0
        log("walk stats: %s; wave-steps / ideal = %.3f, steps per walk = %.1f"
            % (st, st["wave_steps"] / ideal, st["wave_steps"] / max(1, st["walks"])))
    peak = mixed_peak() if prec == TG_MIXED else PEAK_FP64
    achieved = n_pairs * FLOP_PER_PAIR / kern
    extra = {}
    if not args.no_fp64 and prec == TG_MIXED:
        # FP64 on the same workload, and the mixed result's error against it
        for i in ids:
            ctx.zero_screen(i)
        if not args.no_autotune:
            log("autotune fp64: %s" % ctx.autotune(TG_FP64))
        w64, k64, t64 = time_integrate(ctx, TG_FP64, args.fp64_steps, 0)
        # one clean pass of each precision for the error (timed passes accumulate)
        log("error check: one FP64 pass")
        for i in ids:
            ctx.zero_screen(i)
        ctx.integrate(TG_FP64)
        S64 = [ctx.screen_poynting(i) for i in ids]
        T64 = [ctx.read_traces(i) for i in ids]
        log("error check: one mixed pass")
        for i in ids:
            ctx.zero_screen(i)
        ctx.integrate(TG_MIXED, tuning)
        Sm = [ctx.screen_poynting(i) for i in ids]
        Tm = [ctx.read_traces(i) for i in ids]
        epsW = max(energy_rel_error(W.energy_density(a, s.normal), W.energy_density(b, s.normal))[0]
                   for a, b, s in zip(Sm, S64, w.screens))
        epsE = max(trace_rel_error(a, b) for a, b in zip(Tm, T64))
        log("fp64: %.3f s/step -> %.4g pairs/s; mixed/fp64 = %.2fx; eps_W = %.3e, eps_E = %.3e"
            % (w64, n_pairs / w64, w64 / wall, epsW, epsE))
        extra = {"fp64_value": n_pairs / w64, "fp64_ms_per_step": w64 * 1e3, "fp64_kernel_ms": k64 * 1e3,
                 "mixed_over_fp64": w64 / wall, "eps_W_mixed_vs_fp64": epsW, "eps_E_mixed_vs_fp64": epsE,
                 "fp64_roofline_frac": n_pairs * FLOP_PER_PAIR / k64 / PEAK_FP64}
    cpu = None
    if not args.no_cpu:
        t = time.time()
        cpu = cpu_baseline(ctx, w)
        log("cpu baseline (%s, %d threads): %.4g pairs/s [%s] in %.1f s"
            % (cpu["kind"], cpu["cores"], cpu["value"], cpu["sample"], time.time() - t))
        extra["gpu_over_cpu"] = value / cpu["value"]
    s0 = w.screens[0]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "LW field evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if prec == TG_MIXED else "f64",
        "data": "synthetic (seeded bunch, Vay-tracked on device through a K=1 planar undulator)",
        "config": {"workload": w.name, "description": w.description, "particles": w.n_particles,
                   "points_per_particle": w.n_points, "pixels": sum(s.n_pix for s in w.screens),
                   "bins": s0.n_bins, "screens": len(w.screens), "n_pairs": n_pairs,
                   "n_inwin": n_inwin,
                   "precision": args.precision, "tuning": tuned, "graph": bool(args.graph)},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": read_traffic(PREC_NAME[prec], w.name),
                     "kernel": KERNEL_NAME[prec], "flop_per_pair": FLOP_PER_PAIR, "kernel_ms": kern * 1e3},
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
