"""Benchmark of the radiation step on one MI355X.

Metric: LW field evaluations per second = (particle-step x pixel) pairs / s,
N_pairs = sum_screens n_pix x sum_particles (NP - 1), with the trajectories
already resident in HBM (generated on the device by the Vay pusher before
timing).

Default workload: c5, the largest configuration (2e5 macro-electrons of
200 pC in all, Gaussian sigma_z 0.1 mm, 34-period undulator, 4000 steps,
128 x 128 pixel screen 10 m behind the undulator, 4096 bins).  The K timed
steps are K particle batches that together cover the whole bunch once:
step i integrates particles [i B, (i+1) B), B = ceil(n / K), through
tg_integrate_range into the one trace store, as a caller streaming batches
would; the first warm-up step captures the launch in a hipGraph and every
later step replays it, shifted to its batch (--no-graph: eager launches).
The W warm-up steps run the first batches into the store, which is zeroed
before the timed steps, so after them the store holds the full
mixed-precision result.  `value` = pairs per step / median step time.  The
FP64 kernel then runs the same K batches, the same way, into a zeroed store (fp64_value),
and the full mixed result's per-pixel energy error against the full FP64
result is reported (eps_W_mixed_vs_fp64, bound 1e-5).

c2 (1e4 electrons, 64 x 64 pixels, 1024 bins), whose whole launch repeats
step after step, is measured after the headline from hipGraph replays of the
captured launch (keys c2_*).  --workload c2|c3|c4 makes any workload the
headline with one whole integrate per step.

The CPU baseline is the reference's own compiled integrateFieldTrace
(oracle/_ref) driven by an OpenMP pixel loop on the host cores (the oracle
restatement where the reference build is absent), timed on a stratified pixel
sample of the headline workload for at least 60 s and scaled by pairs.

roofline: the radiation kernel's algorithmic FP operations (106 per pair) per
launch / its mean launch duration from HIP events on the launch stream,
against the vector peak of the precision mix (VALU-bound: no MFMA shape).
traffic: HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE, read
from profiles/traffic_radiation_kernel_<precision>_<workload>.txt when present.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c5]
       python bench.py --calibrate-copy BYTES   (a plain copy of BYTES, for
                                                 scaling the PMC byte counters)
"""
import argparse
import json
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
METRIC = "LW field evals/s (particle-step×pixel), mixed vs FP64 GPU at ≤1e-5 rel energy"
KERNEL_NAME = {TG_MIXED: "radiation_kernel<1, true>", TG_FP64: "radiation_kernel<0, true>"}
PREC_NAME = {TG_MIXED: "mixed", TG_FP64: "fp64"}
BATCHED = ("c5",)              # workloads whose timed steps are particle batches


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


class Steps:
    """Per-step wall and kernel times of one timed leg."""

    def __init__(self):
        self.wall, self.kernel, self.total, self.pairs = [], [], [], []

    def median_wall(self):
        return float(np.median(self.wall))

    def summary(self):
        return {"median_s": self.median_wall(), "mean_s": float(np.mean(self.wall)),
                "min_s": float(np.min(self.wall)), "max_s": float(np.max(self.wall)),
                "kernel_mean_s": float(np.mean(self.kernel)), "steps": len(self.wall)}


def run_steps(ctx, prec, ranges, pairs_per_particle, tuning=None, label="step"):
    """One integrate (whole, or a particle range) per entry of ranges; each
    step is timed on the host around the call (the call returns after the
    launch has finished) and by HIP events on the launch stream."""
    st = Steps()
    for i, r in enumerate(ranges):
        t = time.perf_counter()
        if r is None:
            ctx.integrate(prec, tuning)
            n = None
        else:
            ctx.integrate_range(prec, r[0], r[1], tuning=tuning)
            n = r[1] - r[0]
        st.wall.append(time.perf_counter() - t)
        k, tot = ctx.last_timing()
        st.kernel.append(k / 1e3)
        st.total.append(tot / 1e3)
        st.pairs.append(None if n is None else n * pairs_per_particle)
        log("  %s %d/%d: %.3f s (kernel %.3f s)" % (label, i + 1, len(ranges), st.wall[-1], st.kernel[-1]))
    ctx.synchronize()
    return st


def batch_ranges(n, k):
    b = -(-n // k)
    return [(i * b, min(n, (i + 1) * b)) for i in range(k) if i * b < n]


def cpu_model():
    """The host CPU's model name with its CPUID family, model and stepping
    (/proc/cpuinfo of the first CPU).  A virtual machine may report a generic
    name such as "AMD EPYC"; the three numbers still identify the core
    generation (family 26 is Zen 5, 25 is Zen 3/4)."""
    keys = {"model name": None, "cpu family": None, "model": None, "stepping": None}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break   # end of the first CPU's block
                k, _, v = line.partition(":")
                k = k.strip()
                if k in keys and keys[k] is None:
                    keys[k] = v.strip()
    except OSError:
        pass
    name = keys["model name"] or "unknown"
    if keys["cpu family"] is not None:
        name += " (family %s, model %s, stepping %s)" % (keys["cpu family"], keys["model"], keys["stepping"])
    return name


def physical_cores():
    """Distinct physical cores among the CPUs this process may run on."""
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        base = "/sys/devices/system/cpu/cpu%d/topology/" % c
        try:
            with open(base + "physical_package_id") as f:
                pkg = f.read().strip()
            with open(base + "core_id") as f:
                core = f.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", str(c)))
    return len(cores), len(cpus)


def cpu_baseline(ctx, w, target_s=65.0, pixel_stride=16):
    """The reference's integrateFieldTrace (compiled from the reference
    sources) on the host cores; the oracle restatement if the reference build
    is absent.  Timed on every pixel_stride-th pixel of the first screen (a
    stratified sample across the screen) and as many particles of the bunch as
    give about target_s seconds; per-pair work is uniform, so the rate scales
    to the whole workload."""
    import oracle
    n_cpus_allowed = physical_cores()
    n_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or n_cpus_allowed[1]
    s = w.screens[0]
    kind = "reference" if oracle.ref_lib() is not None else "port"
    pix = np.arange(0, s.n_pix, pixel_stride)

    def run(npart):
        d = ctx.get_trajectories(0, npart)
        sub = oracle.Trajectories(**d)
        t = time.perf_counter()
        if kind == "reference":
            rb = oracle.RefBeam(sub)
            rb.integrate(s.obs[pix], s.t0[pix], s.dt, s.n_bins, n_threads=n_threads)
            rb.close()
        else:
            oracle.integrate_screen(sub, s.obs[pix], s.t0[pix], s.dt, s.n_bins, n_threads=n_threads)
        return time.perf_counter() - t

    n0 = min(w.n_particles, 4)
    dt0 = run(n0)
    rate0 = len(pix) * n0 * (w.n_points - 1) / dt0
    npart = int(min(w.n_particles, max(1, round(rate0 * target_s / (len(pix) * (w.n_points - 1))))))
    log("cpu baseline: calibration %.3g pairs/s; timing %d particles x %d pixels" % (rate0, npart, len(pix)))
    dt1 = run(npart)
    pairs = len(pix) * npart * (w.n_points - 1)
    return {"value": pairs / dt1, "unit": "LW field evals/s", "cores": n_threads, "kind": kind,
            "cpu": cpu_model(), "physical_cores_available": n_cpus_allowed[0],
            "logical_cpus_available": n_cpus_allowed[1],
            "sample": "%s: %d particles x %d steps x %d of %d pixels (every %d-th), %d bins, %.1f s"
                      % (w.name, npart, w.n_points - 1, len(pix), s.n_pix, pixel_stride, s.n_bins, dt1)}


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


def errors(ctx, ids, w, S_ref, T_ref):
    """eps_W, eps_E of the screens' current traces against (S_ref, T_ref)."""
    S = [ctx.screen_poynting(i) for i in ids]
    T = [ctx.read_traces(i) for i in ids]
    epsW = max(energy_rel_error(W.energy_density(a, s.normal), W.energy_density(b, s.normal))[0]
               for a, b, s in zip(S, S_ref, w.screens))
    epsE = max(trace_rel_error(a, b) for a, b in zip(T, T_ref))
    return epsW, epsE


def c2_graph_keys(steps, warmup):
    """c2 from hipGraph replays: mixed and FP64, each autotuned, one captured
    launch replayed `steps` times after `warmup` replays; the mixed result's
    error against the FP64 result of the same replays."""
    out = {}
    with teufel_amd.RadiationContext(0) as ctx:
        w, ids = build_workload(ctx, "c2")
        n_pairs = w.n_pairs()
        res = {}
        for prec in (TG_MIXED, TG_FP64):
            name = PREC_NAME[prec]
            t = time.time()
            tuned = ctx.autotune(prec)
            log("c2 autotune %s: %s in %.1f s" % (name, tuned, time.time() - t))
            tun = {"bins_per_window": 0, "use_graph": 1}
            run_steps(ctx, prec, [None] * max(1, warmup), 0, tun, "c2 %s graph warmup" % name)
            for i in ids:
                ctx.zero_screen(i)
            st = run_steps(ctx, prec, [None], 0, tun, "c2 %s graph replay" % name)
            S = [ctx.screen_poynting(i) for i in ids]
            T = [ctx.read_traces(i) for i in ids]
            st2 = run_steps(ctx, prec, [None] * max(0, steps - 1), 0, tun, "c2 %s graph replay" % name)
            for a in ("wall", "kernel", "total"):
                getattr(st, a).extend(getattr(st2, a))
            res[prec] = (st, S, T)
            out["c2_%s_value" % name] = n_pairs / st.median_wall()
            out["c2_%s_ms_per_step" % name] = st.median_wall() * 1e3
            out["c2_%s_kernel_ms" % name] = float(np.mean(st.kernel)) * 1e3
            out["c2_%s_tuning" % name] = tuned
        S64, T64 = res[TG_FP64][1], res[TG_FP64][2]
        Sm, Tm = res[TG_MIXED][1], res[TG_MIXED][2]
        out["c2_eps_W_mixed_vs_fp64"] = max(
            energy_rel_error(W.energy_density(a, s.normal), W.energy_density(b, s.normal))[0]
            for a, b, s in zip(Sm, S64, w.screens))
        out["c2_eps_E_mixed_vs_fp64"] = max(trace_rel_error(a, b) for a, b in zip(Tm, T64))
        out["c2_mixed_over_fp64"] = res[TG_FP64][0].median_wall() / res[TG_MIXED][0].median_wall()
        out["c2_n_pairs"] = n_pairs
        out["c2_graph"] = True
        log("c2 (graph replays): mixed %.4g, fp64 %.4g pairs/s, ratio %.2f, eps_W %.3e"
            % (out["c2_mixed_value"], out["c2_fp64_value"], out["c2_mixed_over_fp64"],
               out["c2_eps_W_mixed_vs_fp64"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c5")
    ap.add_argument("--particles", type=int, default=None, help="shrink the workload (testing only)")
    ap.add_argument("--batch", type=int, default=0,
                    help="particles per batch step (batched workloads; 0: ceil(particles / steps))")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--fp64-steps", type=int, default=0,
                    help="FP64 steps of a non-batched workload (0: as many as --steps, at most 3)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c2", action="store_true", help="skip the c2 graph-replay keys")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--window", type=int, default=0)
    ap.add_argument("--graph", action="store_true", help="replay the launch sequence from a hipGraph")
    ap.add_argument("--no-graph", action="store_true",
                    help="eager launches for the batched workloads (their default is hipGraph replays)")
    ap.add_argument("--groups", type=int, default=0, help="window groups per (tile, chunk); 0 = automatic")
    ap.add_argument("--chunk", type=int, default=0, help="particles per chunk; 0 = automatic")
    ap.add_argument("--walk", type=int, default=0, help="1 = wave-uniform walk, 2 = per-lane walk, 0 = default")
    ap.add_argument("--spill", type=int, default=0, help="spill bucket: 1 = on, 2 = off, 0 = default (on)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the radiation step shards by pixel tiles "
              "across GPUs by design, but no multi-GPU path is built); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        print("bench.py: --steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        sys.exit(2)

    ctx = teufel_amd.RadiationContext(0)
    dev, n_cu = ctx.device_info()
    log("device %s, %d CUs" % (dev.strip(), n_cu))
    w, ids = build_workload(ctx, args.workload, args.particles)
    n_pairs = w.n_pairs()
    pairs_per_particle = n_pairs // w.n_particles
    t = time.time()
    n_inwin = ctx.count_in_window(-1)   # pairs passing the window test (particle.cpp:793)
    log("pairs: %d, passing the window test: %d (%.4f), counted in %.1f s"
        % (n_pairs, n_inwin, n_inwin / n_pairs, time.time() - t))
    prec = TG_MIXED if args.precision == "mixed" else TG_FP64
    batched = w.name in BATCHED
    if batched:
        b = args.batch or -(-w.n_particles // args.steps)
        timed = batch_ranges(w.n_particles, -(-w.n_particles // b))[:args.steps]
        warm = [timed[i % len(timed)] for i in range(args.warmup)]
        log("batched steps: %d batches of %d particles (%d timed, %d warm-up)"
            % (-(-w.n_particles // b), b, len(timed), len(warm)))
    else:
        timed, warm = [None] * args.steps, [None] * args.warmup
    # batched steps replay one captured launch per batch size (the library
    # shifts its particle range), unless --no-graph
    graph = args.graph or (batched and not args.no_graph)
    tuning = None
    if args.window or graph or args.groups or args.chunk or args.walk or args.spill:
        tuning = {"bins_per_window": args.window, "use_graph": 1 if graph else 0,
                  "window_groups": args.groups, "particles_per_chunk": args.chunk, "walk": args.walk,
                  "spill": args.spill}
    tuned = None
    if not args.no_autotune and not args.window:
        t = time.time()
        tuned = ctx.autotune(prec)
        log("autotune %s: %s in %.1f s" % (args.precision, tuned, time.time() - t))
        if graph:
            tuning = {"bins_per_window": 0, "use_graph": 1}

    run_steps(ctx, prec, warm, pairs_per_particle, tuning, "warmup")
    for i in ids:
        ctx.zero_screen(i)
    st = run_steps(ctx, prec, timed, pairs_per_particle, tuning, "step")
    covered = sum(r[1] - r[0] for r in timed) if batched else w.n_particles * len(timed)
    pairs_step = (n_pairs if not batched else float(np.median(st.pairs)))
    value = pairs_step / st.median_wall()
    log("%s: median %.3f s/step (kernel mean %.3f s) -> %.4g pairs/s; %d particles covered"
        % (args.precision, st.median_wall(), float(np.mean(st.kernel)), value, covered))
    stats = ctx.last_stats()
    if stats["work_items"]:
        ideal = (st.pairs[-1] if batched else n_pairs) / 64.0
        log("walk stats (last step): %s; wave-steps / ideal = %.3f, steps per walk = %.1f"
            % (stats, stats["wave_steps"] / ideal, stats["wave_steps"] / max(1, stats["walks"])))
    peak = mixed_peak() if prec == TG_MIXED else PEAK_FP64
    kern_mean = float(np.mean(st.kernel))
    pairs_launch = float(np.mean(st.pairs)) if batched else float(n_pairs)
    achieved = pairs_launch * FLOP_PER_PAIR / kern_mean
    extra = {"timing": st.summary(), "value_from_total_wall": (sum(st.pairs) if batched else n_pairs * len(timed))
             / sum(st.wall)}
    if not args.no_fp64 and prec == TG_MIXED:
        # the store now holds the timed steps' result: the whole bunch once
        # (batched) or len(timed) whole passes
        S_m = [ctx.screen_poynting(i) for i in ids]
        T_m = [ctx.read_traces(i) for i in ids]
        for i in ids:
            ctx.zero_screen(i)
        if not args.no_autotune:
            t = time.time()
            log("autotune fp64: %s in %.1f s" % (ctx.autotune(TG_FP64), time.time() - t))
        if batched:
            # the same timing path as the mixed steps (warm-up captures)
            tun64 = {"bins_per_window": 0, "use_graph": 1} if graph else None
            if graph:
                run_steps(ctx, TG_FP64, timed[:1], pairs_per_particle, tun64, "fp64 warmup")
                for i in ids:
                    ctx.zero_screen(i)
            st64 = run_steps(ctx, TG_FP64, timed, pairs_per_particle, tun64, "fp64 step")
        else:
            n64 = args.fp64_steps or min(3, len(timed))
            st64 = run_steps(ctx, TG_FP64, [None] * n64, pairs_per_particle, None, "fp64 step")
            if n64 != len(timed):   # the error check compares equal numbers of passes
                for i in ids:
                    ctx.zero_screen(i)
                log("error check: %d FP64 passes" % len(timed))
                run_steps(ctx, TG_FP64, timed, pairs_per_particle, None, "fp64 pass")
        log("error check: mixed result against the FP64 result")
        epsW, epsE = errors(ctx, ids, w, S_m, T_m)
        w64 = st64.median_wall()
        fp64_value = (float(np.median(st64.pairs)) if batched else n_pairs) / w64
        log("fp64: median %.3f s/step -> %.4g pairs/s; mixed/fp64 = %.2fx; eps_W = %.3e, eps_E = %.3e"
            % (w64, fp64_value, value / fp64_value, epsW, epsE))
        extra.update({"fp64_value": fp64_value, "fp64_ms_per_step": w64 * 1e3,
                      "fp64_kernel_ms": float(np.mean(st64.kernel)) * 1e3, "fp64_timing": st64.summary(),
                      "mixed_over_fp64": value / fp64_value, "eps_W_mixed_vs_fp64": epsW,
                      "eps_E_mixed_vs_fp64": epsE,
                      "fp64_roofline_frac": pairs_launch * FLOP_PER_PAIR / float(np.mean(st64.kernel)) / PEAK_FP64,
                      "eps_scope": ("full bunch, all %d pixels" % sum(s.n_pix for s in w.screens))})
    cpu = None
    if not args.no_cpu:
        t = time.time()
        cpu = cpu_baseline(ctx, w)
        log("cpu baseline (%s, %d threads): %.4g pairs/s [%s] in %.1f s"
            % (cpu["kind"], cpu["cores"], cpu["value"], cpu["sample"], time.time() - t))
        extra["gpu_over_cpu"] = value / cpu["value"]
    s0 = w.screens[0]
    cfg = {"workload": w.name, "description": w.description, "particles": w.n_particles,
           "points_per_particle": w.n_points, "pixels": sum(s.n_pix for s in w.screens),
           "bins": s0.n_bins, "screens": len(w.screens), "n_pairs": n_pairs, "n_inwin": n_inwin,
           "precision": args.precision, "tuning": tuned, "graph": bool(graph),
           "step": ("one particle batch of %d through tg_integrate_range%s; the %d timed steps cover %d of %d "
                    "particles" % (timed[0][1] - timed[0][0],
                                   " (a replay of the captured launch, shifted to the batch)" if graph else "",
                                   len(timed), covered, w.n_particles))
           if batched else "one whole tg_integrate"}
    ctx.close()
    if not args.no_c2 and w.name != "c2":
        extra.update(c2_graph_keys(5, 1))
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "LW field evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": st.median_wall() * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if prec == TG_MIXED else "f64",
        "data": "synthetic (seeded bunch, Vay-tracked on device through a K=1 planar undulator)",
        "config": cfg,
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": read_traffic(PREC_NAME[prec], w.name),
                     "kernel": KERNEL_NAME[prec], "flop_per_pair": FLOP_PER_PAIR, "kernel_ms": kern_mean * 1e3,
                     "pairs_per_launch": pairs_launch,
                     "timed_by": ("HIP events around each graph replay (the radiation kernel and the partial-"
                                  "trace reduction after it)" if graph else
                                  "HIP events around the radiation kernel on its stream")},
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
