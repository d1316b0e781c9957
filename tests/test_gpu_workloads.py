"""GPU parity on the benchmark workloads themselves (c2-c5 of
teufel_amd.workloads): the real screens, trace lengths, bunches and
undulators, with fewer particles, against the oracle on a stratified pixel
subset of every screen.

* c2: the whole 64 x 64 screen with 1024 bins;
* c3: the 128 x 128 screen with 2048 bins;
* c4: all 32 seeded uneven screens (flat, tilted, toroidal; 16^2..128^2
  pixels; 256..4096 bins; 2..20 m) in one launch;
* c5: -6241.5 e macro-particles (the charge of tests/teufel.DiffractionScreen.cpp:78
  style macro-particles, 200 pC in all at full size) with a Gaussian bunch,
  the 34-period undulator and 4096 bins.

The device pusher generates the trajectories (it is checked against the
oracle pusher elsewhere); both the HIP kernels and the oracle integrate those
same trajectories.  Bounds: FP64 eps_E <= 1e-12 against the oracle, mixed
eps_W <= 1e-5 per pixel against the oracle (energy through the pixel, pixels
>= 1e-3 of the maximum), the stated error budget.  The long-sum
cases put 1000-2000 particles into every cell so that the mixed kernel's FP32
window sums and their flushes run at the lengths of the full workloads.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from teufel_amd import TG_FP64, TG_MIXED
from teufel_amd import workloads as W
from teufel_amd.metrics import energy_rel_error, trace_rel_error

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_ENERGY_TOL = 1e-5
MU0 = 4e-7 * np.pi


def say(msg):
    print("\n[case] " + msg, flush=True)


def stratified(n_pix, k):
    """k pixel indices spread evenly over a screen (row-major grid order), the
    first and last pixel included."""
    return np.unique(np.linspace(0, n_pix - 1, min(k, n_pix)).round().astype(int))


def energy(traces, dt, normal):
    S = dt * np.cross(traces[..., 0:3], traces[..., 3:6]).sum(axis=1) / MU0
    return W.energy_density(S, normal)


def setup(ctx, name, n_particles):
    w = W.make(name, n_particles=n_particles)
    u = w.und
    ctx.remove_screens()
    ctx.track_undulator(w.bunch.x0, w.bunch.u0, w.dt, w.n_steps, u.bpeak, u.lambda_u, u.n_periods, u.origin,
                        charge=w.bunch.charge, mass=w.bunch.mass)
    tr = oracle.Trajectories(**ctx.get_trajectories())
    ids = [ctx.add_screen(s.obs, s.t0, s.dt, s.n_bins) for s in w.screens]
    return w, tr, ids


def oracle_subset(tr, w, per_screen, threads=16):
    """Oracle traces of a stratified pixel subset of every screen: one screen
    with OpenMP threads over its pixels, several screens in a pool of host
    threads (the oracle call releases the interpreter lock)."""
    def one(s, n_threads):
        pick = stratified(s.n_pix, per_screen)
        return pick, oracle.integrate_screen(tr, s.obs[pick], s.t0[pick], s.dt, s.n_bins, n_threads=n_threads)

    if len(w.screens) == 1:
        return [one(w.screens[0], threads)]
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda s: one(s, 1), w.screens))


def check(ctx, w, ids, ref, prec, label):
    worst_e, worst_w = 0.0, 0.0
    for sid, s, (pick, r) in zip(ids, w.screens, ref):
        got = ctx.read_traces(sid)[pick]
        assert np.abs(r).max() > 0, (label, s.name)
        e = trace_rel_error(got, r)
        worst_e = max(worst_e, e)
        if prec == TG_FP64:
            assert e <= FP64_TOL, (label, s.name, e)
        else:
            ew, _ = energy_rel_error(energy(got, s.dt, s.normal[pick]), energy(r, s.dt, s.normal[pick]))
            worst_w = max(worst_w, ew)
            assert ew <= MIXED_ENERGY_TOL, (label, s.name, ew)
    print("%s: eps_E = %.3e, eps_W = %.3e over %d screens" % (label, worst_e, worst_w, len(ids)), flush=True)


CASES = [("c2", 24, 96), ("c3", 12, 96), ("c4", 6, 6), ("c5", 10, 96)]


@pytest.mark.parametrize("name,n_particles,per_screen", CASES)
@pytest.mark.parametrize("prec", [TG_FP64, TG_MIXED])
def test_workload_vs_oracle(ctx, name, n_particles, per_screen, prec):
    say("%s, %d particles, all screens on the device, %d oracle pixels per screen, prec=%d"
        % (name, n_particles, per_screen, prec))
    w, tr, ids = setup(ctx, name, n_particles)
    ref = oracle_subset(tr, w, per_screen)
    ctx.integrate(prec)
    check(ctx, w, ids, ref, prec, "%s prec=%d" % (name, prec))


@pytest.mark.parametrize("name,n_particles,per_screen", [("c5", 2000, 16), ("c4", 1000, 2)])
def test_long_sums_vs_oracle(ctx, name, n_particles, per_screen):
    say("%s long sums: %d particles into every cell, %d oracle pixels per screen; FP64 and mixed, "
        "default plan and one particle chunk" % (name, n_particles, per_screen))
    w, tr, ids = setup(ctx, name, n_particles)
    ref = oracle_subset(tr, w, per_screen)
    ctx.integrate(TG_FP64)
    check(ctx, w, ids, ref, TG_FP64, name + " fp64")
    for tun in (None, {"particles_per_chunk": n_particles, "spill": 1}, {"particles_per_chunk": n_particles, "spill": 2}):
        for i in ids:
            ctx.zero_screen(i)
        ctx.integrate(TG_MIXED, tuning=tun)
        check(ctx, w, ids, ref, TG_MIXED, "%s mixed %s" % (name, tun))
