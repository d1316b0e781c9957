"""GPU checks at the full size of the headline workload, through properties
that hold at any size (the oracle would take hours here).

c5 as the benchmark runs it: 2e5 macro-electrons of 200 pC in all tracked on
the device through the 34-period undulator (4001 points each), the 128 x 128
screen at 10 m with 4096 bins, and one benchmark step, the batch of particles
[0, 1e4), through tg_integrate_range.

* Linearity (the reference sums particles into the traces,
  src/bunch.cpp:739-745): the batch integrated as two halves, into the same
  store, equals the batch in one launch.  FP64; the two sums differ only in
  how the particles are grouped before they are added (chunk and partial-
  trace boundaries), so the bound is the FP64 parity bound, 1e-12.
* Precision: the mixed result of the batch against the FP64 result, per-pixel
  energy over all 16384 pixels (pixels >= 1e-3 of the maximum): eps_W <= 1e-5,
  the stated error budget.
* Determinism: the mixed batch integrated twice gives bitwise equal traces.
"""
import numpy as np
import pytest

import teufel_amd
from teufel_amd import TG_FP64, TG_MIXED
from teufel_amd import workloads as W
from teufel_amd.metrics import energy_rel_error, trace_rel_error

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_ENERGY_TOL = 1e-5
BATCH = 10000


def say(msg):
    print("\n[case] " + msg, flush=True)


def test_c5_batch_linearity_precision_determinism():
    say("c5 full size (2e5 macro-electrons tracked, 128x128 pixels, 4096 bins), batch [0, %d): linearity, "
        "mixed vs FP64 on all pixels, determinism" % BATCH)
    w = W.make("c5")
    u, s = w.und, w.screens[0]
    assert w.n_particles == 200000 and s.n_pix == 16384 and s.n_bins == 4096
    with teufel_amd.RadiationContext(0) as ctx:
        ctx.track_undulator(w.bunch.x0, w.bunch.u0, w.dt, w.n_steps, u.bpeak, u.lambda_u, u.n_periods, u.origin,
                            charge=w.bunch.charge, mass=w.bunch.mass)
        sid = ctx.add_screen(s.obs, s.t0, s.dt, s.n_bins)

        ctx.integrate_range(TG_FP64, 0, BATCH, screen_id=sid)
        whole = ctx.read_traces(sid)
        S64 = ctx.screen_poynting(sid)
        ctx.zero_screen(sid)
        ctx.integrate_range(TG_FP64, 0, BATCH // 2, screen_id=sid)
        ctx.integrate_range(TG_FP64, BATCH // 2, BATCH, screen_id=sid)
        halves = ctx.read_traces(sid)
        assert np.abs(whole).max() > 0
        e = trace_rel_error(halves, whole)
        print("linearity (FP64, two halves against one launch): eps_E = %.3e" % e, flush=True)
        assert e <= FP64_TOL
        del halves

        ctx.zero_screen(sid)
        ctx.integrate_range(TG_MIXED, 0, BATCH, screen_id=sid)
        mixed = ctx.read_traces(sid)
        Sm = ctx.screen_poynting(sid)
        ew, below = energy_rel_error(W.energy_density(Sm, s.normal), W.energy_density(S64, s.normal))
        print("mixed against FP64: eps_W = %.3e (%.1f %% of the pixels below the floor), eps_E = %.3e"
              % (ew, 100 * below, trace_rel_error(mixed, whole)), flush=True)
        assert below < 0.5
        assert ew <= MIXED_ENERGY_TOL
        del whole

        ctx.zero_screen(sid)
        ctx.integrate_range(TG_MIXED, 0, BATCH, screen_id=sid)
        assert np.array_equal(ctx.read_traces(sid), mixed)
