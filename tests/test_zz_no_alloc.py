"""tg_integrate_range_static: the integrate entry point whose launch scratch
lives in the library's static device array (no device allocation).

It runs the kernels of tg_integrate_range with one particle chunk and must
give that entry point's result to the parity bar of the path: FP64 traces
within 1e-12 of each pixel's largest field (the two order the particle sums
differently: chunks and partial traces against one chunk), mixed precision
within 1e-5 of the per-pixel energy, and both against the oracle to the same
bars.  Two calls in a row, the second with fewer particles and only one of
two screens, check that the shared array's contents from the first call do
not leak into the second.  The error paths: before set-up (records not
packed) TG_ERR_STATE, and a particle range whose cursor slots do not fit the
array TG_ERR_RANGE.
"""
import numpy as np
import pytest

import oracle
import teufel_amd
from teufel_amd import TG_FP64, TG_MIXED
from teufel_amd import workloads as W
from teufel_amd.metrics import energy_rel_error, trace_rel_error

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_ENERGY_TOL = 1e-5


def say(msg):
    print("\n[case] " + msg, flush=True)


def _setup(n_particles=40, n_steps=3000):
    w = W.make("c2", n_particles=n_particles, n_steps=n_steps, seed=91)
    u = w.und
    tr = oracle.track_undulator(w.bunch.x0, w.bunch.u0, w.dt, w.n_steps, u.bpeak, u.lambda_u, u.n_periods,
                                u.origin)
    t_c, x_c = W._centre(u, 60.0)
    obs1, n1, _ = W.grid_screen(16, 0.2, 10.0)
    t01, dt1 = W.set_windows(obs1, 256, u, 60.0, t_c, x_c, 0.3e-3)
    obs2, n2, _ = W.grid_screen(9, 0.05, 5.0, tilt_deg=20.0)
    t02, dt2 = W.set_windows(obs2, 131, u, 60.0, t_c, x_c, 0.3e-3)
    return tr, [(obs1, t01, dt1, 256, n1), (obs2, t02, dt2, 131, n2)]


def _energy(traces, dt, normal):
    S = dt * np.cross(traces[..., 0:3], traces[..., 3:6]).sum(axis=1) / (4e-7 * np.pi)
    return W.energy_density(S, normal)


def _check(prec, got, ref, dt, normal, what):
    if prec == TG_FP64:
        e = trace_rel_error(got, ref)
        print("%s: eps_E = %.3e" % (what, e))
        assert e <= FP64_TOL, what
    else:
        ew, _ = energy_rel_error(_energy(got, dt, normal), _energy(ref, dt, normal))
        print("%s: eps_W = %.3e" % (what, ew))
        assert ew <= MIXED_ENERGY_TOL, what


@pytest.mark.parametrize("prec", [TG_FP64, TG_MIXED])
def test_static_scratch_entry_matches_range_entry(prec):
    say("tg_integrate_range_static vs tg_integrate_range, two calls (40 then 17 particles, 2 then 1 screen), "
        "prec=%d" % prec)
    tr, geo = _setup()
    with teufel_amd.RadiationContext(0) as ctx:
        ctx.set_trajectories(tr)
        ids = [ctx.add_screen(o, t, d, nb) for o, t, d, nb, _ in geo]
        # before set-up: the records of this precision are not packed
        with pytest.raises(teufel_amd.TgError, match="not packed"):
            ctx.integrate_range_static(prec, 0, tr.n_particles)
        calls = [((0, tr.n_particles), -1), ((11, 28), ids[1])]
        for (p0, p1), sid in calls:
            targets = ids if sid < 0 else [sid]
            for i in ids:
                ctx.zero_screen(i)
            ctx.integrate_range(prec, p0, p1, screen_id=sid)      # also the set-up of the first call
            ref = {i: ctx.read_traces(i) for i in targets}
            for i in ids:
                ctx.zero_screen(i)
            ctx.integrate_range_static(prec, p0, p1, screen_id=sid)
            for i in ids:
                got = ctx.read_traces(i)
                o, t, d, nb, nrm = geo[i]
                if i not in targets:
                    assert not got.any(), "screen %d outside the call was written" % i
                    continue
                _check(prec, got, ref[i], d, nrm, "static vs range, particles %d-%d, screen %d" % (p0, p1, i))
                orc = oracle.integrate_screen(tr.subset(p0, p1), o, t, d, nb)
                _check(prec, got, orc, d, nrm, "static vs oracle, particles %d-%d, screen %d" % (p0, p1, i))


def test_static_scratch_entry_rejects_what_does_not_fit():
    say("tg_integrate_range_static: 300000 one-step particles on 4096 pixels need more cursor slots than the "
        "array holds; 2000 fit")
    n = 300000
    tr = oracle.Trajectories.empty(n, 2)
    tr.t[:, 1] = 1e-12
    tr.z[:, 1] = 2.99e-4                       # one step at beta = 0.9973
    tr.x[:] = np.linspace(-1e-9, 1e-9, n)[:, None]
    tr.uz[:] = 60.0
    # 64 pixel tiles x 1024 bins: thousands of work items, so thousands of
    # persistent waves, each with a cursor slot for all particles of the call
    obs, _, _ = W.grid_screen(64, 0.2, 10.0)
    t0 = np.linalg.norm(obs, axis=1) / W.C_LIGHT - 2e-16   # the step arrives in the window
    with teufel_amd.RadiationContext(0) as ctx:
        ctx.set_trajectories(tr)
        sid = ctx.add_screen(obs, t0, 1e-16, 1024)
        ctx.integrate_range(TG_FP64, 0, 64, screen_id=sid)        # set-up
        with pytest.raises(teufel_amd.TgError, match="particles per call"):
            ctx.integrate_range_static(TG_FP64, 0, n, screen_id=sid)
        # what fits still runs, and matches the allocating entry point
        ctx.zero_screen(sid)
        ctx.integrate_range(TG_FP64, 0, 2000, screen_id=sid)
        ref = ctx.read_traces(sid)
        ctx.zero_screen(sid)
        ctx.integrate_range_static(TG_FP64, 0, 2000, screen_id=sid)
        assert np.abs(ref).max() > 0
        assert trace_rel_error(ctx.read_traces(sid), ref) <= FP64_TOL
