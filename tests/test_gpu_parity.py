"""GPU parity: the HIP path (through the C ABI) against the oracle and the
reference's golden fixtures.

FP64 tolerance: eps_E <= 1e-12 (teufel_amd.metrics.trace_rel_error).  The
retarded time, bucket index and bucket edges are computed bit-identically to
the reference, so what remains is amplitude rounding (~1e-15).
Mixed tolerance: eps_W <= 1e-5 per pixel (energy), the stated budget.
"""
import math

import numpy as np
import pytest

import oracle
import teufel_amd
from teufel_amd import TG_FP64, TG_MIXED
from teufel_amd import workloads as W
from teufel_amd.metrics import energy_rel_error, trace_rel_error
from tracking_cases import exb_case, magnet_case

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_ENERGY_TOL = 1e-5


def say(msg):
    print("\n[case] " + msg, flush=True)


def _traj_from(d):
    return oracle.Trajectories(**{n: d[n] for n in oracle.Trajectories.NAMES}, charge=d["charge"])


def test_deposit_kats_on_device(ctx, golden):
    say("deposit KATs tests/teufel.integrate_field.cpp")
    g = golden("kat.npz")
    for i in range(5):
        t0, dt, n = g["k%d_par" % i]
        tr = ctx.deposit_only(g["k%d_ts" % i], g["k%d_f" % i], t0, dt, int(n))
        assert np.allclose(tr[:, 0], g["k%d_expected" % i], atol=1e-6, rtol=0), i
        assert np.abs(tr - g["k%d_ref" % i]).max() <= 1e-15, i


def test_c1_fp64_matches_reference_fixture(ctx, golden):
    say("c1 FP64 vs reference fixture")
    g = golden("c1_ref.npz")
    tr = _traj_from(g)
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    sid = ctx.add_screen(g["obs"], g["t0"], float(g["dt"]), int(g["n_bins"]))
    ctx.integrate(TG_FP64)
    out = ctx.read_traces(sid)
    err = trace_rel_error(out[g["pick"]], g["traces"])
    print("eps_E(fixture) = %.3e" % err)
    assert err <= FP64_TOL
    S = ctx.screen_poynting(sid)
    assert np.abs(S - g["poynting"]).max() <= 1e-12 * np.abs(g["poynting"]).max()
    ref = oracle.integrate_screen(tr, g["obs"], g["t0"], float(g["dt"]), int(g["n_bins"]))
    err2 = trace_rel_error(out, ref)
    print("eps_E(oracle, all 441 px) = %.3e" % err2)
    assert err2 <= FP64_TOL


def test_bunch_fixture_fp64(ctx, golden):
    say("3-electron bunch FP64 vs reference fixture")
    g = golden("bunch_ref.npz")
    ctx.remove_screens()
    ctx.set_trajectories(_traj_from(g))
    sid = ctx.add_screen(g["obs"], g["t0"], float(g["dt"]), int(g["n_bins"]))
    ctx.integrate(TG_FP64)
    err = trace_rel_error(ctx.read_traces(sid), g["traces"])
    print("eps_E = %.3e" % err)
    assert err <= FP64_TOL


def test_loop_point_observer(ctx, golden):
    say("circular loop, 2-bin point observer (tests/teufel.loop.cpp)")
    g = golden("loop_ref.npz")
    ctx.remove_screens()
    ctx.set_trajectories(_traj_from(g))
    sid = ctx.add_screen(g["obs"], [float(g["t0"])], float(g["dt"]), int(g["n_bins"]))
    ctx.integrate(TG_FP64)
    out = ctx.read_traces(sid)
    assert trace_rel_error(out, g["traces"]) <= FP64_TOL
    B = np.linalg.norm(out[0, 0, 3:])
    assert abs(B / float(g["b_expected"]) - 1.0) < 1e-3


def _small_bunch(n_particles=48, n_steps=4000, npx=16, n_bins=256, seed=77):
    w = W.make("c2", n_particles=n_particles, n_steps=n_steps, seed=seed)
    u = w.und
    tr = oracle.track_undulator(w.bunch.x0, w.bunch.u0, w.dt, w.n_steps, u.bpeak, u.lambda_u, u.n_periods,
                                u.origin, charge=-1.0)
    obs, nrm, area = W.grid_screen(npx, 0.2, 10.0)
    t_c, x_c = W._centre(u, 60.0)
    t0, dt = W.set_windows(obs, n_bins, u, 60.0, t_c, x_c, 0.3e-3)
    return tr, obs, t0, dt, nrm


@pytest.mark.parametrize("window,chunk,groups,walk,spill", [
    (8, 0, 1, 2, 0), (4, 7, 3, 2, 0), (16, 1000, 2, 1, 0), (32, 16, 5, 2, 2), (6, 0, 0, 1, 0), (12, 0, 0, 2, 2),
    (5, 3, 4, 1, 0), (2, 5, 7, 2, 1), (1, 0, 9, 2, 1), (3, 48, 1, 2, 2)])
def test_bunch_fp64_vs_oracle_launch_shapes(ctx, window, chunk, groups, walk, spill):
    say("48 e- x 4000 steps, 16x16 px, FP64, window=%d chunk=%d groups=%d walk=%d spill=%d"
        % (window, chunk, groups, walk, spill))
    tr, obs, t0, dt, _ = _small_bunch()
    ref = oracle.integrate_screen(tr, obs, t0, dt, 256)
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    sid = ctx.add_screen(obs, t0, dt, 256)
    ctx.integrate(TG_FP64, tuning={"bins_per_window": window, "particles_per_chunk": chunk,
                                   "window_groups": groups, "walk": walk, "spill": spill})
    err = trace_rel_error(ctx.read_traces(sid), ref)
    print("eps_E = %.3e" % err)
    assert err <= FP64_TOL


@pytest.mark.parametrize("walk", [1, 2])
def test_mixed_energy_error_and_determinism(ctx, walk):
    say("48 e- mixed vs FP64, energy per pixel; repeat bitwise; walk=%d" % walk)
    tr, obs, t0, dt, nrm = _small_bunch()
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    sid = ctx.add_screen(obs, t0, dt, 256)
    ctx.integrate(TG_FP64)
    S64 = ctx.screen_poynting(sid)
    T64 = ctx.read_traces(sid)
    ctx.zero_screen(sid)
    ctx.integrate(TG_MIXED, tuning={"walk": walk})
    Tm = ctx.read_traces(sid)
    Sm = ctx.screen_poynting(sid)
    ctx.zero_screen(sid)
    ctx.integrate(TG_MIXED, tuning={"walk": walk})
    Tm2 = ctx.read_traces(sid)
    assert np.array_equal(Tm, Tm2)
    epsW, excl = energy_rel_error(W.energy_density(Sm, nrm), W.energy_density(S64, nrm))
    epsE = trace_rel_error(Tm, T64)
    print("mixed: eps_W = %.3e (excluded %.3f), eps_E = %.3e" % (epsW, excl, epsE))
    assert epsW <= MIXED_ENERGY_TOL


def test_count_in_window_matches_oracle(ctx):
    say("pairs passing the window test: 48 e- x 4000 steps, 16x16 px, full and narrow windows, vs oracle")
    tr, obs, t0, dt, _ = _small_bunch()
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    s_full = ctx.add_screen(obs, t0, dt, 256)
    s_narrow = ctx.add_screen(obs, t0 + 60 * dt, dt / 4, 128)   # a quarter of the pulse
    s_one = ctx.add_screen(obs, t0, dt, 1)                      # one bin: a silent no-op
    n_full = oracle.count_in_window(tr, obs, t0, dt, 256)
    n_narrow = oracle.count_in_window(tr, obs, t0 + 60 * dt, dt / 4, 128)
    print("oracle: full %d, narrow %d of %d pairs" % (n_full, n_narrow, obs.shape[0] * 48 * 4000))
    assert 0 < n_narrow < n_full
    assert ctx.count_in_window(s_full) == n_full
    assert ctx.count_in_window(s_narrow) == n_narrow
    assert ctx.count_in_window(s_one) == 0
    assert ctx.count_in_window(-1) == n_full + n_narrow


@pytest.mark.parametrize("n_bins", [256, 250, 97])
def test_screen_spectra_match_oracle(ctx, n_bins):
    say("per-pixel spectra H/V on device (rocFFT) vs the numpy restatement of scripts/teufel.py, %d bins" % n_bins)
    tr, obs, t0, dt, _ = _small_bunch(n_particles=8, npx=8, n_bins=n_bins)
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    sid = ctx.add_screen(obs, t0, dt, n_bins)
    ctx.integrate(TG_FP64)
    ex, ey = np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0])
    f, dh, dv = ctx.screen_spectra(sid, ex, ey)
    fr, rh, rv = oracle.spectral_density(ctx.read_traces(sid), dt, ex, ey)
    assert np.array_equal(f, fr)
    assert dh.shape == rh.shape == (64, n_bins // 2)
    for got, ref in ((dh, rh), (dv, rv)):
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("spill", [1, 2])
def test_mixed_energy_error_long_sums(ctx, spill):
    say("600 e- x 4000 steps in one chunk, 8x8 px, 64 coarse bins: long FP32 sums per cell; mixed vs FP64, "
        "spill=%d" % spill)
    tr, obs, t0, dt, nrm = _small_bunch(n_particles=600, npx=8, n_bins=64)
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    sid = ctx.add_screen(obs, t0, dt, 64)
    ctx.integrate(TG_FP64)
    S64 = ctx.screen_poynting(sid)
    ctx.zero_screen(sid)
    ctx.integrate(TG_MIXED, tuning={"particles_per_chunk": 600, "bins_per_window": 8, "spill": spill})
    Sm = ctx.screen_poynting(sid)
    epsW, excl = energy_rel_error(W.energy_density(Sm, nrm), W.energy_density(S64, nrm))
    print("mixed: eps_W = %.3e (excluded %.3f)" % (epsW, excl))
    assert epsW <= MIXED_ENERGY_TOL


@pytest.mark.parametrize("prec,walk,spill", [(TG_FP64, 2, 1), (TG_MIXED, 2, 1), (TG_FP64, 2, 2), (TG_MIXED, 2, 2),
                                             (TG_FP64, 1, 0), (TG_MIXED, 1, 0)])
def test_persistent_waves_and_lane_offsets_bitwise(ctx, prec, walk, spill, monkeypatch):
    say("few persistent waves (many items each), per-lane offsets off: bitwise equal; prec=%d walk=%d spill=%d"
        % (prec, walk, spill))
    tr, obs, t0, dt, _ = _small_bunch()
    tun = {"bins_per_window": 4, "particles_per_chunk": 7, "window_groups": 3, "walk": walk, "spill": spill}
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    sid = ctx.add_screen(obs, t0, dt, 256)
    ctx.integrate(prec, tuning=tun)
    ref = ctx.read_traces(sid)
    if prec == TG_FP64:
        assert trace_rel_error(ref, oracle.integrate_screen(tr, obs, t0, dt, 256)) <= FP64_TOL
    for env in ({"TG_SLOTS": "3"}, {"TG_LANE_CURSORS": "0"}, {"TG_SLOTS": "1", "TG_LANE_CURSORS": "0"}):
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        with teufel_amd.RadiationContext(0) as c2:
            c2.set_trajectories(tr)
            s2 = c2.add_screen(obs, t0, dt, 256)
            c2.integrate(prec, tuning=tun)
            assert np.array_equal(c2.read_traces(s2), ref), env
        for k in env:
            monkeypatch.delenv(k)


@pytest.mark.parametrize("prec", [TG_FP64, TG_MIXED])
def test_padded_trajectories_add_zeros(ctx, prec):
    say("particles padded by repeating their last point (zero-length steps), prec=%d" % prec)
    tr, obs, t0, dt, _ = _small_bunch(n_particles=6, n_steps=4000, npx=8, n_bins=128)
    cut = [4001, 2800, 4001, 1400, 2, 3601]
    for p, n in enumerate(cut):
        for name in oracle.Trajectories.NAMES:
            a = getattr(tr, name)
            a[p, n:] = a[p, n - 1]
    ref = oracle.integrate_screen(tr, obs, t0, dt, 128)
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    sid = ctx.add_screen(obs, t0, dt, 128)
    ctx.integrate(prec)
    out = ctx.read_traces(sid)
    assert np.all(np.isfinite(out))
    if prec == TG_FP64:
        assert trace_rel_error(out, ref) <= FP64_TOL
    else:
        assert trace_rel_error(out, ref) <= 1e-4


def test_fields_at_matches_oracle(ctx):
    say("single-time field RetardedField(t, x) of a 4-electron bunch at 400 (t, x), all branches")
    w = W.make("c2", n_particles=4, n_steps=600, seed=3)
    u, b = w.und, w.bunch
    tr = oracle.track_undulator(b.x0, b.u0, w.dt, w.n_steps, u.bpeak, u.lambda_u, u.n_periods, u.origin)
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    rng = np.random.default_rng(21)
    n = 400
    pts = np.stack([rng.normal(size=n) * 0.05, rng.normal(size=n) * 0.05, rng.uniform(-1.0, 12.0, n)], axis=1)
    times = rng.uniform(-2e-9, 4.5e-8, n)
    got = ctx.fields_at(times, pts)
    ref = np.zeros((n, 6))
    for i in range(n):
        for p in range(tr.n_particles):
            ref[i] += oracle.field_at(tr, p, times[i], pts[i])
    zero = ~ref.any(axis=1)
    assert zero.any() and (~zero).any()
    assert not got[zero].any()
    assert trace_rel_error(got[~zero][:, None, :], ref[~zero][:, None, :]) <= FP64_TOL
    # accumulate and particle ranges
    half = ctx.fields_at(times, pts, p_range=(0, 2))
    ctx.fields_at(times, pts, p_range=(2, 4), out=half, accumulate=True)
    assert trace_rel_error(half[~zero][:, None, :], ref[~zero][:, None, :]) <= FP64_TOL


def test_gpu_pusher_matches_oracle(ctx):
    say("device Vay pusher vs oracle pusher, 8 e- x 4000 steps")
    w = W.make("c2", n_particles=8, n_steps=4000, seed=5)
    u = w.und
    ctx.track_undulator(w.bunch.x0, w.bunch.u0, w.dt, w.n_steps, u.bpeak, u.lambda_u, u.n_periods, u.origin)
    d = ctx.get_trajectories()
    tr = oracle.track_undulator(w.bunch.x0, w.bunch.u0, w.dt, w.n_steps, u.bpeak, u.lambda_u, u.n_periods,
                                u.origin)
    assert np.array_equal(d["t"], tr.t)
    for n in ("x", "y", "z"):
        assert np.abs(d[n] - getattr(tr, n)).max() <= 1e-12
    for n in ("ux", "uy", "uz"):
        assert np.abs(d[n] - getattr(tr, n)).max() <= 1e-10 * 60
    amax = np.abs(tr.ax).max()
    for n in ("ax", "ay", "az"):
        assert np.abs(d[n] - getattr(tr, n)).max() <= 1e-10 * amax


@pytest.mark.parametrize("prec", [TG_FP64, TG_MIXED])
def test_extreme_shapes_vs_oracle(ctx, prec):
    say("extreme shapes vs oracle: 2-point trajectories, 2 bins, 1 pixel, 65 pixels, 20000 bins; prec=%d" % prec)
    tr, obs, t0, dt, _ = _small_bunch(n_particles=5, n_steps=600, npx=16, n_bins=64)
    cases = [
        (tr.subset(0, 5), obs[:1], t0[:1] + 15.5 * dt, 32 * dt, 2),  # 1 pixel, 2 bins over the pulse
        (tr.subset(0, 5), obs[:65], t0[:65], dt, 64),              # one full tile and one pixel
        (tr.subset(0, 2), obs[:70], t0[:70], dt / 300.0, 20000),   # very long traces, fine bins
    ]
    short = oracle.Trajectories(**{n: np.ascontiguousarray(getattr(tr, n)[:, 300:302])
                                   for n in oracle.Trajectories.NAMES}, charge=tr.charge)
    ts = [oracle.retarded_time(short, p, k, obs[0]) for p in range(5) for k in range(2)]
    cases.append((short, obs[:64], np.full(64, min(ts)), (max(ts) - min(ts)) / 3.0, 4))  # one step per particle
    for trc, o, t0c, dtc, nb in cases:
        ref = oracle.integrate_screen(trc, o, t0c, dtc, nb)
        ctx.remove_screens()
        ctx.set_trajectories(trc)
        sid = ctx.add_screen(o, t0c, dtc, nb)
        ctx.integrate(prec)
        out = ctx.read_traces(sid)
        assert np.abs(ref).max() > 0, (o.shape, nb)
        if prec == TG_FP64:
            assert trace_rel_error(out, ref) <= FP64_TOL, (o.shape, nb)
        else:
            assert trace_rel_error(out, ref) <= 1e-4, (o.shape, nb)


def test_batched_screens_equal_single(ctx):
    say("3 uneven screens batched vs one at a time, FP64, bitwise")
    tr, obs, t0, dt, _ = _small_bunch(n_particles=24, n_steps=2000)
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    geo = [(obs, t0, dt, 256)]
    o2, _, _ = W.grid_screen(7, 0.05, 5.0, tilt_deg=20.0)
    u = W.Undulator()
    t_c, x_c = W._centre(u, 60.0)
    t02, dt2 = W.set_windows(o2, 100, u, 60.0, t_c, x_c, 0.3e-3)
    geo.append((o2, t02, dt2, 100))
    o3, _, _ = W.grid_screen(9, 0.1, 3.0, toroid_f=2.0)
    t03, dt3 = W.set_windows(o3, 513, u, 60.0, t_c, x_c, 0.3e-3)
    geo.append((o3, t03, dt3, 513))
    ids = [ctx.add_screen(*g) for g in geo]
    ctx.integrate(TG_FP64, tuning={"bins_per_window": 8, "particles_per_chunk": 8})
    batched = [ctx.read_traces(i) for i in ids]
    for i, g in zip(ids, geo):
        ctx.zero_screen(i)
    singles = []
    for i in ids:
        ctx.integrate_range(TG_FP64, 0, tr.n_particles, screen_id=i, tuning={"bins_per_window": 8, "particles_per_chunk": 8})
        singles.append(ctx.read_traces(i))
    for a, b_, g in zip(batched, singles, geo):
        assert np.array_equal(a, b_)
        ref = oracle.integrate_screen(tr, g[0], g[1], g[2], g[3])
        assert trace_rel_error(a, ref) <= FP64_TOL


def test_graph_replay_matches_eager(ctx):
    say("hipGraph replays vs eager, mixed")
    tr, obs, t0, dt, _ = _small_bunch(n_particles=16, n_steps=2000)
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    sid = ctx.add_screen(obs, t0, dt, 256)
    ctx.integrate(TG_MIXED)
    eager = ctx.read_traces(sid)
    for rep in range(4):
        ctx.zero_screen(sid)
        ctx.integrate(TG_MIXED, tuning={"use_graph": 1})
        out = ctx.read_traces(sid)
        np.abs(out).sum()  # host work between replays
        assert np.array_equal(out, eager), rep


def test_graph_replay_after_eager_launch_of_another_plan(ctx):
    say("hipGraph capture, eager launches of other plans (FP64, a larger screen), replay: bitwise equal to eager")
    tr, obs, t0, dt, _ = _small_bunch(n_particles=16, n_steps=2000)
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    sid = ctx.add_screen(obs, t0, dt, 256)
    ctx.integrate(TG_MIXED)
    eager = ctx.read_traces(sid)
    ctx.zero_screen(sid)
    ctx.integrate(TG_MIXED, tuning={"use_graph": 1})          # capture + launch
    assert np.array_equal(ctx.read_traces(sid), eager)
    # eager launches of other plans between the replays: another precision,
    # then a plan whose work list and scratch are larger (finer chunks)
    ctx.zero_screen(sid)
    ctx.integrate(TG_FP64)
    ref64 = oracle.integrate_screen(tr, obs, t0, dt, 256)
    assert trace_rel_error(ctx.read_traces(sid), ref64) <= FP64_TOL
    for tun in ({"use_graph": 1}, {"particles_per_chunk": 1, "bins_per_window": 3}, {"use_graph": 1}):
        ctx.zero_screen(sid)
        ctx.integrate(TG_MIXED, tuning=tun)
        out = ctx.read_traces(sid)
        if tun.get("use_graph"):
            assert np.array_equal(out, eager), tun
        else:
            assert trace_rel_error(out, eager) <= 1e-5


@pytest.mark.parametrize("prec", [TG_FP64, TG_MIXED])
def test_graph_replays_over_particle_batches(ctx, prec):
    say("one captured launch replayed over equal particle batches (shifted ranges), prec=%d: bitwise equal to "
        "eager batches; a smaller last batch captures its own launch" % prec)
    tr, obs, t0, dt, _ = _small_bunch(n_particles=27, n_steps=1500)
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    sid = ctx.add_screen(obs, t0, dt, 256)
    batches = [(0, 8), (8, 16), (16, 24), (24, 27), (8, 16)]
    tun = {"particles_per_chunk": 3, "bins_per_window": 6}   # several chunks: partial traces in the graph
    eager = []
    for p0, p1 in batches:
        ctx.zero_screen(sid)
        ctx.integrate_range(prec, p0, p1, tuning=tun)
        eager.append(ctx.read_traces(sid))
    ctx.zero_screen(sid)
    for (p0, p1), ref in zip(batches, eager):
        ctx.zero_screen(sid)
        ctx.integrate_range(prec, p0, p1, tuning=dict(tun, use_graph=1))
        out = ctx.read_traces(sid)
        np.abs(out).sum()  # host work between replays
        assert np.array_equal(out, ref), (p0, p1)
    # the batches accumulate into one store as a streaming caller's would
    ctx.zero_screen(sid)
    for p0, p1 in batches[:4]:
        ctx.integrate_range(prec, p0, p1, tuning=dict(tun, use_graph=1))
    whole = ctx.read_traces(sid)
    ref = oracle.integrate_screen(tr, obs, t0, dt, 256)
    if prec == TG_FP64:
        assert trace_rel_error(whole, ref) <= FP64_TOL
    else:
        assert trace_rel_error(whole, ref) <= 1e-4


def test_autotune_failed_trial_restores_traces(ctx, monkeypatch):
    say("tg_autotune whose 3rd trial fails (TG_AUTOTUNE_FAIL_TRIAL): error reported, traces unchanged")
    tr, obs, t0, dt, _ = _small_bunch(n_particles=16, n_steps=1000)
    monkeypatch.setenv("TG_AUTOTUNE_FAIL_TRIAL", "3")
    with teufel_amd.RadiationContext(0) as c2:
        c2.set_trajectories(tr)
        sid = c2.add_screen(obs, t0, dt, 256)
        c2.integrate(TG_FP64)
        before = c2.read_traces(sid)
        with pytest.raises(teufel_amd.TgError, match="TG_AUTOTUNE_FAIL_TRIAL"):
            c2.autotune(TG_MIXED)
        assert np.array_equal(c2.read_traces(sid), before)
        # the context still works, untuned
        c2.integrate(TG_FP64)
        assert np.allclose(c2.read_traces(sid), 2 * before, rtol=1e-14, atol=0)


@pytest.mark.parametrize("prec", [TG_FP64, TG_MIXED])
def test_autotune_keeps_traces_and_tuned_launch_matches_oracle(ctx, prec):
    say("autotune (window, groups, chunk) on 2 screens, then tuned launch vs oracle, prec=%d" % prec)
    tr, obs, t0, dt, nrm = _small_bunch(n_particles=40, n_steps=2000)
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    s0 = ctx.add_screen(obs, t0, dt, 256)
    o2, n2, _ = W.grid_screen(6, 0.05, 5.0, tilt_deg=15.0)
    u = W.Undulator()
    t_c, x_c = W._centre(u, 60.0)
    t02, dt2 = W.set_windows(o2, 130, u, 60.0, t_c, x_c, 0.3e-3)
    s1 = ctx.add_screen(o2, t02, dt2, 130)
    ctx.integrate(TG_FP64, tuning={"bins_per_window": 8})
    before = [ctx.read_traces(s0), ctx.read_traces(s1)]
    tuned = ctx.autotune(prec)
    print("tuned:", tuned)
    assert 1 <= tuned["bins_per_window"] <= 256
    assert tuned["window_groups"] >= 0 and tuned["particles_per_chunk"] >= 0
    # the timing trials restore every screen's traces
    assert np.array_equal(ctx.read_traces(s0), before[0])
    assert np.array_equal(ctx.read_traces(s1), before[1])
    ctx.zero_screen(s0)
    ctx.zero_screen(s1)
    ctx.integrate(prec)   # tuning=None: the autotuned shape
    def energy(tr_, d_t, nrm_):   # dt sum E x B / mu0 through each pixel
        S = d_t * np.cross(tr_[:, :, 0:3], tr_[:, :, 3:6]).sum(axis=1) / (4e-7 * math.pi)
        return W.energy_density(S, nrm_)

    for sid, (o, t_0, d_t, nb, nrm_) in ((s0, (obs, t0, dt, 256, nrm)), (s1, (o2, t02, dt2, 130, n2))):
        ref = oracle.integrate_screen(tr, o, t_0, d_t, nb)
        out = ctx.read_traces(sid)
        if prec == TG_FP64:
            assert trace_rel_error(out, ref) <= FP64_TOL, sid
        else:
            epsW, _ = energy_rel_error(energy(out, d_t, nrm_), energy(ref, d_t, nrm_))
            print("screen %d: eps_W = %.3e" % (sid, epsW))
            assert epsW <= MIXED_ENERGY_TOL, sid


def test_accumulate_and_noops(ctx):
    say("accumulate semantics, n_bins<2 and NP<2 no-ops")
    tr, obs, t0, dt, _ = _small_bunch(n_particles=8, n_steps=1000)
    ctx.remove_screens()
    ctx.set_trajectories(tr)
    sid = ctx.add_screen(obs, t0, dt, 64)
    s1 = ctx.add_screen(obs[:5], t0[:5], dt, 1)
    ctx.integrate(TG_FP64)
    a = ctx.read_traces(sid)
    ctx.integrate(TG_FP64)
    b = ctx.read_traces(sid)
    assert np.allclose(b, 2 * a, rtol=1e-14, atol=0)
    host = np.ones_like(a)
    ctx.read_traces(sid, out=host, accumulate=True)
    assert np.array_equal(host, 1.0 + b)
    assert ctx.read_traces(s1).shape == (5, 1, 6)
    one = oracle.Trajectories.empty(3, 1)
    ctx.set_trajectories(one)
    ctx.zero_screen(sid)
    ctx.integrate(TG_FP64)
    assert not ctx.read_traces(sid).any()


def test_errors_are_reported(ctx):
    say("argument errors map to TgError")
    with pytest.raises(teufel_amd.TgError):
        ctx.add_screen(np.zeros((2, 3)), np.zeros(2), -1.0, 10)
    with pytest.raises(teufel_amd.TgError):
        ctx.add_screen(np.array([[0, 0, math.nan]]), np.zeros(1), 1e-15, 10)
    with pytest.raises(teufel_amd.TgError):
        ctx._chk(ctx.lib.tg_zero_screen(ctx.h, 9999))
    # more points per particle than the kernel's 32-bit record offsets allow
    # (no particles, so no array is read)
    with pytest.raises(teufel_amd.TgError, match="points per particle"):
        ctx._chk(ctx.lib.tg_set_trajectories(ctx.h, 0, 23_000_000, *([None] * 10), None))


def test_unphysical_trajectories_are_rejected():
    """The walks rely on the retarded time increasing along every trajectory,
    which holds for any particle slower than light.  A step faster than light
    or back in time makes integrate fail with TgError in both precisions
    (the pack kernels check every step); repeated points (zero-length steps)
    are accepted.  A beam far off the z axis is rejected by TG_MIXED only."""
    say("unphysical trajectories: faster-than-light and backward steps rejected, repeated points accepted")
    n, k = 3, 5
    base = oracle.Trajectories.empty(n, k)
    base.t[:] = np.arange(k) * 1e-12
    base.z[:] = np.arange(k) * 2.99e-4            # beta = 0.9973 along z
    base.uz[:] = 14.0
    obs, _, _ = W.grid_screen(4, 0.02, 10.0)
    t0 = np.linalg.norm(obs, axis=1) / W.C_LIGHT - 1e-15

    def variant(change=()):
        """A copy of base with entries (array name, particle, point) set."""
        tr = oracle.Trajectories(**{nm: getattr(base, nm).copy() for nm in oracle.Trajectories.NAMES},
                                 charge=base.charge.copy())
        for name, p, j, v in change:
            getattr(tr, name)[p, j] = v
        return tr

    cases = {
        "superluminal": variant([("x", 1, 3, 0.5)]),          # a 0.5 m jump within 1 ps
        "backward": variant([("t", 2, 2, 0.5e-12), ("t", 2, 3, 0.4e-12)]),
    }
    with teufel_amd.RadiationContext(0) as c:
        sid = c.add_screen(obs, t0, 1e-15, 64)
        for name, tr in cases.items():
            c.set_trajectories(tr)
            for prec in (TG_FP64, TG_MIXED):
                with pytest.raises(teufel_amd.TgError, match="faster than light"):
                    c.integrate(prec)
        # repeated points: a particle resting for a step (dt = 0, dX = 0)
        rest = variant([("t", 0, 4, 3e-12), ("z", 0, 4, 3 * 2.99e-4)])
        c.set_trajectories(rest)
        c.zero_screen(sid)
        c.integrate(TG_FP64)
        ref = oracle.integrate_screen(rest, obs, t0, 1e-15, 64)
        assert np.abs(ref).max() > 0
        assert trace_rel_error(c.read_traces(sid), ref) <= FP64_TOL
        # a beam far off the z axis: FP64 only
        off = variant()
        off.ux[:] = 10.0
        c.set_trajectories(off)
        c.zero_screen(sid)
        with pytest.raises(teufel_amd.TgError, match="close to the z axis"):
            c.integrate(TG_MIXED)
        c.integrate(TG_FP64)


def test_electrostatic_single_time_field():
    """The reference's tests/teufel.electrostatic.cpp through tg_fields_at
    (ChargedParticle::RetardedField(t, x), src/particle.cpp:612-769), with its
    closed forms and tolerances:
    * -1 pC at rest (a one-point trajectory): the Coulomb field,
      8987.55 V/m at 1 mm, along +-x (1e-3 V/m) and along z (1e-6 relative);
    * the same charge at gamma = 10 in direction (2, 3, 5), pushed through a
      field-free region for 4 ns (2000 Vay steps): at t = 0, 1 mm across the
      direction of motion, before the trajectory's light cone (the back-
      extrapolation), E and B of the uniformly moving charge,
      gamma E_Coulomb across and v x E / c^2 (1e-5 relative); at t = 3 ns on
      the trajectory (the interpolation and its refinement) the same within
      1e-3, and the transverse E along 11 points of the line of motion equal
      to its values at t = 0 (1e-5 relative);
    and every value against the oracle (1e-12 of the field)."""
    say("electrostatic test case: Coulomb at rest, uniformly moving charge at t = 0 and t = 3 ns")
    e_charge, eps0, c_light = 1.6021766208e-19, 8.854187817e-12, 2.99792458e8
    mu0 = 4e-7 * math.pi
    q = -1.0e-12 / e_charge
    e_ref = 1e-12 / (4.0 * math.pi * eps0 * 1.0e-3 * 1.0e-3)
    with teufel_amd.RadiationContext(0) as c:
        rest = oracle.Trajectories.empty(1, 1, q)
        c.set_trajectories(rest)
        pts = np.array([[1e-3, 0.0, 0.0], [-1e-3, 0.0, 0.0], [0.0, 0.0, 1e-3]])
        f = c.fields_at(np.zeros(3), pts)
        assert np.linalg.norm(f[0, :3] - [-e_ref, 0, 0]) < 1e-3
        assert np.linalg.norm(f[1, :3] - [e_ref, 0, 0]) < 1e-3
        assert np.linalg.norm(f[2, :3] - [0, 0, -e_ref]) / e_ref < 1e-6
        for i in range(3):
            assert np.linalg.norm(f[i] - oracle.field_at(rest, 0, 0.0, pts[i])) <= 1e-12 * e_ref

        gamma = 10.0
        beta = math.sqrt(1.0 - 1.0 / (gamma * gamma))
        bg = math.sqrt(gamma * gamma - 1.0)
        d = np.array([2.0, 3.0, 5.0]) / math.sqrt(38.0)
        perp = np.cross(d, [1.0, 0.0, 0.0])
        perp /= np.linalg.norm(perp)
        c.track_undulator([[0.0, 0.0, 0.0]], [d * bg], 4.0e-9 / 2000, 2000, 0.0, 0.3, 1, [0.0, 0.0, 0.0],
                          charge=q, mass=1.0)
        tr = oracle.Trajectories(**c.get_trajectories())
        e_v = perp * e_ref * -1.0 * gamma
        b_v = np.cross(d * beta * c_light, perp) * -1.0e-12 * mu0 / (4 * math.pi) / (1e-3 * 1e-3) * gamma

        def rel(a, b):
            return np.linalg.norm(a - b) / np.linalg.norm(b)

        o0 = perp * 1e-3
        t_obs = 3.0e-9
        oT = d * t_obs * beta * c_light + perp * 1e-3
        line = [-0.005 + i * 0.001 for i in range(11)]
        times = np.array([0.0, t_obs] + [0.0] * 11 + [t_obs] * 11)
        pts = np.array([o0, oT] + [o0 + d * l for l in line] + [oT + d * l for l in line])
        f = c.fields_at(times, pts)
        assert rel(f[0, :3], e_v) < 1e-5 and rel(f[0, 3:], b_v) < 1e-5
        assert rel(f[1, :3], e_v) < 1e-3 and rel(f[1, 3:], b_v) < 1e-3
        tr0 = f[2:13, :3] @ perp
        trT = f[13:24, :3] @ perp
        assert np.all((trT - tr0) / tr0 <= 1e-5)
        for i in range(len(times)):
            ref = oracle.field_at(tr, 0, times[i], pts[i])
            assert np.linalg.norm(f[i, :3] - ref[:3]) <= 1e-12 * np.linalg.norm(ref[:3])
            assert np.linalg.norm(f[i, 3:] - ref[3:]) <= 1e-12 * np.linalg.norm(ref[3:])


def test_undulator_trajectory_closed_forms():
    """The reference's tests/teufel.undulator.cpp through tg_track_undulator:
    one electron at gamma = 30.358 from the origin along z through a planar
    undulator of 54 periods of 37 mm at 0.532 T centred at z = 1.5 m, 3000
    Vay steps over the time that brings it to z = 3.0 m (the slippage of one
    radiation wavelength per period).  With the reference test's tolerances:
    the final z within 1e-5 m of 3.0 m, the largest x inside the undulator
    within 1e-5 m of e B/(beta gamma m c k_u^2), the final time within one
    step of that time, the final x within 1e-5 m of 0; and the positions
    within 1e-12 m of the oracle's pusher."""
    say("undulator test case: 54 x 37 mm at 0.532 T, gamma 30.358, 3000 steps")
    e_charge, c_light = 1.6021766208e-19, 2.99792458e8
    B, lam, N = 0.532, 0.037, 54
    ku = 2 * math.pi / lam
    gamma = 30.35823388162631
    beta = math.sqrt(1.0 - 1.0 / (gamma * gamma))
    bg = math.sqrt(gamma * gamma - 1.0)
    K = 0.934 * 100 * lam * B
    lam_r = (lam / (2 * gamma * gamma)) * (1 + K * K / 2)
    xdis = e_charge * B * c_light / (bg * 9.1e-31 * c_light * c_light * ku * ku)
    tau = N * (lam + lam_r) / c_light + (3.0 - N * lam) / (beta * c_light)
    steps = 3000
    with teufel_amd.RadiationContext(0) as c:
        c.track_undulator([[0.0, 0.0, 0.0]], [[0.0, 0.0, bg]], tau / steps, steps, B, lam, N, [0.0, 0.0, 1.5],
                          charge=-1.0, mass=1.0)
        tr = c.get_trajectories()
    x, z, t = tr["x"][0], tr["z"][0], tr["t"][0]
    inside = (z > lam) & (z < N * lam - lam)
    assert abs(z[-1] - 3.0) <= 1e-5
    assert abs(x[inside].max() - xdis) <= 1e-5
    assert abs(t[-1] - tau) <= tau / steps
    assert abs(x[-1]) <= 1e-5
    ref = oracle.track_undulator([[0.0, 0.0, 0.0]], [[0.0, 0.0, bg]], tau / steps, steps, B, lam, N, [0.0, 0.0, 1.5],
                                 charge=-1.0, mass=1.0)
    # positions within 1e-12 m, as test_gpu_pusher_matches_oracle (the
    # undulator field's sin/cos/cosh differ from the host's by rounding)
    assert np.abs(x - ref.x[0]).max() <= 1e-12
    assert np.abs(z - ref.z[0]).max() <= 1e-12


def _track_lattice_gpu(x0, u0, dt, n_steps, **kw):
    with teufel_amd.RadiationContext(0) as c:
        c.track_lattice(x0, u0, dt, n_steps, **kw)
        return c.get_trajectories()


def test_lattice_pusher_homogeneous_fields():
    """The reference's tests/teufel.magnet.cpp and tests/teufel.EcrossB.cpp
    through tg_track_lattice (HomogeneousField, src/fields.cpp:155-163).
    Magnet: gamma = 10 across 0.1 T, one revolution in 1000 steps; the half-
    way point 2 R_gyro from the start (1e-5 m), the final time (0.01 ps), the
    return to the start (1e-5 m), |u| (1e-4).  E x B: gamma = 30.358 in
    crossed fields that cancel, 3000 steps over 10 ns; the final time (one
    step), the straight line v t (2e-6 m), |u| (1.5e-3).  Both are the
    reference tests' own bounds.  Every operation of the pusher in a
    homogeneous field is correctly rounded, so the trajectories equal the
    oracle's (which equals the compiled reference's) bitwise."""
    B, u0, dt, n, r_gyro, tau, bg = magnet_case()
    say("magnet test case: gamma 10 in |B| = 0.1 T, %d steps" % n)
    tr = _track_lattice_gpu([[0.0, 0.0, 0.0]], [u0], dt, n, B0=B)
    X = np.stack([tr["x"][0], tr["y"][0], tr["z"][0]], axis=1)
    assert abs(0.5 * np.linalg.norm(X[n // 2]) - r_gyro) < 1e-5
    assert 1e12 * abs(tr["t"][0, -1] - tau) < 0.01
    assert np.linalg.norm(X[-1]) < 1e-5
    assert abs(math.sqrt(tr["ux"][0, -1] ** 2 + tr["uy"][0, -1] ** 2 + tr["uz"][0, -1] ** 2) - bg) < 1e-4
    ref = oracle.track_lattice([[0.0, 0.0, 0.0]], [u0], dt, n, B0=B)
    for name in oracle.Trajectories.NAMES:
        assert np.array_equal(tr[name], getattr(ref, name)), name

    E, B, u0, v, dt, n, tau = exb_case()
    say("E x B test case: gamma 30.358 in crossed fields, %d steps" % n)
    tr = _track_lattice_gpu([[0.0, 0.0, 0.0]], [u0], dt, n, E0=E, B0=B)
    assert abs(tr["t"][0, -1] - tau) < dt
    Xf = np.array([tr["x"][0, -1], tr["y"][0, -1], tr["z"][0, -1]])
    assert np.abs(Xf - v * tau).max() < 2e-6
    U = np.stack([tr["ux"][0], tr["uy"][0], tr["uz"][0]], axis=1)
    assert np.abs(np.linalg.norm(U, axis=1) - np.linalg.norm(u0)).max() < 1.5e-3
    ref = oracle.track_lattice([[0.0, 0.0, 0.0]], [u0], dt, n, E0=E, B0=B)
    for name in oracle.Trajectories.NAMES:
        assert np.array_equal(tr[name], getattr(ref, name)), name


def test_lattice_pusher_undulator_with_homogeneous_field():
    """An undulator and a homogeneous E, B in one lattice (Lattice::Field,
    src/fields.cpp:229-239, the undulator's field first): 16 electrons of the
    c2 bunch, 600 steps, against the oracle's pusher within 1e-12 m and
    1e-12 relative in u (the undulator's sin/cos/cosh differ from the host's
    by rounding); without the homogeneous element tg_track_lattice equals
    tg_track_undulator bitwise."""
    w = W.make("c2", n_particles=16, n_steps=600)
    u, b = w.und, w.bunch
    und = (u.bpeak, u.lambda_u, u.n_periods, u.origin)
    E0, B0 = [3e4, -1e4, 2e3], [1e-3, -2e-3, 5e-4]
    say("undulator + homogeneous field: 16 electrons, 600 steps")
    tr = _track_lattice_gpu(b.x0, b.u0, w.dt, w.n_steps, undulator=und, E0=E0, B0=B0)
    ref = oracle.track_lattice(b.x0, b.u0, w.dt, w.n_steps, undulator=und, E0=E0, B0=B0)
    plain = oracle.track_undulator(b.x0, b.u0, w.dt, w.n_steps, *und)
    assert np.abs(ref.x - plain.x).max() > 1e-6          # the homogeneous field moves the beam
    for name in ("x", "y", "z"):
        assert np.abs(tr[name] - getattr(ref, name)).max() <= 1e-12, name
    assert np.array_equal(tr["t"], ref.t)
    for name in ("ux", "uy", "uz"):
        assert np.abs(tr[name] - getattr(ref, name)).max() <= 1e-12 * np.abs(ref.uz).max(), name
    with teufel_amd.RadiationContext(0) as c:
        c.track_undulator(b.x0, b.u0, w.dt, w.n_steps, *und)
        a = c.get_trajectories()
    lat = _track_lattice_gpu(b.x0, b.u0, w.dt, w.n_steps, undulator=und)
    for name in oracle.Trajectories.NAMES:
        assert np.array_equal(a[name], lat[name]), name
