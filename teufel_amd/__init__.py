"""teufel_amd - MI355X (gfx950) Lienard-Wiechert radiation step of TEUFEL.

Python access to the C ABI (include/teufel_gpu.h) for tests and benchmarks.
The C++ host mirror of the reference API (ChargedParticle, Bunch, Beam,
MeshedScreen, ScreenObserver, PointObserver) lives in teufel_amd/host/.

Every compute call goes to the HIP library; if it is not built, or no GPU is
present, the call raises - there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _native
from ._native import TG_FP64, TG_MIXED, TgTuning

__all__ = ["TG_FP64", "TG_MIXED", "TgTuning", "TgError", "RadiationContext", "load_library"]

_ERRNAMES = {1: "ARG", 2: "OOM", 3: "HIP", 4: "RANGE", 5: "STATE"}


class TgError(IOError):
    """Non-zero status of a tg_* call (the reference raises IOexception)."""

    def __init__(self, code, msg):
        super().__init__("TG_ERR_%s (%d): %s" % (_ERRNAMES.get(code, "?"), code, msg))
        self.code = code


def load_library():
    return _native.load()


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a, shape=None):
    """An input array as C-contiguous float64 (a copy where needed)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _check_shape(a, shape, name):
    if a.shape != tuple(shape):
        raise ValueError("%s: shape %s, expected %s" % (name, a.shape, tuple(shape)))
    return a


def _out(a, shape, name):
    """An output buffer handed to the library: it is written in place, so it
    must already be C-contiguous float64 of exactly the expected shape (a
    converted copy would silently receive the result instead)."""
    if not isinstance(a, np.ndarray):
        raise ValueError("%s: expected a numpy array, got %s" % (name, type(a).__name__))
    if a.dtype != np.float64:
        raise ValueError("%s: dtype %s, expected float64" % (name, a.dtype))
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("%s: array is not C-contiguous" % name)
    if not a.flags["WRITEABLE"]:
        raise ValueError("%s: array is read-only" % name)
    return _check_shape(a, shape, name)


class RadiationContext:
    """One device context: trajectories, screens and their on-device traces."""

    def __init__(self, device=0):
        self.lib = _native.load()
        h = ctypes.c_void_p()
        rc = self.lib.tg_create(device, ctypes.byref(h))
        if rc != 0:
            raise TgError(rc, "tg_create(%d) failed (is a GPU present?)" % device)
        self.h = h
        self.screens = []  # (n_pix, n_bins)

    # ---- plumbing ----
    def _chk(self, rc):
        if rc != 0:
            raise TgError(rc, self.lib.tg_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.tg_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        buf = ctypes.create_string_buffer(256)
        n = ctypes.c_int()
        self._chk(self.lib.tg_device_info(self.h, buf, 256, ctypes.byref(n)))
        return buf.value.decode(), n.value

    # ---- trajectories ----
    def set_trajectories(self, tr):
        """tr: object with t,x,y,z,ux,uy,uz,ax,ay,az (n_particles, n_points) and charge."""
        names = ("t", "x", "y", "z", "ux", "uy", "uz", "ax", "ay", "az")
        arrs = [_f64(getattr(tr, n)) for n in names]
        if arrs[0].ndim == 1:
            arrs = [a.reshape(1, -1) if a.ndim == 1 else a for a in arrs]
        if arrs[0].ndim != 2:
            raise ValueError("t: expected (n_particles, n_points), got shape %s" % (arrs[0].shape,))
        n_p, n_pts = arrs[0].shape
        for n, a in zip(names, arrs):
            _check_shape(a, (n_p, n_pts), n)
        q = _check_shape(_f64(tr.charge).reshape(-1), (n_p,), "charge")
        self._keep = arrs + [q]
        self._chk(self.lib.tg_set_trajectories(self.h, n_p, n_pts, *[_dptr(a) for a in arrs], _dptr(q)))

    def track_undulator(self, x0, u0, dt, n_steps, bpeak, lambda_u, n_periods, origin,
                        charge=-1.0, mass=1.0, t_start=0.0):
        x0 = _f64(x0, (-1, 3))
        u0 = _check_shape(_f64(u0, (-1, 3)), x0.shape, "u0")
        o = _check_shape(_f64(origin).reshape(-1), (3,), "origin")
        self._chk(self.lib.tg_track_undulator(self.h, x0.shape[0], _dptr(x0), _dptr(u0), t_start, charge,
                                              mass, dt, n_steps, bpeak, lambda_u, n_periods, _dptr(o)))

    def track_lattice(self, x0, u0, dt, n_steps, undulator=None, E0=None, B0=None,
                      charge=-1.0, mass=1.0, t_start=0.0):
        """Vay pusher through a lattice of an undulator (bpeak, lambda_u,
        n_periods, origin) and a homogeneous field E0 [V/m], B0 [T]; any of
        the three may be None."""
        x0 = _f64(x0, (-1, 3))
        u0 = _check_shape(_f64(u0, (-1, 3)), x0.shape, "u0")
        if undulator is not None:
            bpeak, lambda_u, n_periods, origin = undulator
            oa = _check_shape(_f64(origin).reshape(-1), (3,), "origin")
            o = _dptr(oa)
        else:
            bpeak, lambda_u, n_periods, o = 0.0, 0.0, 0, None
        e = None if E0 is None else _check_shape(_f64(E0).reshape(-1), (3,), "E0")
        b = None if B0 is None else _check_shape(_f64(B0).reshape(-1), (3,), "B0")
        self._chk(self.lib.tg_track_lattice(self.h, x0.shape[0], _dptr(x0), _dptr(u0), t_start, charge, mass, dt,
                                            n_steps, bpeak, lambda_u, n_periods, o,
                                            None if e is None else _dptr(e), None if b is None else _dptr(b)))

    def trajectory_shape(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self._chk(self.lib.tg_trajectory_shape(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def get_trajectories(self, p_begin=0, p_end=None):
        n_p, n_pts = self.trajectory_shape()
        p_end = n_p if p_end is None else p_end
        out = {n: np.zeros((p_end - p_begin, n_pts)) for n in
               ("t", "x", "y", "z", "ux", "uy", "uz", "ax", "ay", "az")}
        out["charge"] = np.zeros(p_end - p_begin)
        names = ("t", "x", "y", "z", "ux", "uy", "uz", "ax", "ay", "az", "charge")
        self._chk(self.lib.tg_get_trajectories(self.h, p_begin, p_end, *[_dptr(out[n]) for n in names]))
        return out

    # ---- screens ----
    def add_screen(self, obs, t0, dt, n_bins):
        obs = _f64(obs)
        if obs.ndim != 2 or obs.shape[1] != 3:
            raise ValueError("obs: expected (n_pix, 3), got shape %s" % (obs.shape,))
        t0 = _check_shape(_f64(t0).reshape(-1), (obs.shape[0],), "t0")
        sid = ctypes.c_int()
        self._chk(self.lib.tg_add_screen(self.h, obs.shape[0], _dptr(obs), _dptr(t0), float(dt), int(n_bins),
                                         ctypes.byref(sid)))
        self.screens.append((obs.shape[0], int(n_bins)))
        return sid.value

    def zero_screen(self, sid):
        self._chk(self.lib.tg_zero_screen(self.h, sid))

    def remove_screens(self):
        self._chk(self.lib.tg_remove_screens(self.h))
        self.screens = []

    # ---- integration ----
    @staticmethod
    def _tuning(tuning):
        if tuning is None:
            return None
        if isinstance(tuning, TgTuning):
            return ctypes.byref(tuning)
        t = TgTuning(**tuning)
        return ctypes.byref(t)

    def integrate(self, precision=TG_FP64, tuning=None):
        self._chk(self.lib.tg_integrate(self.h, precision, self._tuning(tuning)))

    def integrate_range(self, precision, p_begin, p_end, screen_id=-1, tuning=None):
        self._chk(self.lib.tg_integrate_range(self.h, precision, self._tuning(tuning), screen_id, p_begin, p_end))

    def integrate_range_static(self, precision, p_begin, p_end, screen_id=-1, tuning=None):
        """tg_integrate_range with its launch scratch in the library's static
        device array (no device allocation; see include/teufel_gpu.h)."""
        self._chk(self.lib.tg_integrate_range_static(self.h, precision, self._tuning(tuning), screen_id, p_begin,
                                                     p_end))

    def synchronize(self):
        self._chk(self.lib.tg_synchronize(self.h))

    def last_timing(self):
        k, t = ctypes.c_float(), ctypes.c_float()
        self._chk(self.lib.tg_last_timing(self.h, ctypes.byref(k), ctypes.byref(t)))
        return k.value, t.value

    def last_stats(self):
        v = [ctypes.c_longlong() for _ in range(3)]
        self._chk(self.lib.tg_last_stats(self.h, *[ctypes.byref(x) for x in v]))
        return {"wave_steps": v[0].value, "walks": v[1].value, "work_items": v[2].value}

    def autotune(self, precision=TG_MIXED):
        t = TgTuning()
        self._chk(self.lib.tg_autotune(self.h, precision, ctypes.byref(t)))
        return {"bins_per_window": t.bins_per_window, "particles_per_chunk": t.particles_per_chunk,
                "window_groups": t.window_groups, "use_graph": t.use_graph, "spill": t.spill}

    def read_traces(self, sid, out=None, accumulate=False):
        n_pix, n_bins = self.screens[sid]
        if out is None:
            out = np.zeros((n_pix, n_bins, 6))
            accumulate = False
        _out(out, (n_pix, n_bins, 6), "out")
        self._chk(self.lib.tg_read_traces(self.h, sid, _dptr(out), 1 if accumulate else 0))
        return out

    def screen_poynting(self, sid):
        n_pix, _ = self.screens[sid]
        out = np.zeros((n_pix, 3))
        self._chk(self.lib.tg_screen_poynting(self.h, sid, _dptr(out)))
        return out

    def screen_spectra(self, sid, ex, ey):
        """Spectral power density [J/(m² Hz)] of the horizontal (E.ex, B.ey)
        and vertical (E.ey, B.ex) polarisations of every pixel, as
        spectralDensityH/V of scripts/teufel.py:198-226; returns
        (f[n_bins//2], dens_h[n_pix, n_bins//2], dens_v[n_pix, n_bins//2])."""
        n_pix, n_bins = self.screens[sid]
        nh = n_bins // 2
        ex = _check_shape(_f64(ex).reshape(-1), (3,), "ex")
        ey = _check_shape(_f64(ey).reshape(-1), (3,), "ey")
        f = np.zeros(max(nh, 1))
        dh = np.zeros((n_pix, max(nh, 1)))
        dv = np.zeros((n_pix, max(nh, 1)))
        self._chk(self.lib.tg_screen_spectra(self.h, sid, _dptr(ex), _dptr(ey), _dptr(f), _dptr(dh), _dptr(dv)))
        return f[:nh], dh[:, :nh], dv[:, :nh]

    def count_in_window(self, sid=-1):
        """(pixel, particle step) pairs passing the window test ts2 > tmin and
        ts1 < tmax of src/particle.cpp:793, for screen sid (-1: all)."""
        n = ctypes.c_longlong(0)
        self._chk(self.lib.tg_count_in_window(self.h, sid, ctypes.byref(n)))
        return n.value

    def fields_at(self, times, points, p_range=None, out=None, accumulate=False):
        """Field (E, B) at points[i] at times[i] summed over the particles
        (ChargedParticle/Bunch/Beam::RetardedField(t, x)); returns (n, 6)."""
        points = _f64(points)
        if points.ndim == 1:
            points = points.reshape(1, -1)
        if points.ndim != 2 or points.shape[1] != 3:
            raise ValueError("points: expected (n, 3), got shape %s" % (points.shape,))
        times = np.asarray(times, dtype=np.float64)
        if times.ndim == 0:
            times = np.full(points.shape[0], float(times))
        times = _check_shape(_f64(times), (points.shape[0],), "times")
        if out is None:
            out = np.zeros((points.shape[0], 6))
            accumulate = False
        _out(out, (points.shape[0], 6), "out")
        n_particles, _ = self.trajectory_shape()
        p0, p1 = p_range if p_range is not None else (0, n_particles)
        self._chk(self.lib.tg_fields_at(self.h, points.shape[0], _dptr(times), _dptr(points), int(p0), int(p1),
                                        _dptr(out), 1 if accumulate else 0))
        return out

    def deposit_only(self, ts, fields, t0, dt, n_bins, trace=None):
        ts = _f64(ts)
        fields = _f64(fields)
        if ts.ndim == 1:
            ts = ts[None]
        if fields.ndim == 2:
            fields = fields[None]
        if ts.ndim != 2:
            raise ValueError("ts: expected (n_pseudo, n_points), got shape %s" % (ts.shape,))
        _check_shape(fields, ts.shape + (6,), "fields")
        if trace is None:
            trace = np.zeros((max(int(n_bins), 0), 6))
        _out(trace, (max(int(n_bins), 0), 6), "trace")
        self._chk(self.lib.tg_deposit_only(self.h, ts.shape[0], ts.shape[1], _dptr(ts), _dptr(fields),
                                           float(t0), float(dt), int(n_bins), _dptr(trace)))
        return trace
