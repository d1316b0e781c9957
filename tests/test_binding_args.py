"""The Python binding rejects malformed arrays before they reach the C ABI.

The library reads and writes host buffers with sizes taken from the context
(trace shapes, trajectory shape) or from other arguments, so a short, float32
or non-contiguous array would be read or written past its end.  These checks
run on the CPU: the context below has no device handle and a library stub
that fails the test if any call reaches it.
"""
import numpy as np
import pytest

import teufel_amd


class _NoLib:
    def __getattr__(self, name):
        raise AssertionError("library call %s reached with invalid arguments" % name)


def _ctx(screens=((4, 8),), shape=(3, 10)):
    c = teufel_amd.RadiationContext.__new__(teufel_amd.RadiationContext)
    c.lib = _NoLib()
    c.h = None
    c.screens = list(screens)
    c.trajectory_shape = lambda: shape
    return c


class _Traj:
    def __init__(self, n_p=3, n_pts=10, **over):
        for n in ("t", "x", "y", "z", "ux", "uy", "uz", "ax", "ay", "az"):
            setattr(self, n, np.zeros((n_p, n_pts)))
        self.charge = -np.ones(n_p)
        for k, v in over.items():
            setattr(self, k, v)


@pytest.mark.parametrize("bad", [
    np.zeros((4, 8, 6), dtype=np.float32),            # wrong dtype
    np.zeros((4, 8, 5)),                              # short
    np.zeros((4, 16, 6))[:, ::2, :],                  # non-contiguous view
    np.asfortranarray(np.zeros((4, 8, 6))),           # Fortran order
    np.zeros(4 * 8 * 6).tolist(),                     # not an array
])
def test_read_traces_rejects_bad_out(bad):
    with pytest.raises(ValueError):
        _ctx().read_traces(0, out=bad, accumulate=True)


def test_read_traces_rejects_read_only_out():
    a = np.zeros((4, 8, 6))
    a.flags.writeable = False
    with pytest.raises(ValueError, match="read-only"):
        _ctx().read_traces(0, out=a)


@pytest.mark.parametrize("field,value", [
    ("x", np.zeros((3, 9))),          # shorter than t
    ("az", np.zeros((2, 10))),        # fewer particles
    ("charge", -np.ones(2)),          # charge per particle
    ("t", np.zeros((2, 3, 4))),       # not 2-D
])
def test_set_trajectories_rejects_mismatched_arrays(field, value):
    with pytest.raises(ValueError):
        _ctx().set_trajectories(_Traj(**{field: value}))


def test_add_screen_rejects_mismatched_geometry():
    with pytest.raises(ValueError):
        _ctx().add_screen(np.zeros((5, 3)), np.zeros(4), 1e-15, 16)
    with pytest.raises(ValueError):
        _ctx().add_screen(np.zeros((5, 2)), np.zeros(5), 1e-15, 16)


def test_fields_at_rejects_bad_arrays():
    c = _ctx()
    with pytest.raises(ValueError):
        c.fields_at(np.zeros(3), np.zeros((4, 3)))                      # times vs points
    with pytest.raises(ValueError):
        c.fields_at(np.zeros(4), np.zeros((4, 3)), out=np.zeros((4, 6), dtype=np.float32))
    with pytest.raises(ValueError):
        c.fields_at(np.zeros(4), np.zeros((4, 3)), out=np.zeros((3, 6)))
    with pytest.raises(ValueError):
        c.fields_at(np.zeros(4), np.zeros((4, 2)))


def test_deposit_only_rejects_bad_arrays():
    c = _ctx()
    with pytest.raises(ValueError):
        c.deposit_only(np.zeros((2, 5)), np.zeros((2, 4, 6)), 0.0, 1.0, 8)
    with pytest.raises(ValueError):
        c.deposit_only(np.zeros((2, 5)), np.zeros((2, 5, 6)), 0.0, 1.0, 8, trace=np.zeros((7, 6)))
    with pytest.raises(ValueError):
        c.deposit_only(np.zeros((2, 5)), np.zeros((2, 5, 6)), 0.0, 1.0, 8,
                       trace=np.zeros((8, 6), dtype=np.float32))


def test_track_and_spectra_reject_bad_vectors():
    c = _ctx()
    with pytest.raises(ValueError):
        c.track_undulator(np.zeros((4, 3)), np.zeros((3, 3)), 1e-12, 10, 0.03, 0.3, 8, (0, 0, 1.5))
    with pytest.raises(ValueError):
        c.track_undulator(np.zeros((4, 3)), np.zeros((4, 3)), 1e-12, 10, 0.03, 0.3, 8, (0, 0))
    with pytest.raises(ValueError):
        c.track_lattice(np.zeros((4, 3)), np.zeros((4, 3)), 1e-12, 10, E0=(1.0, 2.0))
    with pytest.raises(ValueError):
        c.track_lattice(np.zeros((4, 3)), np.zeros((4, 3)), 1e-12, 10, B0=np.zeros(4))
    with pytest.raises(ValueError):
        c.track_lattice(np.zeros((4, 3)), np.zeros((4, 3)), 1e-12, 10, undulator=(0.03, 0.3, 8, (0, 0)))
    with pytest.raises(ValueError):
        c.screen_spectra(0, np.array([1.0, 0.0]), np.array([0.0, 1.0, 0.0]))
