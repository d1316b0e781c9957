"""Synthetic inputs of the radiation step: undulator, bunches, screens, windows.

Pure numpy (no GPU needed).  Counter-based random numbers (SplitMix64 +
Box-Muller) make every sample a function of (seed, stream, index) only, so the
same bunch comes out on any machine and in any order.

Workloads (all electrons, charge -1 e unless stated; planar undulator
K = 1, lambda_u = 0.3 m, field region (N+1) lambda_u with one-period linear end
ramps as src/undulator.cpp:107-126 builds it):

  c1  1 electron, gamma 60 on axis, N = 8, 100 steps per period (1000 steps);
      flat 21 x 21 screen, 0.1 x 0.1 m at z = 10 m, 256 bins
  c2  1e4 electrons: gamma ~ N(60, 0.06), x,y ~ N(0, 0.5 mm), x',y' ~ N(0, 50 urad),
      z ~ U(-0.15, 0.15) mm; 4000 steps; 64 x 64 screen 0.2 x 0.2 m at 10 m, 1024 bins
  c3  as c2 with 1e5 electrons, 128 x 128 screen, 2048 bins
  c4  32 screens (flat / tilted 0-30 deg / toroidal f = 1-5 m), 16^2..128^2 pixels,
      256..4096 bins, 2..20 m; 2e4 electrons of the c2 distribution
  c5  2e5 macro-electrons of -6241.5 e (200 pC), Gaussian sigma_z = 0.1 mm,
      N = 34, 4000 steps, 128 x 128 screen 10 m behind the undulator, 4096 bins

The undulator field of c1-c4 spans z = 0.15 .. 2.85 m (centre 1.5 m),
particles start at z ~ 0 and travel 3.0 m.  Windows follow the arrival-time
curve per pixel (MeshedScreen-style per-pixel t0, src/mesh.cpp:169-175):
t0_p = T_p - W/2 + dt/2 with T_p = t_c + |x_p - x_c|/c the arrival of the
emission from the undulator centre, W = 1.3 [(N+1) lambda_u (1 + K^2/2 +
gamma^2 theta_max^2) / (2 gamma^2 c) + L_b / c], dt = W / n_bins.
"""
import math
from dataclasses import dataclass, field

import numpy as np

MEC2_EV = 0.5109989461e6          # src/global.h:38
C_LIGHT = 2.99792458e8            # src/global.h:39
E_CHARGE = 1.6021766208e-19       # src/global.h:37

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """SplitMix64 finaliser of counter values x (uint64 array)."""
    with np.errstate(over="ignore"):
        z = (np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed, stream, n, offset=0):
    """U(0,1) open interval: value i depends only on (seed, stream, offset+i)."""
    i = np.arange(offset, offset + n, dtype=np.uint64)
    ctr = (np.uint64(seed) << np.uint64(40)) ^ (np.uint64(stream) << np.uint64(32)) ^ i
    z = splitmix64(ctr)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def gaussian(seed, stream, n, offset=0):
    """N(0,1) by Box-Muller from streams (2*stream, 2*stream+1)."""
    u1 = uniform(seed, 2 * stream, n, offset)
    u2 = uniform(seed, 2 * stream + 1, n, offset)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


@dataclass
class Undulator:
    K: float = 1.0
    lambda_u: float = 0.3
    n_periods: int = 8
    origin: tuple = (0.0, 0.0, 1.5)

    @property
    def bpeak(self):
        # K = lambda_u c B / (2 pi m c^2)  (src/undulator.cpp:75 with Krms*sqrt 2)
        return self.K * 2.0 * math.pi * MEC2_EV / (C_LIGHT * self.lambda_u)

    @property
    def length(self):
        return (self.n_periods + 1) * self.lambda_u


@dataclass
class Bunch:
    x0: np.ndarray          # (n, 3) initial positions [m]
    u0: np.ndarray          # (n, 3) initial beta*gamma
    charge: float = -1.0    # per particle, units of e
    gamma0: float = 60.0

    @property
    def mass(self):
        """Per particle, units of m_e: a macro-particle of N electrons has
        charge -N and mass N, keeping the electron's charge-to-mass ratio (as
        tests/teufel.DiffractionScreen.cpp:78 builds one)."""
        return abs(self.charge)


def single_electron(gamma=60.0):
    bg = math.sqrt(gamma * gamma - 1.0)
    return Bunch(np.zeros((1, 3)), np.array([[0.0, 0.0, bg]]), -1.0, gamma)


def sample_bunch(n, seed, gamma0=60.0, rel_sigma_gamma=1e-3, sigma_xy=0.5e-3, sigma_div=50e-6,
                 z_uniform=0.3e-3, sigma_z=None, charge=-1.0):
    """Gaussian transverse phase space, uniform (or Gaussian) longitudinal."""
    g = gamma0 * (1.0 + rel_sigma_gamma * gaussian(seed, 0, n))
    x = sigma_xy * gaussian(seed, 1, n)
    y = sigma_xy * gaussian(seed, 2, n)
    xp = sigma_div * gaussian(seed, 3, n)
    yp = sigma_div * gaussian(seed, 4, n)
    if sigma_z is None:
        z = (uniform(seed, 10, n) - 0.5) * z_uniform
    else:
        z = sigma_z * gaussian(seed, 5, n)
    bg = np.sqrt(g * g - 1.0)
    norm = np.sqrt(1.0 + xp * xp + yp * yp)
    u = np.stack([bg * xp / norm, bg * yp / norm, bg / norm], axis=1)
    return Bunch(np.stack([x, y, z], axis=1), u, charge, gamma0)


@dataclass
class Screen:
    obs: np.ndarray         # (n_pix, 3)
    t0: np.ndarray          # (n_pix,)
    dt: float
    n_bins: int
    normal: np.ndarray = None   # (n_pix, 3), facing the source
    area: np.ndarray = None     # (n_pix,)
    name: str = ""

    @property
    def n_pix(self):
        return self.obs.shape[0]


def grid_screen(n, size, distance, tilt_deg=0.0, toroid_f=None):
    """n x n pixel centres over size x size, ScreenObserver cell convention
    (src/screen.cpp:98-103: offsets ix - n/2 with integer division), optional
    toroidal sag z = sqrt((sqrt(f^2-y^2)+f)^2 - x^2) - 2f, tilted about y and
    placed at z = distance.  Returns obs, normal (towards -z side), area."""
    d = size / n
    off = (np.arange(n) - n // 2) * d
    X, Y = np.meshgrid(off, off, indexing="ij")
    X, Y = X.reshape(-1), Y.reshape(-1)
    if toroid_f:
        f = toroid_f
        Z = np.sqrt((np.sqrt(f * f - Y * Y) + f) ** 2 - X * X) - 2.0 * f
        r1 = np.sqrt(f * f - Y * Y) + f
        dzdx = -X / (Z + 2.0 * f)
        dzdy = -(r1 * Y / np.sqrt(f * f - Y * Y)) / (Z + 2.0 * f)
        nrm = np.stack([dzdx, dzdy, -np.ones_like(X)], axis=1)
    else:
        Z = np.zeros_like(X)
        nrm = np.stack([np.zeros_like(X), np.zeros_like(X), -np.ones_like(X)], axis=1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    a = math.radians(tilt_deg)
    ca, sa = math.cos(a), math.sin(a)
    Xr = ca * X + sa * Z
    Zr = -sa * X + ca * Z
    obs = np.stack([Xr, Y, Zr + distance], axis=1)
    nr = np.stack([ca * nrm[:, 0] + sa * nrm[:, 2], nrm[:, 1], -sa * nrm[:, 0] + ca * nrm[:, 2]], axis=1)
    area = np.full(n * n, d * d)
    return obs, nr, area


def set_windows(obs, n_bins, und, gamma, t_c, x_c, bunch_len):
    """Per-pixel t0 and common dt from the arrival curve of the emission at x_c."""
    x_c = np.asarray(x_c, dtype=np.float64)
    rel = obs - x_c
    R = np.linalg.norm(rel, axis=1)
    theta = np.arccos(np.clip(rel[:, 2] / R, -1.0, 1.0))
    thmax = float(theta.max()) if theta.size else 0.0
    W = 1.3 * (und.length * (1.0 + und.K ** 2 / 2.0 + (gamma * thmax) ** 2) / (2.0 * gamma * gamma * C_LIGHT)
               + bunch_len / C_LIGHT)
    dt = W / n_bins
    T = t_c + R / C_LIGHT
    t0 = T - W / 2.0 + dt / 2.0
    return t0, dt


@dataclass
class Workload:
    name: str
    und: Undulator
    bunch: Bunch
    dt: float
    n_steps: int
    screens: list = field(default_factory=list)
    description: str = ""

    @property
    def n_particles(self):
        return self.bunch.x0.shape[0]

    @property
    def n_points(self):
        return self.n_steps + 1

    def n_pairs(self):
        """Sum_screens n_pix x Sum_particles (NP - 1): the metric's pair count."""
        return sum(s.n_pix for s in self.screens) * self.n_particles * self.n_steps


def _centre(und, gamma):
    """Time/position of the nominal electron at the undulator centre (x_c, t_c)."""
    beta = math.sqrt(1.0 - 1.0 / (gamma * gamma))
    zc = und.origin[2]
    # drift from z=0 to the field start, then the undulator half with the
    # average longitudinal velocity reduced by K^2/(4 gamma^2)
    z_start = zc - und.length / 2.0
    bz_und = beta * (1.0 - und.K ** 2 / (4.0 * gamma * gamma))
    t_c = z_start / (beta * C_LIGHT) + (und.length / 2.0) / (bz_und * C_LIGHT)
    return t_c, np.array([0.0, 0.0, zc])


def make(name, n_particles=None, seed=None, n_steps=None):
    """Build one of the named workloads (optionally shrunk for tests)."""
    name = name.lower()
    if name == "c1":
        und = Undulator(n_periods=8, origin=(0.0, 0.0, 1.5))
        b = single_electron(60.0)
        beta = math.sqrt(1.0 - 1.0 / 3600.0)
        dt = und.lambda_u / (beta * C_LIGHT) / 100.0
        steps = 1000 if n_steps is None else n_steps
        w = Workload("c1", und, b, dt, steps, description="1 e-, 21x21 px, 256 bins, FP64")
        obs, nrm, area = grid_screen(21, 0.1, 10.0)
        t_c, x_c = _centre(und, 60.0)
        t0, sdt = set_windows(obs, 256, und, 60.0, t_c, x_c, 0.0)
        w.screens.append(Screen(obs, t0, sdt, 256, nrm, area, "flat21"))
        return w
    if name in ("c2", "c3", "c4", "c5"):
        cfg = {"c2": (10000, 1002, 64, 0.2, 1024), "c3": (100000, 1003, 128, 0.2, 2048),
               "c4": (20000, 1004, None, None, None), "c5": (200000, 1005, 128, 0.2, 4096)}[name]
        n = cfg[0] if n_particles is None else n_particles
        sd = cfg[1] if seed is None else seed
        if name == "c5":
            und = Undulator(n_periods=34, origin=(0.0, 0.0, 0.15 + 35 * 0.3 / 2.0))
            b = sample_bunch(n, sd, sigma_z=0.1e-3, charge=-6241.5)
            path = und.origin[2] + und.length / 2.0 + 0.15
            blen = 0.6e-3
            screen_z = und.origin[2] + und.length / 2.0 + 10.0
        else:
            und = Undulator(n_periods=8, origin=(0.0, 0.0, 1.5))
            b = sample_bunch(n, sd)
            path = 3.0
            blen = 0.3e-3
            screen_z = 10.0
        beta = math.sqrt(1.0 - 1.0 / 3600.0)
        steps = 4000 if n_steps is None else n_steps
        dt = path / (beta * C_LIGHT) / steps
        w = Workload(name, und, b, dt, steps)
        t_c, x_c = _centre(und, 60.0)
        if name != "c4":
            npx, size, nb = cfg[2], cfg[3], cfg[4]
            obs, nrm, area = grid_screen(npx, size, screen_z)
            t0, sdt = set_windows(obs, nb, und, 60.0, t_c, x_c, blen)
            w.screens.append(Screen(obs, t0, sdt, nb, nrm, area, "flat%d" % npx))
            w.description = "%d e-, %dx%d px, %d bins" % (n, npx, npx, nb)
        else:
            w.screens = batched_screens(32, 1004, und, t_c, x_c, blen)
            w.description = "%d e-, 32 uneven screens" % n
        return w
    raise ValueError("unknown workload %r" % name)


def batched_screens(count, seed, und, t_c, x_c, blen, sizes=(16, 24, 32, 48, 64, 96, 128),
                    bins=(256, 512, 1024, 2048, 4096)):
    """Seeded uneven screens: flat, tilted (0-30 deg about y) or toroidal (f 1-5 m)."""
    u = [uniform(seed, s, count) for s in range(6)]
    out = []
    for i in range(count):
        n = sizes[int(u[0][i] * len(sizes))]
        nb = bins[int(u[1][i] * len(bins))]
        dist = 2.0 + 18.0 * u[2][i]
        kind = int(u[3][i] * 3)
        tilt = 30.0 * u[4][i] if kind == 1 else 0.0
        f = 1.0 + 4.0 * u[5][i] if kind == 2 else None
        size = 0.02 * dist
        obs, nrm, area = grid_screen(n, size, x_c[2] + dist, tilt, f)
        t0, sdt = set_windows(obs, nb, und, 60.0, t_c, x_c, blen)
        out.append(Screen(obs, t0, sdt, nb, nrm, area,
                          "%s%d_b%d_d%.1f" % (("flat", "tilt", "toroid")[kind], n, nb, dist)))
    return out


def energy_density(S, normal):
    """Radiated energy per unit area through each pixel, W_p = -S_p . n_p
    (n facing the source; MeshedScreen::totalEnergy sign, src/mesh.cpp:744-754)."""
    return -np.einsum("ij,ij->i", S, normal)
