"""The reference's homogeneous-field tracking test cases as inputs and
closed-form answers, shared by the oracle's CPU tests and the GPU pusher's
tests."""
import math

import numpy as np

from teufel_amd import workloads as W

C = W.C_LIGHT


def magnet_case():
    """tests/teufel.magnet.cpp: an electron at gamma = 10 starting at the
    origin perpendicular to a homogeneous B of 0.1 T, 2 x 500 steps over one
    revolution.  Returns (B, u0, dt, n_steps, gyro radius, period, beta gamma)."""
    B = np.array([0.033166247903554, 0.05, 0.08])
    p = np.cross([0.08, 0.05, 0.03], B)
    gamma = 10.0
    bg, beta = math.sqrt(gamma * gamma - 1.0), math.sqrt(1.0 - 1.0 / gamma ** 2)
    r_gyro = bg * W.MEC2_EV / C / np.linalg.norm(B)
    tau = 2 * math.pi * r_gyro / (beta * C)
    return B, p * bg / np.linalg.norm(p), 0.5 * tau / 500, 1000, r_gyro, tau, bg


def exb_case():
    """tests/teufel.EcrossB.cpp: an electron at gamma = 30.358 in crossed
    E = B x v and B, 3000 steps over 10 ns.  Returns (E, B, u0, v, dt, n_steps,
    tau)."""
    B = np.array([0.12, 0.52, 0.45])
    gamma = 30.35823388162631
    bg, beta = math.sqrt(gamma * gamma - 1.0), math.sqrt(1.0 - 1.0 / gamma ** 2)
    d = np.array([2.0, 3.0, 5.0])
    d = d / np.linalg.norm(d)
    v = d * beta * C
    return np.cross(B, v), B, d * bg, v, 10e-9 / 3000, 3000, 10e-9
