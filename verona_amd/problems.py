"""Initial states of the benchmark problems (host arrays in Verona layout).

kelvin_helmholtz(n): the 2D relativistic Kelvin-Helmholtz setup, n x n
periodic cells on [0, 1]^2, gamma = 4/3, p = 1 everywhere.  With
T(y) = tanh((y - 1/4)/0.01) - tanh((y - 3/4)/0.01) (= 2 inside the band
1/4 < y < 3/4, 0 outside): rho = 0.1 + 0.45 T, v_x = -0.5 + 0.5 T,
v_y = 0.01 sin(4 pi x) [exp(-(y - 1/4)^2 / 0.1^2) + exp(-(y - 3/4)^2 / 0.1^2)],
plus uniform noise in [-1e-4, 1e-4) on v_x and v_y drawn from SplitMix64
(seed 42) in row order (j outer, i inner; v_x draw then v_y draw per cell).
Primitives u^i = Gamma v^i, u = p / (gamma - 1).

shock_tube(n): the 1D relativistic blast wave (Marti & Mueller problem 1)
on [0, 1] with the diaphragm at x = 1/2: rho = 10, p = 13.33 on the left,
rho = 1, p = 1e-6 on the right, at rest, gamma = 5/3.
"""
import math

import numpy as np

from .bc import uniform_array
from .device import host_array


def kelvin_helmholtz(n, seed=42, gamma=4.0 / 3.0):
    P = host_array(2, n, n)
    dx = 1.0 / n
    c = (np.arange(n) + 0.5) * dx
    x, y = np.meshgrid(c, c, indexing="ij")  # [i, j]
    T = np.tanh((y - 0.25) / 0.01) - np.tanh((y - 0.75) / 0.01)
    rho = 0.1 + 0.45 * T
    vx = -0.5 + 0.5 * T
    vy = 0.01 * np.sin(4.0 * math.pi * x) * (np.exp(-((y - 0.25) ** 2) / 0.01) +
                                            np.exp(-((y - 0.75) ** 2) / 0.01))
    noise = uniform_array(seed, 2 * n * n).reshape(n, n, 2)  # [j, i, draw]
    vx = vx + (2.0 * noise[:, :, 0].T - 1.0) * 1e-4
    vy = vy + (2.0 * noise[:, :, 1].T - 1.0) * 1e-4
    lor = 1.0 / np.sqrt(1.0 - vx * vx - vy * vy)
    s = slice(3, n + 3)
    P[0, s, s] = rho
    P[1, s, s] = 1.0 / (gamma - 1.0)
    P[2, s, s] = lor * vx
    P[3, s, s] = lor * vy
    return P


def shock_tube(n=1024, gamma=5.0 / 3.0):
    rho = np.where(np.arange(n) < n // 2, 10.0, 1.0)
    p = np.where(np.arange(n) < n // 2, 13.33, 1e-6)
    return rho, p / (gamma - 1.0), np.zeros(n)
