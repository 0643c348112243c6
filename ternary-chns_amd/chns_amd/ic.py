"""Initial-condition generators (host side of init.c) producing packed real
[Nx][Ny] fields for ``Context.set_real_state`` (which applies dfft +
truncation on the device, init.c:253-262,291-294).

* ``ternary_noise``: the reference default, ``srand(10 + 2*rank)`` then
  ``1/4 + 0.001*rand()/RAND_MAX`` for c1 and c2 interleaved per point over each
  rank's x-slab (init.c:265-289), with the C library's rand().
* ``binary_droplet``: the tanh droplet of init.c:227-246.
* ``noise``: mean + U[-amp, amp] from SplitMix64 (53-bit uniforms), points
  row-major with x slow, c1/c2 interleaved per point for ternary.
* ``lens``: two c3 lenses on a c1/c2 band interface (liquid-lens coalescence).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(seed: int, count: int) -> np.ndarray:
    """U[0,1) doubles from SplitMix64 (Steele, Lea & Flood, 2014), vectorised."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def noise(Nx: int, Ny: int, seed: int, mean: float, amp: float, nfields: int):
    u = _splitmix64(seed, nfields * Nx * Ny).reshape(Nx, Ny, nfields)
    return tuple(mean + amp * (2.0 * u[..., f] - 1.0) for f in range(nfields))


def ternary_noise(Nx: int, Ny: int, nranks: int = 1):
    if Nx % nranks:
        raise ValueError("nranks must divide Nx")
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand.argtypes = [ctypes.c_uint]
    ln = Nx // nranks
    c1 = np.empty((Nx, Ny))
    c2 = np.empty((Nx, Ny))
    for r in range(nranks):
        libc.srand(10 + r * 2)                                    # init.c:268-269
        v = np.array([libc.rand() for _ in range(2 * ln * Ny)], dtype=np.float64) / 2147483647.0
        v = v.reshape(ln, Ny, 2)
        c1[r * ln:(r + 1) * ln] = (1.0 / 4.0) + 0.001 * v[..., 0]  # init.c:285
        c2[r * ln:(r + 1) * ln] = (1.0 / 4.0) + 0.001 * v[..., 1]  # init.c:286
    return c1, c2


def binary_droplet(Nx: int, Ny: int, lx: float, ly: float, epsilon: float):
    dx, dy = lx / Nx, ly / Ny
    x = (np.arange(Nx) * dx)[:, None]
    y = (np.arange(Ny) * dy)[None, :]
    r = np.sqrt((x - lx / 2.0) * (x - lx / 2.0) + (y - ly / 2.0) * (y - ly / 2.0))
    return 0.5 * (1.0 - np.tanh((r - (math.pi / 4.5)) / (0.5 * epsilon)))


def lens(Nx: int, Ny: int, lx: float, ly: float, epsilon: float, radius: float = 0.4, sep: float = 1.2):
    dx, dy = lx / Nx, ly / Ny
    x = (np.arange(Nx) * dx)[:, None] * np.ones((1, Ny))
    y = (np.arange(Ny) * dy)[None, :] * np.ones((Nx, 1))
    w = 0.5 * epsilon
    H = 0.5 * (np.tanh((y - ly / 4.0) / w) - np.tanh((y - 3.0 * ly / 4.0) / w))
    c3 = np.zeros_like(x)
    for sgn in (-1.0, 1.0):
        xc, yc = lx / 2.0 + sgn * sep / 2.0, 3.0 * ly / 4.0
        r = np.sqrt((x - xc) * (x - xc) + (y - yc) * (y - yc))
        c3 = np.maximum(c3, 0.5 * (1.0 - np.tanh((r - radius) / w)))
    return (1.0 - c3) * (1.0 - H), (1.0 - c3) * H
