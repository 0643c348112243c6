"""The product's initial-condition generators (chns_amd.ic, which bench.py and
the driver feed to the library) equal the oracle's, bit for bit, so that the
GPU and CPU legs of a comparison start from the same fields."""
import math

import numpy as np
import pytest

import oracle as O
from chns_amd import ic

TWO_PI = 2.0 * math.pi


@pytest.mark.parametrize("model,N,seed,mean,amp", [("binary", 32, 5, 0.5, 0.1), ("ternary", 64, 2, 1 / 3, 0.05),
                                                   ("ternary", 16, 2003, 1 / 3, 0.05)])
def test_noise(model, N, seed, mean, amp):
    p = O.Params(model=model, Nx=N, Ny=N)
    nf = 1 if model == "binary" else 2
    got = ic.noise(N, N, seed, mean, amp, nf)
    ref = O.init_noise_real(p, seed, mean, amp)
    assert len(got) == len(ref) == nf
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("N,L", [(64, TWO_PI), (256, TWO_PI), (128, 4 * math.pi)])
def test_lens(N, L):
    p = O.Params(model="ternary", Nx=N, Ny=N, lx=L, ly=L, epsilon=1.5 * L / N)
    for kw in ({}, dict(radius=0.45)):
        got = ic.lens(N, N, L, L, p.epsilon, **kw)
        ref = O.init_lens_real(p, **kw)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)


def test_reference_ics():
    p = O.Params(model="ternary", Nx=32, Ny=32)
    for nr in (1, 2):
        for a, b in zip(ic.ternary_noise(32, 32, nr), O.init_ternary_noise_real(p, nr)):
            assert np.array_equal(a, b)
    q = O.Params(model="binary", Nx=32, Ny=32)
    assert np.array_equal(ic.binary_droplet(32, 32, q.lx, q.ly, q.epsilon), O.init_binary_droplet_real(q))
