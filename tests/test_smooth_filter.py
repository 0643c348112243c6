"""Pins of the oracle's SMOOTH_FILTER form of truncation() (macro.h:15,
functions.c:54,70-73): every mode is multiplied by e = exp(-36 pp^18) twice,
pp = ksqr/ksqrmax, wherever the reference calls truncation().

What the mathematics fixes, with each input of a product filtered exactly once:
two vorticity modes make u.grad w out of cross terms only (a single Fourier
mode has u perpendicular to grad w), so the filtered right-hand side is
e1^2 e2^2 times the unfiltered one; the advection of a phase mode by a
vorticity mode scales the same way; a small phase mode enters the chemical
potential linearly, so its right-hand side scales by e^2.  A factor applied
zero or two times per field, or to the output, fails these.  "Unfiltered" is
the oracle with truncation() replaced by the identity.
"""
import math

import numpy as np
import pytest

import chns_amd as C
import oracle as O
import oracle.chns_oracle as OC

TWO_PI = 2.0 * math.pi
N = 32


def base(model, **kw):
    d = dict(model=model, Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI, nu=0.01, dt=1e-3, lam=1e-3,
             epsilon=1.5 * TWO_PI / N, sigma=0.1, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06,
             smooth_filter=True)
    d.update(kw)
    return O.Params(**d)


def grid(p):
    x = (np.arange(p.Nx) * p.dx)[:, None] * np.ones((1, p.Ny))
    y = (np.arange(p.Ny) * p.dy)[None, :] * np.ones((p.Nx, 1))
    return x, y


def efac(p, kx, ky):
    """e at wavenumber (kx, ky), from the oracle's table."""
    return O.smooth_filter_factor(p)[kx % p.Nx, ky]


def unfiltered(monkeypatch):
    monkeypatch.setattr(OC, "truncation", lambda p, cv: np.array(cv, copy=True))


def test_factor_identity_and_corner():
    """e == 1.0 exactly where 36 pp^18 is below half an ulp of 1; e = exp(-36) at the corner."""
    p = base("ns")
    f = O.smooth_filter_factor(p)
    pp = O.ksqr_grid(p) / (2.0 * (N / 2) ** 2)
    assert np.all(f[pp < 0.1] == 1.0)
    assert f[N // 2, N // 2] == math.exp(-36.0)
    assert np.all(f[pp > 0.3] < 1.0)
    # monotone in |k| along a row
    assert np.all(np.diff(f[:N // 2 + 1, 5]) <= 0.0)


@pytest.mark.parametrize("N_, L", [(16, TWO_PI), (64, TWO_PI), (32, 3.0)])
def test_host_table_bit_exact(N_, L):
    """chns_filter_table (C library exp/pow) equals the oracle's factor bit for bit."""
    kw = dict(Nx=N_, Ny=N_, lx=L, ly=L)
    t = C.filter_table(C.params("ternary", **kw))
    o = O.smooth_filter_factor(O.Params(model="ternary", **kw))
    assert np.array_equal(t, o)


def test_two_vorticity_modes(monkeypatch):
    p = base("ns")
    x, y = grid(p)
    k1, k2 = (11, 10), (-7, 13)                       # pp ~ 0.43: e^4 ~ 1 - 6e-5
    w = 0.7 * np.cos(k1[0] * x + k1[1] * y + 0.3) + 0.5 * np.cos(k2[0] * x + k2[1] * y - 1.1)
    s = O.State(O.dfft(w))
    got = O.rhs(p, s)[0]
    unfiltered(monkeypatch)
    ref = O.rhs(p, s)[0]
    scale = efac(p, *k1) ** 2 * efac(p, *k2) ** 2
    assert abs(1.0 - scale) > 1e-5
    assert O.rel_l2(got, scale * ref) < 1e-12


@pytest.mark.parametrize("model", ["binary", "ternary"])
def test_advection_of_a_phase_mode(model, monkeypatch):
    """N_c(w) - N_c(0) = -F(u.grad c) with u and grad c each filtered once."""
    p = base(model)
    x, y = grid(p)
    kw, kc = (9, 12), (-12, 8)
    w = 0.6 * np.cos(kw[0] * x + kw[1] * y + 0.2)
    c = 0.05 * np.cos(kc[0] * x + kc[1] * y - 0.4)
    if model == "binary":
        ph = (O.dfft(0.5 + c),)
    else:
        ph = (O.dfft(1.0 / 3.0 + c), O.dfft(1.0 / 3.0 - 0.5 * c))

    def adv():
        a = O.rhs(p, O.State(O.dfft(w), *ph))
        b = O.rhs(p, O.State(0.0 * O.dfft(w), *ph))
        return [u - v for u, v in zip(a[1:], b[1:])]

    got = adv()
    unfiltered(monkeypatch)
    ref = adv()
    scale = efac(p, *kw) ** 2 * efac(p, *kc) ** 2
    assert abs(1.0 - scale) > 1e-5
    for g, r in zip(got, ref):
        assert O.rel_l2(g, scale * r) < 1e-11


def test_potential_input_filtered_once(monkeypatch):
    """Binary, w = 0: N_c = -lambda 24 (sigma/eps) k^2 F(P(c)), c = T(c-hat): a small mode
    c = 1/2 + d cos(k.x) gives N_c(k) = e^2 times the unfiltered value to O(d)."""
    p = base("binary")
    x, y = grid(p)
    k = (12, 10)
    d = 1e-8
    s = O.State(np.zeros((N, N // 2 + 1), complex), O.dfft(0.5 + d * np.cos(k[0] * x + k[1] * y)))
    got = O.rhs(p, s)[1][k[0], k[1]]
    unfiltered(monkeypatch)
    ref = O.rhs(p, s)[1][k[0], k[1]]
    e2 = efac(p, *k) ** 2
    assert abs(1.0 - e2) > 1e-5
    # round-off ~1e-15 against an effect (1 - e^2)|ref| ~ 6e-10: a factor e^0 or e^4 misses by ~6e-10
    assert abs(got - e2 * ref) < 1e-3 * abs(1.0 - e2) * abs(ref)
