"""Pins of the oracle's optional physics switches (macro.h GRAVITY, VPM, SPONGE).

Each check compares with what the mathematics fixes, not with a retyped
formula: the buoyancy torque of x-only fields in closed form, the wall
penalisation reducing to linear damping -omega/eta when chi = 1 (and to the
plain right-hand side when chi = 0), the sponge reducing to -s*omega for a
constant sponge, and the symmetry of the init.c wall profiles.
"""
import math
from dataclasses import replace

import numpy as np
import pytest

import oracle as O

TWO_PI = 2.0 * math.pi


def base(model, N=32, **kw):
    d = dict(model=model, Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI, nu=0.01, dt=1e-3, lam=1e-3,
             epsilon=1.5 * TWO_PI / N, sigma=0.1, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06)
    d.update(kw)
    return O.Params(**d)


def grid(p):
    x = (np.arange(p.Nx) * p.dx)[:, None] * np.ones((1, p.Ny))
    y = (np.arange(p.Ny) * p.dy)[None, :] * np.ones((p.Nx, 1))
    return x, y


@pytest.mark.parametrize("model", ["binary", "ternary"])
def test_gravity_torque_of_x_only_fields(model):
    """omega = 0, c = c(x): advection and the curl of the surface-tension force vanish,
    and N_w = -g d/dx rho(c) = g k sin(kx) sum_i (rho_i - rho_ref) A_i."""
    p = base(model, gravity=True, rho1=1.3, rho2=0.9, rho3=1.1, g=2.5)
    x, _ = grid(p)
    k = 3
    A, B = 0.07, -0.04
    c1 = 1.0 / 3.0 + A * np.cos(k * x)
    c2 = 1.0 / 3.0 + B * np.cos(k * x)
    if model == "binary":
        s = O.State(np.zeros((p.Nx, p.Ny2), complex), O.dfft(c1))
        expect = p.g * k * np.sin(k * x) * (p.rho1 - p.rho2) * A
    else:
        s = O.State(np.zeros((p.Nx, p.Ny2), complex), O.dfft(c1), O.dfft(c2))
        expect = p.g * k * np.sin(k * x) * ((p.rho1 - p.rho3) * A + (p.rho2 - p.rho3) * B)
    nw = O.ifft(p, O.rhs(p, s)[0])
    assert np.max(np.abs(nw - expect)) < 1e-12 * max(1.0, np.max(np.abs(expect)))
    # layered fields c = c(y) feel no torque
    _, y = grid(p)
    cy = 1.0 / 3.0 + A * np.cos(k * y)
    fields = [O.dfft(cy)] if model == "binary" else [O.dfft(cy), O.dfft(cy)]
    s2 = O.State(np.zeros((p.Nx, p.Ny2), complex), *fields)
    assert np.max(np.abs(O.ifft(p, O.gravity_term(p, *fields)))) < 1e-12


@pytest.mark.parametrize("model", ["ns", "binary", "ternary"])
def test_vpm_full_wall_is_linear_damping(model):
    """chi = 1 everywhere: N_w = -omega/eta for a single-mode omega (no self-advection),
    the surface tension and the advection of c vanish (the (1 - chi) factors)."""
    p = base(model, vpm=True, chi_override=np.ones(32))
    x, y = grid(p)
    w = 0.8 * np.cos(2 * x + 3 * y)
    rng = np.random.default_rng(1)
    cs = [O.dfft(1.0 / 3.0 + 0.05 * np.cos(x + 2 * y + rng.uniform())) for _ in range(2)]
    fields = [O.dfft(w)] + cs[: {"ns": 0, "binary": 1, "ternary": 2}[model]]
    got = O.rhs(p, O.State(*fields))
    eta = 1.5 * p.dt
    assert np.max(np.abs(O.ifft(p, got[0]) + w / eta)) < 1e-9 * np.max(np.abs(w / eta))
    if model != "ns":
        # the c equations keep only -(lambda/gamma) k^2 M: equal to the plain model at omega = 0
        still = O.rhs(replace(p, vpm=False, chi_override=None), O.State(0 * fields[0], *fields[1:]))
        for a, b in zip(got[1:], still[1:]):
            assert O.rel_l2(a, b) < 1e-12


@pytest.mark.parametrize("model", ["ns", "ternary"])
def test_vpm_no_wall_and_constant_sponge(model):
    """chi = 0: the plain right-hand side; a constant sponge s adds exactly -s T(omega)."""
    p0 = base(model)
    x, y = grid(p0)
    w = O.random_vorticity(p0, 4)
    cs = O.init_noise_real(p0, 3, 1.0 / 3.0, 0.05)
    s = O.state_from_real(p0, w, *(cs if model == "ternary" else ()))
    plain = O.rhs(p0, s)
    p1 = replace(p0, vpm=True, chi_override=np.zeros(32))
    for a, b in zip(O.rhs(p1, s), plain):
        assert O.rel_l2(a, b) < 1e-13
    s0 = 7.5
    p2 = replace(p1, sponge=True, sponge_override=np.full(32, s0))
    got = O.rhs(p2, s)
    assert O.rel_l2(got[0] - plain[0], -s0 * O.truncation(p0, s.cwz)) < 1e-12
    for a, b in zip(got[1:], plain[1:]):
        assert O.rel_l2(a, b) < 1e-13


def test_wall_profiles():
    """init.c:184-211 profiles: mirror-symmetric, 1/2 at the walls x = 6dx and lx - 6dx,
    ~1 at the boundary and ~0 in the interior; the sponge carries 1/eta."""
    p = base("ternary", N=64)
    chi, sp, etainv = O.vpm_fields(p)
    N = p.Nx
    assert np.allclose(chi[1:], chi[1:][::-1], rtol=0, atol=1e-13)   # x = i dx vs lx - x round differently
    assert abs(chi[6] - 0.5) < 1e-6 and abs(chi[N - 6] - 0.5) < 1e-6
    assert chi[0] > 0.999 and chi[N // 2] < 1e-12
    assert np.allclose(sp / etainv, chi, rtol=0, atol=1e-15)   # same profile in y on a square grid
    assert etainv == 1.0 / (1.5 * p.dt)
