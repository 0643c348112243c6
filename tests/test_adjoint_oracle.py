"""Pins of the oracle's discrete adjoint (oracle/chns_adjoint.py).

The adjoint is checked against what the forward oracle fixes, never against a
restatement of itself: central finite differences of the oracle loss for the
initial fields and sigma_12/13/23, and the transpose identity
<g, J d> = <J^T g, d> of the right-hand side with J d from central
differences of ``oracle.rhs``.
"""
import math
from dataclasses import replace

import numpy as np
import pytest

import oracle as O
from oracle import chns_adjoint as OA

TWO_PI = 2.0 * math.pi


def params(N=16, **kw):
    base = dict(model="ternary", Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI, nu=0.02, dt=2e-3, lam=0.02,
                epsilon=1.5 * TWO_PI / N, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06, Xlambda=2.0)
    base.update(kw)
    return O.Params(**base)


def fields(p, seed=5, amp=0.1):
    c1, c2 = O.init_noise_real(p, seed, 1.0 / 3.0, amp)
    return [O.random_vorticity(p, seed + 1, amp=2.0), c1, c2]


def loss_fn(p, nsteps, target):
    def f(wz, c1, c2):
        s = O.run(p, O.state_from_real(p, wz, c1, c2), nsteps)
        d = O.spectrum(p, s)[0]["sph1"] - target
        return float(np.sum(d * d))
    return f


def test_rhs_vjp_is_transpose_of_rhs():
    p = params()
    s = O.state_from_real(p, *fields(p))
    rng = np.random.default_rng(3)
    d = O.state_from_real(p, *[rng.standard_normal((p.Nx, p.Ny)) for _ in range(3)])
    g = [O.truncation(p, O.dfft(rng.standard_normal((p.Nx, p.Ny)))) for _ in range(3)]
    h = 1e-5

    def shifted(sign):
        return O.State(*[a + sign * h * b for a, b in zip(s.fields(), d.fields())])

    jd = [(a - b) / (2 * h) for a, b in zip(O.rhs(p, shifted(1)), O.rhs(p, shifted(-1)))]
    lhs = sum(OA.spec_dot(p, gk, jk) for gk, jk in zip(g, jd))
    jt, _ = OA.rhs_vjp(p, s, *g)
    rhs = sum(OA.spec_dot(p, a, b) for a, b in zip(jt, d.fields()))
    assert abs(lhs - rhs) <= 1e-7 * abs(rhs)
    # every state adjoint lives on the kept modes
    mask = O.dealias_mask(p)
    assert all(np.all(x[~mask] == 0) for x in jt)


def test_spectrum_loss_seed_matches_fd():
    p = params()
    f = fields(p)
    target = np.abs(np.random.default_rng(9).standard_normal(p.nshell)) * 1e-4
    loss, grads, _, _ = OA.loss_and_grad(p, O.state_from_real(p, *f), 0, target)
    L = loss_fn(p, 0, target)
    assert math.isclose(loss, L(*f), rel_tol=1e-14)
    d = np.random.default_rng(2).standard_normal((p.Nx, p.Ny))
    h = 1e-4
    fd = (L(f[0], f[1] + h * d, f[2]) - L(f[0], f[1] - h * d, f[2])) / (2 * h)
    assert abs(fd - np.sum(grads[1] * d)) <= 1e-7 * abs(fd)
    assert np.all(grads[0] == 0) and np.all(grads[2] == 0)


@pytest.mark.parametrize("N,nsteps", [(16, 3), (32, 2)])
def test_loss_gradient_matches_central_differences(N, nsteps):
    p = params(N)
    f = fields(p)
    tgt = O.state_from_real(p, None, *O.init_noise_real(p, 7, 1.0 / 3.0, 0.1))
    target = O.spectrum(p, O.run(p, tgt, nsteps))[0]["sph1"]
    loss, grads, gsig, _ = OA.loss_and_grad(p, O.state_from_real(p, *f), nsteps, target)
    L = loss_fn(p, nsteps, target)
    assert math.isclose(loss, L(*f), rel_tol=1e-14)
    rng = np.random.default_rng(11)
    for k in range(3):
        d = rng.standard_normal((p.Nx, p.Ny))
        h = 1e-4 if k else 1e-3
        plus = [x.copy() for x in f]
        minus = [x.copy() for x in f]
        plus[k] += h * d
        minus[k] -= h * d
        fd = (L(*plus) - L(*minus)) / (2 * h)
        ad = float(np.sum(grads[k] * d))
        assert abs(fd - ad) <= 1e-6 * abs(ad), (k, fd, ad)
    for k, name in enumerate(("sigma_12", "sigma_13", "sigma_23")):
        h = 1e-4 * getattr(p, name)
        Lp = loss_fn(replace(p, **{name: getattr(p, name) + h}), nsteps, target)(*f)
        Lm = loss_fn(replace(p, **{name: getattr(p, name) - h}), nsteps, target)(*f)
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - gsig[k]) <= 1e-6 * abs(gsig[k]), (name, fd, gsig[k])


def test_sigma_partials_planted_error_is_caught():
    """A plausible slip (dropping the del(gamma) dependence of fac) must fail the FD check."""
    p = params()
    f = fields(p)
    nsteps = 2
    tgt = O.state_from_real(p, None, *O.init_noise_real(p, 7, 1.0 / 3.0, 0.1))
    target = O.spectrum(p, O.run(p, tgt, nsteps))[0]["sph1"]
    _, _, gsig, acc = OA.loss_and_grad(p, O.state_from_real(p, *f), nsteps, target)
    bad = dict(acc)
    bad["fac"] = 0.0
    gbad = OA.sigma_gradient(p, bad)
    name = "sigma_12"
    h = 1e-4 * p.sigma_12
    L = loss_fn
    fd = (L(replace(p, sigma_12=p.sigma_12 + h), nsteps, target)(*f)
          - L(replace(p, sigma_12=p.sigma_12 - h), nsteps, target)(*f)) / (2 * h)
    assert abs(fd - gsig[0]) <= 1e-6 * abs(fd)
    assert abs(fd - gbad[0]) > 1e-3 * abs(fd), name


def fe_loss_fn(p, nsteps):
    """Total free energy of the truncated phase fields after nsteps (the
    reference's free_energy.c sums, evaluated on T(c))."""
    def f(wz, c1, c2):
        s = O.run(p, O.state_from_real(p, wz, c1, c2), nsteps)
        t = O.State(s.cwz, O.truncation(p, s.cphi1), O.truncation(p, s.cphi2))
        return sum(O.free_energy(p, t))
    return f


def test_free_energy_loss_value_and_seed():
    """The loss equals the oracle's free_energy() of the truncated state, and
    its seed gives the central difference along a random state direction."""
    p = params()
    s = O.state_from_real(p, *fields(p))
    loss, ephi, egphi, ubar, _ = OA.free_energy_loss(p, s)
    t = O.State(s.cwz, O.truncation(p, s.cphi1), O.truncation(p, s.cphi2))
    e1, e2 = O.free_energy(p, t)
    assert math.isclose(ephi, e1, rel_tol=1e-13) and math.isclose(egphi, e2, rel_tol=1e-13)
    rng = np.random.default_rng(4)
    d = [O.truncation(p, O.dfft(rng.standard_normal((p.Nx, p.Ny)))) for _ in range(2)]
    h = 1e-5

    def L(sign):
        q = O.State(s.cwz, s.cphi1 + sign * h * d[0], s.cphi2 + sign * h * d[1])
        return OA.free_energy_loss(p, q)[0]

    fd = (L(1) - L(-1)) / (2 * h)
    ad = OA.spec_dot(p, ubar[1], d[0]) + OA.spec_dot(p, ubar[2], d[1])
    assert abs(fd - ad) <= 1e-8 * abs(ad)


@pytest.mark.parametrize("N,nsteps", [(16, 3), (32, 2)])
def test_free_energy_gradient_matches_central_differences(N, nsteps):
    """Gradients of the final free energy with respect to the initial fields
    and sigma_12/13/23 (whose direct part, the gamma_i weights of the energy
    itself, a dropped term would miss: checked by planting it below)."""
    p = params(N)
    f = fields(p)
    loss, grads, gsig, acc = OA.loss_and_grad(p, O.state_from_real(p, *f), nsteps, None)
    L = fe_loss_fn(p, nsteps)
    assert math.isclose(loss, L(*f), rel_tol=1e-13)
    rng = np.random.default_rng(12)
    for k in range(3):
        d = rng.standard_normal((p.Nx, p.Ny))
        h = 1e-4 if k else 1e-3
        plus = [x.copy() for x in f]
        minus = [x.copy() for x in f]
        plus[k] += h * d
        minus[k] -= h * d
        fd = (L(*plus) - L(*minus)) / (2 * h)
        ad = float(np.sum(grads[k] * d))
        assert abs(fd - ad) <= 1e-6 * abs(ad), (k, fd, ad)
    no_direct = OA.sigma_gradient(p, acc)   # without the loss's own gamma dependence
    for k, name in enumerate(("sigma_12", "sigma_13", "sigma_23")):
        h = 1e-4 * getattr(p, name)
        fd = (fe_loss_fn(replace(p, **{name: getattr(p, name) + h}), nsteps)(*f)
              - fe_loss_fn(replace(p, **{name: getattr(p, name) - h}), nsteps)(*f)) / (2 * h)
        assert abs(fd - gsig[k]) <= 1e-6 * abs(gsig[k]), (name, fd, gsig[k])
        assert abs(fd - no_direct[k]) > 1e-3 * abs(fd)
