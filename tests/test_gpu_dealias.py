"""GPU parity of the 2/3 rule (dealias = 1) and the N/3 circle (dealias = 2)
against the oracle.  Both replace the reference's N/4 circle of truncation()
(functions.c:56-59,75-78) by another keep-set, pinned by lattice counts in
tests/test_oracle_pins.py; the library runs them through its unpruned kernels
with a 0/1 factor table.

Bar: one-step increments and 20-step trajectories at rel-L2 1e-10, as for the
reference rule; the truncated state itself bit for bit.
"""
import math

import numpy as np
import pytest

import chns_amd as C
import oracle as O

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi


def case(model, N, rule, **over):
    kw = dict(Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI, nu=5e-3, dt=5e-4 * (32.0 / N) ** 2, lam=1e-3,
              epsilon=1.5 * TWO_PI / N, sigma=0.1, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06, dealias=rule)
    kw.update(over)
    op = O.Params(model=model, **kw)
    wz = O.random_vorticity(op, 1)
    if model == "ternary":
        reals = (wz,) + tuple(O.init_noise_real(op, 2, 1.0 / 3.0, 0.05))
    elif model == "binary":
        reals = (wz,) + O.init_noise_real(op, 1, 0.5, 0.1)
    else:
        reals = (wz,)
    return op, C.params(model, **kw), reals


@pytest.mark.parametrize("rule", [1, 2])
@pytest.mark.parametrize("model", ["ns", "binary", "ternary"])
@pytest.mark.parametrize("N", [32, 64])
def test_step_parity(rule, model, N):
    op, cp, reals = case(model, N, rule)
    s0 = O.State(*[O.dfft(r) for r in reals])      # content in every mode, cut by the step
    s1 = O.etd2rk(op, s0)
    ref = O.run(op, s0, 20, spectrum_every=5)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        ctx.step(1)
        g1 = ctx.get_state()
        ctx.step(4)
        ctx.truncate_state()            # spectrum()'s side effect every 5 steps
        for _ in range(3):
            ctx.step(5)
            ctx.truncate_state()
        g20 = ctx.get_state()
    for a, b, a0 in zip(g1, s1.fields(), s0.fields()):
        assert O.rel_l2(a - a0, b - a0) < 1e-10
    for a, b in zip(g20, ref.fields()):
        assert O.rel_l2(a, b) < 1e-10


@pytest.mark.parametrize("rule", [0, 1, 2])
def test_initial_state_and_truncation(rule):
    """dfft + truncation of real fields, and the state truncation, under each rule
    (the mask bit for bit: kept modes unchanged, cut modes exactly zero)."""
    op, cp, reals = case("ternary", 64, rule)
    s0 = O.state_from_real(op, *reals)
    full = O.State(*[O.dfft(r) for r in reals])
    mask = O.dealias_mask(op)
    with C.Context(cp) as ctx:
        ctx.set_real_state(*reals)
        g = ctx.get_state()
        ctx.set_state(*full.fields())
        ctx.truncate_state()
        t = ctx.get_state()
    for a, b in zip(g, s0.fields()):
        assert O.rel_l2(a, b) < 1e-14
        assert np.all(a[~mask] == 0)
    assert np.array_equal(t[0], full.cwz)                       # omega is not truncated
    for a, b in zip(t[1:], full.fields()[1:]):
        assert np.array_equal(a[mask], b[mask]) and np.all(a[~mask] == 0)


@pytest.mark.parametrize("rule", [1, 2])
def test_diagnostics(rule):
    op, cp, reals = case("ternary", 64, rule)
    s = O.run(op, O.state_from_real(op, *reals), 3)
    ref_sp, s2 = O.spectrum(op, s)
    with C.Context(cp) as ctx:
        ctx.set_state(*s.fields())
        ke = ctx.kinetic_energy()
        fe = ctx.free_energy()
        sp = ctx.spectrum(truncate_state=True)
    for x, y in zip(ke, O.kinetic_energy(op, s)):
        assert abs(x - y) <= 1e-12 * max(abs(y), 1e-300)
    for x, y in zip(fe, O.free_energy(op, s)):
        assert abs(x - y) <= 1e-12 * max(abs(y), 1e-300)
    for k in ref_sp:
        assert O.rel_l2(sp[k][: len(ref_sp[k])], ref_sp[k]) < 1e-12, k


@pytest.mark.parametrize("rule", [1, 2])
def test_step_2048(rule):
    """2048^2: one ternary step (Jt = 683 dealiased columns, the full-width
    row stager)."""
    N = 2048
    op, cp, _ = case("ternary", N, rule, nu=2e-3, lam=5e-4, dt=1e-6)
    reals = (O.random_vorticity(op, 1),) + tuple(O.init_lens_real(op))
    s0 = O.state_from_real(op, *reals)
    s1 = O.etd2rk(op, s0)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        ctx.step(1)
        g = ctx.get_state()
    for a, b, a0 in zip(g, s1.fields(), s0.fields()):
        assert O.rel_l2(a - a0, b - a0) < 1e-10


def test_rejections():
    kw = dict(Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI, dealias=1)
    with pytest.raises(C.ChnsError):
        C.Ensemble([C.params("ternary", **kw)])                    # single contexts only
    with pytest.raises(C.ChnsError):
        C.Context(C.params("ternary", smooth_filter=1, **kw))      # the filter replaces the cut
    with C.Context(C.params("ternary", **kw)) as ctx:
        ctx.set_real_state(None, np.full((32, 32), 0.3), np.full((32, 32), 0.3))
        with pytest.raises(C.ChnsError):
            ctx.adj_forward(1, 1)                                  # the adjoint covers the reference rule
