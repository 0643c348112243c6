"""GPU parity of the SMOOTH_FILTER form of truncation() (macro.h:15,
functions.c:70-73) against the oracle (pinned in tests/test_smooth_filter.py):
the step, the initial-condition path, the state side effect of spectrum() and
the diagnostics that truncate (kinetic energy, spectra, dumps).

Bar: one-step increments and 20-step trajectories at rel-L2 1e-10, as for the
plain step; the filtered state itself bit for bit (the same two products).
"""
import math

import numpy as np
import pytest

import chns_amd as C
import oracle as O

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi


def case(model, N, **over):
    kw = dict(Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI, nu=5e-3, dt=5e-4 * (32.0 / N) ** 2, lam=1e-3,
              epsilon=1.5 * TWO_PI / N, sigma=0.1, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06, smooth_filter=1)
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


def full_state(op, reals):
    """A state with content in every mode (the filter keeps them all)."""
    return O.State(*[O.dfft(r) for r in reals])


@pytest.mark.parametrize("model", ["ns", "binary", "ternary"])
@pytest.mark.parametrize("N", [32, 64])
def test_filtered_step_parity(model, N):
    op, cp, reals = case(model, N)
    s0 = full_state(op, reals)
    s1 = O.etd2rk(op, s0)
    ref = O.run(op, s0, 20)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        ctx.step(1)
        g1 = ctx.get_state()
        ctx.step(19)
        g20 = ctx.get_state()
    for a, b, a0 in zip(g1, s1.fields(), s0.fields()):
        assert O.rel_l2(a - a0, b - a0) < 1e-10
    for a, b in zip(g20, ref.fields()):
        assert O.rel_l2(a, b) < 1e-10


def test_filter_with_optional_physics():
    """Filter together with the wall, the sponge and buoyancy (ternary, 32^2)."""
    op, cp, reals = case("ternary", 32, vpm=1, sponge=1, gravity=1)
    s0 = full_state(op, reals)
    ref = O.run(op, s0, 10)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        ctx.step(10)
        g = ctx.get_state()
    for a, b in zip(g, ref.fields()):
        assert O.rel_l2(a, b) < 1e-10


def test_filter_changes_the_step():
    """The filtered step differs from the sharp cut on a state with content beyond N/4."""
    op, cp, reals = case("ternary", 32)
    s0 = full_state(op, reals)
    cp0 = C.params("ternary", **{k: getattr(cp, k) for k in ("Nx", "Ny", "lx", "ly", "nu", "dt", "lambda_",
                                                              "epsilon", "sigma_12", "sigma_13", "sigma_23")})
    out = []
    for p in (cp, cp0):
        with C.Context(p) as ctx:
            ctx.set_state(*s0.fields())
            ctx.step(1)
            out.append(ctx.get_state())
    assert O.rel_l2(out[0][1] - s0.cphi1, out[1][1] - s0.cphi1) > 1e-6


@pytest.mark.parametrize("N", [32, 256])
def test_filtered_initial_condition(N):
    """set_real_state: dfft then truncation(), the filter here (init.c:291-294)."""
    op, cp, reals = case("ternary", N)
    ref = O.state_from_real(op, *reals)
    with C.Context(cp) as ctx:
        ctx.set_real_state(*reals)
        g = ctx.get_state()
    for a, b in zip(g, ref.fields()):
        assert O.rel_l2(a, b) < 1e-14


@pytest.mark.parametrize("model", ["ns", "binary", "ternary"])
def test_filtered_diagnostics(model):
    op, cp, reals = case(model, 64)
    s = O.run(op, full_state(op, reals), 3)
    ref_sp, s2 = O.spectrum(op, s)
    with C.Context(cp) as ctx:
        ctx.set_state(*s.fields())
        ke = ctx.kinetic_energy()
        fields = [ctx.real_field(k, True) for k in range(1 + ctx.nph)]
        sp = ctx.spectrum(truncate_state=True)
        g = ctx.get_state()
    ke_ref = O.kinetic_energy(op, s)
    assert max(abs(x - y) / max(abs(y), 1e-300) for x, y in zip(ke, ke_ref)) < 1e-12
    for f, ref in zip(fields, s.fields()):
        assert O.rel_l2(f, O.ifft(op, O.truncation(op, ref))) < 1e-14
    for k in ref_sp:
        assert O.rel_l2(sp[k][: len(ref_sp[k])], ref_sp[k]) < 1e-12, k
    for a, b in zip(g, s2.fields()):
        assert np.array_equal(a, b)


def test_filtered_step_full_size():
    """2048^2 (the 4-wave row transforms without the stager): one step."""
    N = 2048
    op, cp, _ = case("ternary", N, nu=2e-3, lam=5e-4, dt=1e-6)
    reals = (O.random_vorticity(op, 1),) + tuple(O.init_lens_real(op))
    s0 = full_state(op, reals)
    s1 = O.etd2rk(op, s0)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        ctx.step(1)
        g = ctx.get_state()
    for a, b, a0 in zip(g, s1.fields(), s0.fields()):
        assert O.rel_l2(a - a0, b - a0) < 1e-10


def test_filter_rejections():
    kw = dict(Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI, smooth_filter=1)
    with pytest.raises(C.ChnsError):
        C.Ensemble([C.params("ternary", **kw)])                    # single contexts only
    with C.Context(C.params("ternary", **kw)) as ctx:
        ctx.set_real_state(None, np.full((32, 32), 0.3), np.full((32, 32), 0.3))
        with pytest.raises(C.ChnsError):
            ctx.adj_forward(1, 1)                                  # the adjoint covers the plain step


@pytest.mark.parametrize("model", ["binary", "ternary"])
@pytest.mark.parametrize("N", [32, 64])
@pytest.mark.parametrize("graphs", [True, False])
def test_step_after_state_truncation(model, N, graphs):
    """k steps, spectrum()'s truncation of the phase fields, k steps, ... as the
    driver runs it every navg/2 steps (main.c:281-284).  The filter is not
    idempotent, so the steps after a truncation must read the truncated state
    (spectrum.c:24-26, then chem_potential.c:27-32): compared with the oracle's
    loop with the truncation in it."""
    op, cp, reals = case(model, N)
    s0 = full_state(op, reals)
    k = 5
    ref = O.run(op, s0, 4 * k, spectrum_every=k)
    with C.Context(cp) as ctx:
        ctx.graphs(graphs)
        ctx.set_state(*s0.fields())
        for _ in range(4):
            ctx.step(k)
            ctx.truncate_state()
        g = ctx.get_state()
    for a, b in zip(g, ref.fields()):
        assert O.rel_l2(a, b) < 1e-10
    # and the filtered trajectory does differ from one without the truncations
    plain = O.run(op, s0, 4 * k)
    assert max(O.rel_l2(a, b) for a, b in zip(g[1:], plain.fields()[1:])) > 1e-8

