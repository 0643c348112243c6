"""GPU parity of the optional physics switches (macro.h GRAVITY, VPM, SPONGE)
against the oracle (oracle/chns_oracle.py, pinned in tests/test_optional_physics.py).

Bar: one-step increments at 1e-10 and 20-step trajectories at rel-L2 1e-10,
as for the plain step.
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
              epsilon=1.5 * TWO_PI / N, sigma=0.1, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06)
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


SWITCHES = [
    ("binary", dict(gravity=1)),
    ("ternary", dict(gravity=1)),
    ("ns", dict(vpm=1)),
    ("binary", dict(vpm=1)),
    ("ternary", dict(vpm=1)),
    ("ns", dict(vpm=1, sponge=1)),
    ("ternary", dict(vpm=1, sponge=1, gravity=1)),
]


@pytest.mark.parametrize("model,sw", SWITCHES)
@pytest.mark.parametrize("N", [32, 64])
def test_switch_parity(model, sw, N):
    op, cp, reals = case(model, N, **sw)
    s0 = O.state_from_real(op, *reals)
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


def test_switches_change_the_step():
    """Each switch changes the increment (so the tests above are not vacuous)."""
    op, cp, reals = case("ternary", 32)
    s0 = O.state_from_real(op, *reals)
    base = O.etd2rk(op, s0)
    for sw in (dict(gravity=1), dict(vpm=1), dict(vpm=1, sponge=1)):
        op2, cp2, _ = case("ternary", 32, **sw)
        with C.Context(cp2) as ctx:
            ctx.set_state(*s0.fields())
            ctx.step(1)
            g = ctx.get_state()
        assert O.rel_l2(g[0] - s0.cwz, base.cwz - s0.cwz) > 1e-3, sw


def test_full_size_sponge_one_step():
    """2048^2 (row transforms with one row per workgroup): the sponge(y) profile is applied
    at the right real-space points."""
    N = 2048
    op, cp, _ = case("ternary", N, vpm=1, sponge=1, gravity=1, nu=2e-3, lam=5e-4, dt=1e-6)
    reals = (O.random_vorticity(op, 1),) + tuple(O.init_lens_real(op))
    s0 = O.state_from_real(op, *reals)
    s1 = O.etd2rk(op, s0)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        ctx.step(1)
        g = ctx.get_state()
    for a, b, a0 in zip(g, s1.fields(), s0.fields()):
        assert O.rel_l2(a - a0, b - a0) < 1e-10


@pytest.mark.parametrize("model", ["binary", "ternary"])
def test_full_size_walls_one_step(model):
    """2048^2 with the walls and no sponge: pass D's one-row-per-workgroup path with the
    conservative vorticity advection, where the walls scale A_i and u.grad c_i by 1 - chi
    and put -u chi/eta, -v chi/eta into Q and P (the 32^2/64^2 cases run the same form
    through the several-rows-per-workgroup path)."""
    N = 2048
    op, cp, _ = case(model, N, vpm=1, nu=2e-3, lam=5e-4, dt=1e-6)
    if model == "ternary":
        reals = (O.random_vorticity(op, 2),) + tuple(O.init_lens_real(op))
    else:
        reals = (O.random_vorticity(op, 2),) + O.init_noise_real(op, 3, 0.5, 0.1)
    s0 = O.state_from_real(op, *reals)
    s1 = O.etd2rk(op, s0)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        ctx.step(1)
        g = ctx.get_state()
    for a, b, a0 in zip(g, s1.fields(), s0.fields()):
        assert O.rel_l2(a - a0, b - a0) < 1e-10


def test_switch_rejections():
    kw = dict(Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI)
    with pytest.raises(C.ChnsError):
        C.Context(C.params("ternary", sponge=1, **kw))            # sponge needs vpm
    with pytest.raises(C.ChnsError):
        C.Context(C.params("ns", gravity=1, **kw))                # gravity needs a phase field
    with pytest.raises(C.ChnsError):
        C.Ensemble([C.params("ternary", vpm=1, **kw)])            # single contexts only
    with C.Context(C.params("ternary", gravity=1, **kw)) as ctx:
        ctx.set_real_state(None, np.full((32, 32), 0.3), np.full((32, 32), 0.3))
        with pytest.raises(C.ChnsError):
            ctx.adj_forward(1, 1)                                 # the adjoint covers the plain step
