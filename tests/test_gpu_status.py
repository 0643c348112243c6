"""Divergence status (chns_status): the device flag raised by the ETD2RK
update on a non-finite mode, and the free-energy guard of free_energy.c:151-154
(the reference aborts when ephi >= 100).  The conditions are forced from the
host: a NaN placed in the uploaded state, and an initial phase field far
outside [0, 1] (c ~ 10, as a bad input file would give).  Every kernel runs to
its normal end; only the values are bad."""
import math

import numpy as np
import pytest

import chns_amd as C
import oracle as O

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi


def params(model, N=32, **over):
    kw = dict(Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI, nu=5e-3, dt=5e-4, lam=1e-3, epsilon=1.5 * TWO_PI / N,
              sigma=0.1, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06)
    kw.update(over)
    return O.Params(model=model, **kw), C.params(model, **kw)


@pytest.mark.parametrize("model", ["ns", "binary", "ternary"])
@pytest.mark.parametrize("graphs", [True, False])
def test_nonfinite_state_raises_the_flag(model, graphs):
    op, cp = params(model)
    nf = {"ns": 1, "binary": 2, "ternary": 3}[model]
    reals = [O.random_vorticity(op, 1)] + [np.full((32, 32), 0.3)] * (nf - 1)
    with C.Context(cp) as ctx:
        ctx.graphs(graphs)
        ctx.set_real_state(*reals)
        ctx.step(9)
        assert ctx.status() == 0
        st = ctx.get_state()
        st[nf - 1][2, 3] = complex(np.nan, 0.0)      # the last field: c2, c or omega
        ctx.set_state(*st)
        assert ctx.status() == 0                     # an upload clears the flags
        ctx.step(9)
        assert ctx.status() == C.CHNS_FLAG_NONFINITE
        rc = ctx.lib.chns_status(ctx.h, None)
        assert rc == C.CHNS_E_DIVERGED
        assert "non-finite" in ctx.lib.chns_last_error(ctx.h).decode()
        ctx.set_real_state(*reals)
        assert ctx.status() == 0


@pytest.mark.parametrize("model", ["binary", "ternary"])
def test_free_energy_guard(model):
    """ephi >= 100 sets CHNS_FLAG_ENERGY at the next chns_free_energy, as the
    reference's guard would abort the run there."""
    op, cp = params(model)
    nph = 2 if model == "ternary" else 1
    good = O.init_noise_real(op, 3, 1.0 / 3.0 if nph == 2 else 0.5, 0.05)[:nph]
    x = (np.arange(32) * op.dx)[:, None] * np.ones((1, 32))
    bad = [10.0 + 0.1 * np.cos(x)] + list(good[1:])
    with C.Context(cp) as ctx:
        ctx.set_real_state(None, *good)
        e, _ = ctx.free_energy()
        assert e < 100 and ctx.status() == 0
        ctx.set_real_state(None, *bad)
        e, _ = ctx.free_energy()
        ref, _ = O.free_energy(op, O.state_from_real(op, None, *bad))
        assert e >= 100 and abs(e - ref) <= 1e-10 * ref
        assert ctx.status() == C.CHNS_FLAG_ENERGY
        ctx.set_real_state(None, *good)
        assert ctx.status() == 0


def test_ensemble_instance_flag():
    """Each ensemble instance has its own flag, raised by the batched update."""
    op, cp = params("ternary")
    reals = [None] + list(O.init_noise_real(op, 2, 1.0 / 3.0, 0.05))
    with C.Ensemble([cp, cp, cp]) as ens:
        for b in range(3):
            ens[b].set_real_state(*reals)
        st = ens[1].get_state()
        st[1][1, 1] = complex(np.inf, 0.0)
        ens[1].set_state(*st)
        ens.step(9)
        ens.sync()
        assert [ens[b].status() for b in range(3)] == [0, C.CHNS_FLAG_NONFINITE, 0]
