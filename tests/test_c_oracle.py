"""The C/OpenMP restatement of the step (oracle/chns_step.c) against the numpy
oracle: it restates the same reference functions with its own FFT and must
agree to round-off.  It then stands in for the numpy oracle where that is too
slow (GPU parity at 1024^2 and 2048^2 over 100 steps) and is bench.py's CPU
baseline."""
import math

import numpy as np
import pytest

import oracle as O
from oracle import c_oracle as CO

TWO_PI = 2.0 * math.pi


def case(model, N, **over):
    kw = dict(Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI, nu=5e-3, dt=5e-4, lam=1e-3, epsilon=1.5 * TWO_PI / N,
              sigma=0.1, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06)
    kw.update(over)
    op = O.Params(model=model, **kw)
    nph = {"ns": 0, "binary": 1, "ternary": 2}[model]
    c = O.init_noise_real(op, 2, 1 / 3 if nph == 2 else 0.5, 0.05)[:nph] if nph else ()
    return op, O.state_from_real(op, O.random_vorticity(op, 1), *c)


@pytest.mark.parametrize("model", ["ns", "binary", "ternary"])
@pytest.mark.parametrize("N", [16, 32, 64])
def test_step_matches_numpy_oracle(model, N):
    op, s0 = case(model, N)
    r1 = O.etd2rk(op, s0)
    g1 = CO.run(op, s0, 1)
    for a, b, a0 in zip(g1.fields(), r1.fields(), s0.fields()):
        assert O.rel_l2(a - a0, b - a0) < 1e-12
    r = O.run(op, s0, 20, spectrum_every=5)
    g = CO.run(op, s0, 20, spectrum_every=5)
    for a, b in zip(g.fields(), r.fields()):
        assert O.rel_l2(a, b) < 1e-13


@pytest.mark.parametrize("model", ["ns", "ternary"])
def test_forcing_and_reference_defaults(model):
    """Kolmogorov forcing; and the reference defaults (L = 4 pi, gamma_1 < 0)."""
    op, s0 = case(model, 32, f_amp=0.3, f_k=3)
    for a, b in zip(CO.run(op, s0, 10).fields(), O.run(op, s0, 10).fields()):
        assert O.rel_l2(a, b) < 1e-13
    p = O.Params(model=model, Nx=32, Ny=32)
    s = O.state_from_real(p, O.random_vorticity(p, 3), *O.init_ternary_noise_real(p)[: (2 if model == "ternary" else 0)])
    for a, b in zip(CO.run(p, s, 5).fields(), O.run(p, s, 5).fields()):
        assert O.rel_l2(a, b) < 1e-13


def test_threads_do_not_change_results():
    """The OpenMP split is over independent rows and columns: bitwise equal."""
    op, s0 = case("ternary", 64)
    old = CO.threads()
    try:
        CO.threads(1)
        a = CO.run(op, s0, 3)
        CO.threads(max(2, old))
        b = CO.run(op, s0, 3)
    finally:
        CO.threads(old)
    for x, y in zip(a.fields(), b.fields()):
        assert np.array_equal(x, y)


def test_rejects_what_it_does_not_cover():
    with pytest.raises(ValueError):
        CO.Stepper(O.Params(model="ternary", Nx=32, Ny=32, smooth_filter=True))
    with pytest.raises(ValueError):
        CO.Stepper(O.Params(model="ternary", Nx=32, Ny=32, dealias=1))
