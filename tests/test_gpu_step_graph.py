"""chns_step replays a HIP graph of 8 captured steps (the remainder of a
call launched directly); with event timing on it launches every pass
directly.  Both paths run the same kernels in the same order on the same
buffers, so they must agree bit for bit after every call: replays are
interleaved with eager work (state reads, uploads, the spectrum truncation,
diagnostics) and synchronisations, each call checked against the direct
launches.  Parity of the step itself with the oracle is covered by the other
GPU tests, which replay the graph whenever they take 8 steps or more (timing
is off by default).
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


def same(g, d):
    for a, b in zip(g.get_state(), d.get_state()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("model,N,over", [
    ("ternary", 64, {}), ("binary", 128, {}), ("ns", 64, {}),
    ("ternary", 32, dict(gravity=1, vpm=1, sponge=1)), ("ternary", 32, dict(smooth_filter=1)),
    ("binary", 64, dict(dealias=1)),
])
def test_graph_replays_match_direct_launches(model, N, over):
    op, cp, reals = case(model, N, **over)
    s0 = O.State(*[O.dfft(r) for r in reals])
    with C.Context(cp) as g, C.Context(cp) as d:
        d.timing(True)
        for ctx in (g, d):
            ctx.set_state(*s0.fields())
        for k in (8, 11, 3, 16):
            g.step(k)
            d.step(k)
            same(g, d)
            d.timing_read()
        assert np.isfinite(g.get_state()[0]).all()
        # a state upload between replays (pass A runs again, outside the graph)
        for ctx in (g, d):
            ctx.set_state(*s0.fields())
        g.step(9)
        d.step(9)
        same(g, d)
        if model != "ns":
            # the in-place truncation of spectrum() and a diagnostic between replays
            for ctx in (g, d):
                ctx.truncate_state()
                ctx.kinetic_energy()
            g.step(8)
            d.step(8)
            same(g, d)


def test_graph_matches_oracle_after_reupload():
    """The graph captured on the first call keeps stepping correctly after the
    state is replaced (20 steps, rel-L2 1e-10 against the oracle)."""
    op, cp, reals = case("ternary", 64)
    s0 = O.State(*[O.dfft(r) for r in reals])
    ref = O.run(op, s0, 20)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        ctx.step(8)
        ctx.set_real_state(*reals)   # a different state: the truncated one
        ctx.set_state(*s0.fields())
        ctx.step(20)
        g = ctx.get_state()
    for a, b in zip(g, ref.fields()):
        assert O.rel_l2(a, b) < 1e-10


@pytest.mark.parametrize("model", ["ternary", "ns"])
def test_short_first_call_leaves_the_graph_ready(model):
    """The graph is captured on the first call whatever its length: a 5-step
    first call (a warm-up) launches directly, and the next call replays.  The
    results equal a context with graphs switched off, bit for bit."""
    op, cp, reals = case(model, 64)
    s0 = O.State(*[O.dfft(r) for r in reals])
    with C.Context(cp) as g, C.Context(cp) as d:
        d.graphs(False)
        for ctx in (g, d):
            ctx.set_state(*s0.fields())
            ctx.step(5)
        same(g, d)
        for ctx in (g, d):
            ctx.step(20)
        same(g, d)
        g.graphs(False)                      # switching off between calls
        d.graphs(True)
        for ctx in (g, d):
            ctx.step(9)
        same(g, d)
