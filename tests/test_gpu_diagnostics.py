"""Device diagnostics vs the oracle (free_energy.c, kinetic_energy.c,
perimeter.c, spectrum.c, dump path of main.c:332-358), through the C ABI."""
import math

import numpy as np
import pytest

import chns_amd as C
import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12   # reductions in a different order than the reference's loops


def state(model, N, steps=3):
    L = 2 * math.pi
    kw = dict(Nx=N, Ny=N, lx=L, ly=L, nu=5e-3, dt=5e-4, lam=1e-3, epsilon=1.5 * L / N, sigma=0.1,
              sigma_12=0.1, sigma_13=0.08, sigma_23=0.06)
    op = O.Params(model=model, **kw)
    wz = O.random_vorticity(op, 1)
    c = O.init_lens_real(op) if model == "ternary" else (O.init_noise_real(op, 3, 0.5, 0.3) if model == "binary" else ())
    s = O.run(op, O.state_from_real(op, wz, *c), steps)   # untruncated content in every mode
    return op, C.params(model, **kw), s


def relerr(a, b):
    return max(abs(x - y) / max(abs(y), 1e-300) for x, y in zip(a, b))


@pytest.mark.parametrize("model", ["binary", "ternary"])
@pytest.mark.parametrize("N", [32, 128, 512])
def test_free_energy_and_perimeter(model, N):
    op, cp, s = state(model, N)
    with C.Context(cp) as ctx:
        ctx.set_state(*s.fields())
        fe = ctx.free_energy()
        pe = ctx.perimeter()
    assert relerr(fe, O.free_energy(op, s)) < TOL
    assert pe == O.perimeter(op, s)


@pytest.mark.parametrize("model", ["ns", "binary", "ternary"])
@pytest.mark.parametrize("N", [32, 128, 512])
def test_kinetic_energy_and_dump(model, N):
    op, cp, s = state(model, N)
    with C.Context(cp) as ctx:
        ctx.set_state(*s.fields())
        ke = ctx.kinetic_energy()
        ke2, counts = ctx.kinetic_energy_counts()
        fields = [ctx.real_field(k, True) for k in range(1 + ctx.nph)]
    assert relerr(ke, O.kinetic_energy(op, s)) < TOL
    ref_ke, ref_counts = O.kinetic_energy(op, s, counts=True)
    assert ke2 == ke and counts == ref_counts                   # the point counts are exact
    for f, ref in zip(fields, s.fields()):
        assert O.rel_l2(f, O.ifft(op, O.truncation(op, ref))) < 1e-14


@pytest.mark.parametrize("model", ["ns", "binary", "ternary"])
@pytest.mark.parametrize("N", [32, 128, 512])
def test_spectrum_and_state_truncation(model, N):
    op, cp, s = state(model, N)
    ref, s2 = O.spectrum(op, s)
    with C.Context(cp) as ctx:
        ctx.set_state(*s.fields())
        sp = ctx.spectrum(truncate_state=True)
        g = ctx.get_state()
    for k in ref:
        assert O.rel_l2(sp[k][: len(ref[k])], ref[k]) < TOL, k
    for a, b in zip(g, s2.fields()):
        assert np.array_equal(a, b)


def test_diagnostics_repeatable_2048():
    op, cp, _ = state("ternary", 64, 0)
    cp = C.params("ternary", **{k: getattr(cp, k) for k in ("nu", "dt", "sigma_12", "sigma_13", "sigma_23")},
                  Nx=2048, Ny=2048, lx=2 * math.pi, ly=2 * math.pi, lam=5e-4, epsilon=1.5 * 2 * math.pi / 2048)
    from chns_amd import ic
    c1, c2 = ic.lens(2048, 2048, 2 * math.pi, 2 * math.pi, 1.5 * 2 * math.pi / 2048)
    outs = []
    for _ in range(2):
        with C.Context(cp) as ctx:
            ctx.set_real_state(None, c1, c2)
            ctx.step(2)
            outs.append((ctx.free_energy(), ctx.kinetic_energy(), ctx.perimeter(), ctx.spectrum(False)["ek"].tobytes()))
    assert outs[0] == outs[1]
