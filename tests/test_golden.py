"""Committed golden fixtures (tests/golden/chns_<model>_<N>.npz).

CPU: the oracle still reproduces every fixture.  GPU: the HIP library,
started from the stored initial state, reproduces the stored one-step and
100-step states (with spectrum()'s truncation every 10 steps) and the
diagnostics of the final state, without running the oracle.
"""
import ast
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = [(m, n) for m in ("ns", "binary", "ternary") for n in (16, 32)]
NAMES = ("cwz", "cphi1", "cphi2")


def load(model, N):
    z = np.load(os.path.join(HERE, "golden", f"chns_{model}_{N}.npz"))
    kw = dict(ast.literal_eval(str(z["params"])))
    return z, kw


def fields(z, tag):
    return [z[f"{tag}_{n}"] for n in NAMES if f"{tag}_{n}" in z]


@pytest.mark.parametrize("model,N", CASES)
def test_oracle_reproduces_fixture(model, N):
    z, kw = load(model, N)
    op = O.Params(model=model, **kw)
    s0 = O.State(*fields(z, "s0"))
    s1 = O.etd2rk(op, s0)
    sT = O.run(op, s0, int(z["steps"]), spectrum_every=int(z["spec_every"]))
    for a, b in zip(s1.fields(), fields(z, "s1")):
        assert O.rel_l2(a, b) < 1e-13
    for a, b in zip(sT.fields(), fields(z, "sT")):
        assert O.rel_l2(a, b) < 1e-12
    np.testing.assert_allclose(O.kinetic_energy(op, sT), z["ke"], rtol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("model,N", CASES)
def test_library_matches_fixture(model, N):
    import chns_amd as C

    z, kw = load(model, N)
    cp = C.params(model, **kw)
    with C.Context(cp) as ctx:
        ctx.set_state(*fields(z, "s0"))
        ctx.step(1)
        for a, b in zip(ctx.get_state(), fields(z, "s1")):
            assert O.rel_l2(a, b) < 1e-12
        ctx.set_state(*fields(z, "s0"))
        every = int(z["spec_every"])
        for _ in range(int(z["steps"]) // every):
            ctx.step(every)
            ctx.truncate_state()
        sT = ctx.get_state()
        ke = ctx.kinetic_energy()
        fe = ctx.free_energy() if model != "ns" else None
        pe = ctx.perimeter() if model != "ns" else None
        sp = ctx.spectrum(False)
    for a, b in zip(sT, fields(z, "sT")):
        assert O.rel_l2(a, b) <= 1e-10
    np.testing.assert_allclose(ke, z["ke"], rtol=1e-10)
    if model != "ns":
        np.testing.assert_allclose(fe, z["fe"], rtol=1e-10)
        assert [tuple(x) for x in z["perimeter"]] == [tuple(x) for x in pe]
    for k, v in sp.items():
        ref = z["spec_" + k]
        np.testing.assert_allclose(v[:len(ref)], ref[:len(v)], rtol=1e-9, atol=1e-12 * np.max(np.abs(ref)))
