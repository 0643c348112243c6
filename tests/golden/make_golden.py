"""Write the golden fixtures of the CHNS-2D step (oracle outputs, small N).

    python tests/golden/make_golden.py

For each model (ns, binary, ternary) and N in {16, 32}: the seeded initial
state, the state after one ETD2RK step, after 100 steps with spectrum()'s
truncation every 10 steps (main.c:281-284), and the diagnostics of the final
state (free_energy.c, kinetic_energy.c, perimeter.c, spectrum.c).  The
reference ships no fixtures of its own; these freeze the oracle's outputs so
that the GPU tests can check against stored data and a change of the oracle
shows up as a failing CPU test.
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

MODELS = ("ns", "binary", "ternary")
SIZES = (16, 32)
STEPS = 100
SPEC_EVERY = 10


def params(model, N):
    """Fixture parameters: L = 2 pi, epsilon = 1.5 dx, sigma_ij = 0.1/0.08/0.06."""
    L = 2 * math.pi
    return dict(Nx=N, Ny=N, lx=L, ly=L, nu=5e-3, dt=5e-4, lam=1e-3, epsilon=1.5 * L / N,
                sigma=0.1, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06,
                f_amp=0.5 if model == "ns" else 0.0, f_k=2 if model == "ns" else 0)


def initial(op, model):
    wz = O.random_vorticity(op, 1)
    if model == "ternary":
        return O.state_from_real(op, wz, *O.init_noise_real(op, 2, 1.0 / 3.0, 0.05))
    if model == "binary":
        return O.state_from_real(op, wz, *O.init_noise_real(op, 1, 0.5, 0.1))
    return O.state_from_real(op, wz)


def fixture(model, N):
    kw = params(model, N)
    op = O.Params(model=model, **kw)
    s0 = initial(op, model)
    s1 = O.etd2rk(op, s0)
    sT = O.run(op, s0, STEPS, spectrum_every=SPEC_EVERY)
    out = {"params": np.array(repr(sorted(kw.items()))), "steps": STEPS, "spec_every": SPEC_EVERY}
    for tag, s in (("s0", s0), ("s1", s1), ("sT", sT)):
        for name, f in zip(("cwz", "cphi1", "cphi2"), s.fields()):
            out[f"{tag}_{name}"] = f
    out["ke"] = np.array(O.kinetic_energy(op, sT))
    if model != "ns":
        out["fe"] = np.array(O.free_energy(op, sT))
        out["perimeter"] = np.array(O.perimeter(op, sT))
    sp, _ = O.spectrum(op, sT)
    for k, v in sp.items():
        out["spec_" + k] = v
    return out


def path(model, N):
    return os.path.join(HERE, f"chns_{model}_{N}.npz")


def main():
    for model in MODELS:
        for N in SIZES:
            np.savez_compressed(path(model, N), **fixture(model, N))
            print("wrote", os.path.relpath(path(model, N)))


if __name__ == "__main__":
    main()
