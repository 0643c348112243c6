"""The chns2d time-loop driver (main.c:24-368 on the C ABI).

CPU tests check the initial-condition generators (init.c:220-294) and the
raw input path (read_array.c:11-36) through the driver's ``ic_only`` mode,
which stops before any device call.  GPU tests run the full loop and compare
every output file (data/*.dat, spectra/*, vort/*, phase_field/*) with the
oracle's restatement of the same loop.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.environ.get("CHNS2D", os.path.join(ROOT, "ternary-chns_amd", "chns_amd", "chns2d"))


def run(tmp, **kw):
    args = [DRIVER] + [f"{k}={v}" for k, v in kw.items()] + [f"out={tmp}"]
    return subprocess.run(args, capture_output=True, text=True, timeout=300)


def read_raw(path, N):
    return np.fromfile(path, dtype=np.float64).reshape(N, N)


@pytest.fixture(autouse=True)
def _need_driver():
    if not os.path.exists(DRIVER):
        pytest.fail("chns2d is not built (python ternary-chns_amd/build.py)")


# ---------------------------------------------------------------------------
# initial conditions: bit-exact against the oracle (CPU only)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nranks", [1, 2, 4])
def test_ternary_noise_ic_bit_exact(tmp_path, nranks):
    N = 32
    r = run(tmp_path, Nx=N, ic_only=1, nranks=nranks)
    assert r.returncode == 0, r.stderr
    c1, c2 = O.init_ternary_noise_real(O.Params(Nx=N, Ny=N), nranks)
    assert np.array_equal(read_raw(tmp_path / "init_field" / "ic1.bin", N), c1)
    assert np.array_equal(read_raw(tmp_path / "init_field" / "ic2.bin", N), c2)


def test_binary_droplet_ic(tmp_path):
    N = 64
    r = run(tmp_path, model="binary", Nx=N, ic_only=1)
    assert r.returncode == 0, r.stderr
    got = read_raw(tmp_path / "init_field" / "ic1.bin", N)
    ref = O.init_binary_droplet_real(O.Params(Nx=N, Ny=N, model="binary"))
    # C tanh against numpy's tanh: equal to the last bit or one ulp apart
    assert np.max(np.abs(got - ref)) <= 2.3e-16
    # the 0.5 level lies on the circle r = pi/4.5 about the centre
    L, dx = 4 * math.pi, 4 * math.pi / N
    x = np.arange(N)[:, None] * dx - L / 2
    y = np.arange(N)[None, :] * dx - L / 2
    inside = np.sqrt(x * x + y * y) < math.pi / 4.5
    assert np.all(got[inside] > 0.5) and np.all(got[~inside] <= 0.5)


def test_read_path_round_trip(tmp_path):
    N = 32
    rng = np.random.default_rng(5)
    a, b = rng.random((N, N)), rng.random((N, N))
    (tmp_path / "in").mkdir()
    a.tofile(tmp_path / "in" / "a.in")
    b.tofile(tmp_path / "in" / "b.in")
    r = run(tmp_path, Nx=N, init="read", read1=tmp_path / "in" / "a.in", read2=tmp_path / "in" / "b.in", ic_only=1)
    assert r.returncode == 0, r.stderr
    assert np.array_equal(read_raw(tmp_path / "init_field" / "ic1.bin", N), a)
    assert np.array_equal(read_raw(tmp_path / "init_field" / "ic2.bin", N), b)


@pytest.mark.parametrize("args", [
    dict(bogus=1),                                  # unknown parameter
    dict(maxiter=10, nfile=10),                     # navg < 2
    dict(nranks=3, ic_only=1),                      # ranks must divide Nx
    dict(model="binary", init="noise", ic_only=1),  # noise IC is ternary-only
])
def test_driver_rejects_bad_arguments(tmp_path, args):
    kw = dict(Nx=32)
    kw.update(args)
    r = run(tmp_path, **kw)
    assert r.returncode == 2, (r.returncode, r.stderr)


def test_missing_input_file_fails(tmp_path):
    r = run(tmp_path, Nx=32, init="read", read1=tmp_path / "nope", ic_only=1)
    assert r.returncode == 1 and "cannot read" in r.stderr


# ---------------------------------------------------------------------------
# full loop on the GPU against the oracle's main.c restatement
# ---------------------------------------------------------------------------
CASES = {
    # reference defaults (global.h) at N = 32: TERNARY, glibc noise IC
    "ternary": (dict(Nx=32), dict(Nx=32, Ny=32), "noise"),
    # BINARY defaults with the init.c droplet
    "binary": (dict(model="binary", Nx=32), dict(Nx=32, Ny=32, model="binary"), "droplet"),
    # Navier-Stokes with Kolmogorov forcing from rest
    "ns": (dict(model="ns", Nx=32, f_amp=1.0, f_k=2), dict(Nx=32, Ny=32, model="ns", f_amp=1.0, f_k=2), "zero"),
    # SMOOTH_FILTER (macro.h:15): its truncation is not idempotent, so the steps
    # after each spectrum() truncation (every navg/2 = 2 steps) see it
    "ternary_filter": (dict(Nx=32, smooth_filter=1), dict(Nx=32, Ny=32, smooth_filter=True), "noise"),
    "binary_filter": (dict(model="binary", Nx=32, smooth_filter=1),
                      dict(Nx=32, Ny=32, model="binary", smooth_filter=True), "droplet"),
    # the 2/3 rule instead of the reference's N/4 circle
    "ternary_23": (dict(Nx=32, dealias=1), dict(Nx=32, Ny=32, dealias=1), "noise"),
    # the host arrays hold the state across every reference-named call
    # (chns_ref.h mirror mode: upload, run, write back, as the reference does)
    "ternary_mirror": (dict(Nx=32, mirror=1), dict(Nx=32, Ny=32), "noise"),
    "binary_mirror": (dict(model="binary", Nx=32, mirror=1), dict(Nx=32, Ny=32, model="binary"), "droplet"),
}


def _cols(path):
    return np.loadtxt(path, ndmin=2)


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_driver_matches_oracle_loop(tmp_path, case):
    dargs, okw, init = CASES[case]
    maxiter, nfile = 8, 2
    r = run(tmp_path, maxiter=maxiter, nfile=nfile, init=init, **dargs)
    assert r.returncode == 0, r.stderr
    p = O.Params(**okw)
    model = {O.MODEL_NS: "ns", O.MODEL_BINARY: "binary", O.MODEL_TERNARY: "ternary"}[p.model]
    if model == "ternary":
        s0 = O.state_from_real(p, None, *O.init_ternary_noise_real(p))
    elif model == "binary":
        s0 = O.state_from_real(p, None, O.init_binary_droplet_real(p))
    else:
        s0 = O.state_from_real(p)
    rec, _ = O.main_loop(p, s0, maxiter, nfile)
    nph = {"ns": 0, "binary": 1, "ternary": 2}[model]
    d = tmp_path / "data"

    ken = _cols(d / "ken_t.dat")
    ref = np.array(rec["ken"])
    assert ken.shape == ref.shape == (maxiter + 1, 2 + nph)
    assert np.array_equal(ken[:, 0], np.round(ref[:, 0], 6))
    np.testing.assert_allclose(ken[:, 1:], ref[:, 1:], rtol=1e-9, atol=2e-12)
    if nph:
        fe = _cols(d / "free_en_t.dat")
        np.testing.assert_allclose(fe[:, 1:], np.array(rec["fe"])[:, 1:], rtol=1e-9, atol=2e-12)
        names = ["gamma.dat"] if nph == 1 else ["gamma1.dat", "gamma2.dat"]
        for ph, name in enumerate(names):
            # C prints 0/0 (no points inside yet) as "-nan", Python as "nan"
            lines = (d / name).read_text().replace("-nan", "nan").splitlines()
            assert lines == ["%g \t %g \t %g \t %g" % row for row in rec["gamma"][ph]]
        assert (d / ("volume_droplet.dat" if nph == 1 else "volume_phase1.dat")).stat().st_size == 0

    navg = maxiter // nfile
    assert len(rec["spectra"]) == 1 + maxiter // (navg // 2)
    kxm = int(round((p.Nx / 2.0) * p.factor_x))
    keys = {0: ["ek", "ensk"], 1: ["ek", "ensk", "sph"], 2: ["ek", "ensk", "sph1", "sph2", "sp12"]}[nph]
    for i, sp in enumerate(rec["spectra"]):
        got = _cols(tmp_path / "spectra" / f"spectrum.{i}")
        ns = p.nshell
        np.testing.assert_allclose(got[:, 0], np.arange(1, ns) * (2 * math.pi / p.lx), rtol=1e-5)
        for c, k in enumerate(keys):
            refc = sp[k][1:ns]
            np.testing.assert_allclose(got[:, 1 + c], refc, rtol=1e-9, atol=1e-12 * max(np.max(np.abs(refc)), 1e-300))
        for fn, k in (("Xspectrum", "ekx"), ("Yspectrum", "eky")):
            got = _cols(tmp_path / "spectra" / f"{fn}.{i}")
            assert np.array_equal(got[:, 0], np.arange(1, kxm))
            refc = sp[k][1:kxm]
            np.testing.assert_allclose(got[:, 1], refc, rtol=1e-9, atol=1e-12 * max(np.max(np.abs(refc)), 1e-300))

    assert len(rec["dumps"]) == 1 + maxiter // navg
    for fl, fields in enumerate(rec["dumps"]):
        names = ["vort/wz"] + (["phase_field/phi"] if nph == 1 else []) + \
                (["phase_field/phi1", "phase_field/phi2"] if nph == 2 else [])
        for name, ref in zip(names, fields):
            got = read_raw(tmp_path / f"{name}.{fl}", p.Nx)
            scale = max(np.max(np.abs(ref)), 1e-300)
            assert np.max(np.abs(got - ref)) <= 1e-10 * scale + 1e-300, (name, fl)


@pytest.mark.gpu
@pytest.mark.parametrize("mirror", [0, 1])
def test_binary_line_skipped_without_phase_points(tmp_path, mirror):
    """BINARY kinetic_energy() writes its line only when some point has
    c >= 0.5 (kinetic_energy.c:103-108): a read-in field below 0.5 everywhere
    leaves data/ken_t.dat empty while the other files keep their lines, and
    the oracle's loop records no row either."""
    N, maxiter, nfile = 32, 4, 2
    rng = np.random.default_rng(7)
    c = 0.3 + 0.05 * rng.random((N, N))
    (tmp_path / "in").mkdir()
    c.tofile(tmp_path / "in" / "c.in")
    print(f"chns2d binary read, all c < 0.5, mirror={mirror}", flush=True)
    r = run(tmp_path, model="binary", Nx=N, init="read", read1=tmp_path / "in" / "c.in", maxiter=maxiter,
            nfile=nfile, mirror=mirror)
    assert r.returncode == 0, r.stderr
    d = tmp_path / "data"
    assert (d / "ken_t.dat").stat().st_size == 0
    assert len(_cols(d / "free_en_t.dat")) == maxiter + 1
    p = O.Params(Nx=N, Ny=N, model="binary")
    rec, _ = O.main_loop(p, O.state_from_real(p, None, c), maxiter, nfile)
    assert rec["ken"] == [] and len(rec["fe"]) == maxiter + 1


@pytest.mark.gpu
@pytest.mark.parametrize("mirror", [0, 1])
def test_free_energy_guard_exits_3(tmp_path, mirror):
    """free_energy.c:151-154: the run stops once ephi >= 100.  A BINARY run
    with dt far above the explicit k^2 P(c) limit diverges within a few steps;
    the driver (free_energy() of chns_ref.h) writes that step's free-energy
    line, then exits with status 3 and says why on stderr."""
    N = 32
    print(f"chns2d binary droplet, dt = 5 (diverges), mirror={mirror}", flush=True)
    r = run(tmp_path, model="binary", Nx=N, init="droplet", dt=5.0, maxiter=50, nfile=10, mirror=mirror)
    assert r.returncode == 3, (r.returncode, r.stderr)
    assert "free_energy" in r.stderr
    lines = (tmp_path / "data" / "free_en_t.dat").read_text().splitlines()
    assert 2 <= len(lines) <= 51
    last = float(lines[-1].split()[1])
    assert not (last < 100.0)              # the line that tripped the guard (or nan)
    assert all(float(x.split()[1]) < 100.0 for x in lines[:-1])
