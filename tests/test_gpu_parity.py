"""GPU parity of the fused ETD2RK step against the oracle (through the C ABI).

Bar: relative L2 of omega, c1, c2 <= 1e-10 after 100 steps,
mass (the k = 0 mode of each phase field) conserved to 1e-13.
"""
import math

import numpy as np
import pytest

import chns_amd as C
import oracle as O

pytestmark = pytest.mark.gpu

TOL_TRAJ = 1e-10     # rel-L2 after 100 steps (FP64)
TOL_MASS = 1e-13     # relative drift of the k=0 mode


def case(model, N, L=2 * math.pi, **over):
    kw = dict(Nx=N, Ny=N, lx=L, ly=L, nu=5e-3, dt=5e-4, lam=1e-3, epsilon=1.5 * L / N,
              sigma=0.1, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06)
    kw.update(over)
    op = O.Params(model=model, **kw)
    wz = O.random_vorticity(op, 1)
    if model == "ternary":
        c1, c2 = O.init_noise_real(op, 2, 1.0 / 3.0, 0.05)
        reals = (wz, c1, c2)
    elif model == "binary":
        (c1,) = O.init_noise_real(op, 1, 0.5, 0.1)
        reals = (wz, c1)
    else:
        reals = (wz,)
    return op, C.params(model, **kw), reals


def rel(a, b):
    return O.rel_l2(a, b)


@pytest.mark.parametrize("model", ["ns", "binary", "ternary"])
@pytest.mark.parametrize("N", [16, 32, 64, 128])
def test_init_dfft_truncation(model, N):
    op, cp, reals = case(model, N)
    s0 = O.state_from_real(op, *reals)
    with C.Context(cp) as ctx:
        ctx.set_real_state(*reals)
        g = ctx.get_state()
    for a, b in zip(g, s0.fields()):
        assert rel(a, b) < 1e-14


@pytest.mark.parametrize("model", ["ns", "binary", "ternary"])
@pytest.mark.parametrize("N", [16, 32, 64, 128])
def test_one_step_increment(model, N):
    """The change of the state over one step matches to 1e-10 (tests N itself)."""
    op, cp, reals = case(model, N)
    s0 = O.state_from_real(op, *reals)
    s1 = O.etd2rk(op, s0)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        ctx.step(1)
        g = ctx.get_state()
    for a, b, a0 in zip(g, s1.fields(), s0.fields()):
        assert rel(a - a0, b - a0) < 1e-10


@pytest.mark.parametrize("model", ["ns", "binary", "ternary"])
@pytest.mark.parametrize("N", [32, 64, 128])
def test_100_steps(model, N):
    op, cp, reals = case(model, N)
    s0 = O.state_from_real(op, *reals)
    ref = O.run(op, s0, 100)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        ctx.step(100)
        g = ctx.get_state()
    for a, b in zip(g, ref.fields()):
        assert rel(a, b) <= TOL_TRAJ
    for a, b in zip(g[1:], s0.fields()[1:]):
        assert abs(a[0, 0] - b[0, 0]) <= TOL_MASS * abs(b[0, 0])


@pytest.mark.parametrize("model", ["binary", "ternary"])
def test_spectrum_truncation_cadence(model):
    """spectrum() truncates the phase fields every navg/2 steps (main.c:281-284)."""
    op, cp, reals = case(model, 64)
    s0 = O.state_from_real(op, *reals)
    ref = O.run(op, s0, 40, spectrum_every=10)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        for _ in range(4):
            ctx.step(10)
            ctx.truncate_state()
        g = ctx.get_state()
    for a, b in zip(g, ref.fields()):
        assert rel(a, b) <= TOL_TRAJ


def test_reference_defaults_ternary():
    """Reference defaults (L = 4 pi, eps = 3 dx, gamma1 < 0) at N = 64, glibc-rand IC."""
    op = O.Params(model="ternary", Nx=64, Ny=64)
    cp = C.params("ternary", Nx=64, Ny=64)
    c1, c2 = O.init_ternary_noise_real(op, nranks=1)
    s0 = O.state_from_real(op, None, c1, c2)
    ref = O.run(op, s0, 100)
    with C.Context(cp) as ctx:
        ctx.set_real_state(None, c1, c2)
        ctx.step(100)
        g = ctx.get_state()
    for a, b in zip(g, ref.fields()):
        assert rel(a, b) <= TOL_TRAJ


def test_kolmogorov_steady_state():
    """NS + forcing f sin(k_f y): omega -> f/(nu k_f^2) sin(k_f y), exact under ETD."""
    N, kf, F0, nu, dt = 64, 4, 0.1, 5e-3, 0.5
    op, cp, _ = case("ns", N, nu=nu, dt=dt, f_amp=F0, f_k=kf)
    zero = np.zeros((N, N // 2 + 1), np.complex128)
    with C.Context(cp) as ctx:
        ctx.set_state(zero)
        ctx.step(4000)
        (w,) = ctx.get_state()
    t = 4000 * dt
    amp = F0 / (nu * kf * kf) * (1.0 - math.exp(-nu * kf * kf * t))
    expect = np.zeros_like(zero)
    expect[0, kf] = -0.5j * amp * N * N
    assert rel(w, expect) < 1e-12


@pytest.mark.parametrize("model", ["binary", "ternary"])
@pytest.mark.parametrize("N", [256, 512])
def test_larger_grids_10_steps(model, N):
    op, cp, reals = case(model, N, dt=5e-5)
    s0 = O.state_from_real(op, *reals)
    ref = O.run(op, s0, 10)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        ctx.step(10)
        g = ctx.get_state()
    for a, b, a0 in zip(g, ref.fields(), s0.fields()):
        assert rel(a, b) <= TOL_TRAJ
        assert rel(a - a0, b - a0) < 1e-11


@pytest.mark.parametrize("N", [1024, 2048])
@pytest.mark.parametrize("model", ["ns", "binary", "ternary"])
def test_full_size_one_step(model, N):
    """1024^2 (two rows per row-kernel workgroup) and 2048^2 (one row; LDS-staged
    inputs for the phase models, direct row loads for NS): one step vs the oracle."""
    op, cp, _ = case(model, N, nu=2e-3, lam=5e-4, dt=2.5e-6)
    wz = O.random_vorticity(op, 1)
    if model == "ternary":
        reals = (wz,) + tuple(O.init_lens_real(op))
    elif model == "binary":
        reals = (wz,) + O.init_noise_real(op, 7, 0.5, 0.1)
    else:
        reals = (wz,)
    s0 = O.state_from_real(op, *reals)
    s1 = O.etd2rk(op, s0)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        ctx.step(1)
        g = ctx.get_state()
    for a, b, a0 in zip(g, s1.fields(), s0.fields()):
        assert rel(a, b) <= TOL_TRAJ
        assert rel(a - a0, b - a0) < 1e-10


@pytest.mark.parametrize("N", [1024, 2048])
def test_full_size_properties(N):
    """At the benchmark sizes: mass conservation, finiteness, bitwise repeatability."""
    op, cp, _ = case("ternary", N, nu=2e-3, lam=5e-4, dt=1e-6)
    c1, c2 = O.init_lens_real(op)
    outs = []
    for _ in range(2):
        with C.Context(cp) as ctx:
            ctx.set_real_state(None, c1, c2)
            m0 = [f[0, 0] for f in ctx.get_state()[1:]]
            ctx.step(20)
            g = ctx.get_state()
        outs.append(g)
        for f in g:
            assert np.all(np.isfinite(f))
        for f, m in zip(g[1:], m0):
            assert abs(f[0, 0] - m) <= TOL_MASS * abs(m)
    for a, b in zip(*outs):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_full_size_real_io():
    """2048^2 real-space boundary: dfft + truncation of real fields and the real dump (the
    row transforms with one row per workgroup order real-space points differently)."""
    N = 2048
    op, cp, _ = case("ternary", N, nu=2e-3, lam=5e-4, dt=1e-6)
    wz = O.random_vorticity(op, 3)
    c1, c2 = O.init_lens_real(op)
    s0 = O.state_from_real(op, wz, c1, c2)
    with C.Context(cp) as ctx:
        ctx.set_real_state(wz, c1, c2)
        g = ctx.get_state()
        r1 = ctx.real_field(1, truncate=False)
        r0 = ctx.real_field(0, truncate=True)
        fe = ctx.free_energy()
    for a, b in zip(g, s0.fields()):
        assert rel(a, b) < 1e-14
    assert rel(r1, O.ifft(op, s0.cphi1)) < 1e-13
    assert rel(r0, O.ifft(op, O.truncation(op, s0.cwz))) < 1e-13
    ref = O.free_energy(op, s0)
    assert all(abs(x - y) <= 1e-12 * abs(y) for x, y in zip(fe, ref))


@pytest.mark.parametrize("dealias", [0, 1])
def test_config_binary_quench_128_1000_steps(dealias):
    """Binary spinodal quench, 128^2, c0 = 0.5 + U[-0.1, 0.1] (SplitMix64 seed 1),
    eps = 1.5 dx, sigma = 0.1, lambda = 1e-3, nu = 0.01, dt = 1e-3, 1000 steps, with
    the reference's N/4 cut (dealias 0) and with the 2/3 rule (dealias 1): the
    trajectory against the oracle (round-off grows through the coarsening, so the bar
    after 1000 steps is 1e-9) and mass to 1e-13."""
    N = 128
    kw = dict(Nx=N, Ny=N, lx=2 * math.pi, ly=2 * math.pi, nu=0.01, dt=1e-3, lam=1e-3,
              epsilon=1.5 * 2 * math.pi / N, sigma=0.1, dealias=dealias)
    print(f"binary quench 128^2, 1000 steps, dealias {dealias}", flush=True)
    op = O.Params(model="binary", **kw)
    (c0,) = O.init_noise_real(op, 1, 0.5, 0.1)
    s0 = O.state_from_real(op, None, c0)
    ref = O.run(op, s0, 1000)
    with C.Context(C.params("binary", **kw)) as ctx:
        ctx.set_real_state(None, c0)
        ctx.step(1000)
        g = ctx.get_state()
    for a, b in zip(g, ref.fields()):
        assert rel(a, b) < 1e-9
    assert abs(g[1][0, 0] - s0.cphi1[0, 0]) <= TOL_MASS * abs(s0.cphi1[0, 0])


def test_config_ternary_forced_1024():
    """Ternary 1024^2, c1, c2 = 1/3 + U[-0.05, 0.05] (seed 2), all sigma_ij = 0.1,
    eps = 1.5 dx, lambda = 1e-3, nu = 5e-3, Xlambda = 2, Kolmogorov forcing 0.1 sin(4y):
    10 steps against the oracle (dt = 5e-4 diverges at this size for the explicit
    k^2 muF term; dt = 5e-4 (256/1024)^3 is used)."""
    N = 1024
    kw = dict(Nx=N, Ny=N, lx=2 * math.pi, ly=2 * math.pi, nu=5e-3, dt=5e-4 * (256.0 / N) ** 3, lam=1e-3,
              epsilon=1.5 * 2 * math.pi / N, sigma_12=0.1, sigma_13=0.1, sigma_23=0.1, Xlambda=2.0,
              f_amp=0.1, f_k=4)
    op = O.Params(model="ternary", **kw)
    c1, c2 = O.init_noise_real(op, 2, 1.0 / 3.0, 0.05)
    s0 = O.state_from_real(op, None, c1, c2)
    ref = O.run(op, s0, 10)
    with C.Context(C.params("ternary", **kw)) as ctx:
        ctx.set_real_state(None, c1, c2)
        ctx.step(10)
        g = ctx.get_state()
    for a, b in zip(g, ref.fields()):
        assert rel(a, b) <= TOL_TRAJ


@pytest.mark.parametrize("c0", [0.3, 0.85])
def test_binary_linear_growth_rate(c0):
    """One library step of c = c0 + d cos(3x + 2y), w = 0, grows the mode by
    exp(s dt), s = -lambda k^2 (1.5 sigma eps k^2 + 24 (sigma/eps) P'(c0)),
    P(c) = c(1-c)(1-2c) (linear stability of binary Cahn-Hilliard)."""
    N, L = 32, 2 * math.pi
    kw = dict(Nx=N, Ny=N, lx=L, ly=L, nu=5e-3, dt=1e-3, lam=1e-3, epsilon=1.5 * L / N, sigma=0.1)
    op = O.Params(model="binary", **kw)
    x = (np.arange(N) * op.dx)[:, None] * np.ones((1, N))
    y = (np.arange(N) * op.dy)[None, :] * np.ones((N, 1))
    s0 = O.State(np.zeros((N, N // 2 + 1), np.complex128), O.dfft(c0 + 1e-9 * np.cos(3 * x + 2 * y)))
    with C.Context(C.params("binary", **kw)) as ctx:
        ctx.set_state(*s0.fields())
        ctx.step(1)
        g = ctx.get_state()
    dp = 1.0 - 6.0 * c0 + 6.0 * c0 * c0
    rate = -1e-3 * 13 * (1.5 * 0.1 * op.epsilon * 13 + 24.0 * (0.1 / op.epsilon) * dp)
    got = g[1][3, 2] / s0.cphi1[3, 2]
    assert abs((got.real - 1.0) - math.expm1(rate * 1e-3)) < 1e-6 * abs(math.expm1(rate * 1e-3))
    assert abs(got.imag) < 1e-14


@pytest.mark.parametrize("model,N", [("ternary", 16), ("binary", 64), ("ns", 128), ("ternary", 2048)])
def test_state_round_trip_is_exact(model, N):
    """chns_set_state then chns_get_state return every mode bit for bit (the
    IDXC <-> column-major transpose runs on the device; N/2+1 columns leave
    ragged 32 x 32 tiles at every size)."""
    L = 2 * math.pi
    cp = C.params(model, Nx=N, Ny=N, lx=L, ly=L)
    rng = np.random.default_rng(N)
    nf = {"ns": 1, "binary": 2, "ternary": 3}[model]
    f = [rng.standard_normal((N, N // 2 + 1)) + 1j * rng.standard_normal((N, N // 2 + 1)) for _ in range(nf)]
    with C.Context(cp) as ctx:
        ctx.set_state(*f)
        g = ctx.get_state()
        ctx.set_state(*g[::-1] if nf > 1 else g)   # a second upload reuses the staging buffer
        h = ctx.get_state()
    for a, b in zip(f, g):
        assert np.array_equal(a, b)
    for a, b in zip((g[::-1] if nf > 1 else g), h):
        assert np.array_equal(a, b)


def test_ternary_linear_growth_rate():
    """One library step of a small mode on c1, then on c2, of a uniform ternary
    state at rest maps the pair of amplitudes by exp(A dt), with A from the
    linearised ternary chemical potential (tests/test_oracle_pins.py derives A
    with sympy and pins the oracle the same way): this checks
    L_c = -0.75 lambda eps k^4, the beta/z1c/z2c tables on the device and the
    -(lambda/gamma_i) k^2 M_i term independently of the oracle."""
    import test_oracle_pins as T

    p, c10, c20, kx, ky = T.ternary_growth_case()
    expA, A = T.ternary_linear_map(p, c10, c20, kx, ky)
    kw = dict(Nx=p.Nx, Ny=p.Ny, lx=p.lx, ly=p.ly, nu=p.nu, dt=p.dt, lam=p.lam, epsilon=p.epsilon,
              sigma_12=p.sigma_12, sigma_13=p.sigma_13, sigma_23=p.sigma_23)
    G = np.zeros((2, 2))
    with C.Context(C.params("ternary", **kw)) as ctx:
        for j, s0 in enumerate(T.ternary_growth_states(p, c10, c20, kx, ky)):
            ctx.set_state(*s0.fields())
            ctx.step(1)
            g = ctx.get_state()
            d0 = (s0.cphi1, s0.cphi2)[j][kx, ky]
            for i in range(2):
                r = g[1 + i][kx, ky] / d0
                assert abs(r.imag) < 1e-14
                G[i, j] = r.real
            assert np.max(np.abs(g[0])) < 1e-12 * abs(d0)
    T.check_ternary_growth(G, expA, A, p.dt)



def full_size_case(model, N):
    """Workloads at the benchmark sizes that stay finite over 100 steps:
    NS with forcing; the binary spinodal quench and the forced ternary of the
    benchmark configurations (dt scaled as (256/N)^3, the explicit k^2 muF
    rate growing as N^3 at eps = 1.5 dx); at 2048^2 the liquid lens of the
    benchmark (dt = 1e-6) with a random vorticity."""
    import bench

    L = 2 * math.pi
    if model == "ns":
        kw = dict(Nx=N, Ny=N, lx=L, ly=L, nu=5e-3, dt=2e-4, f_amp=0.1, f_k=4)
        op = O.Params(model="ns", **kw)
        return op, C.params("ns", **kw), O.state_from_real(op, O.random_vorticity(op, 1))
    if model == "binary":
        kw = dict(Nx=N, Ny=N, lx=L, ly=L, nu=0.01, dt=5e-4 * (256 / N) ** 3, lam=1e-3, epsilon=1.5 * L / N, sigma=0.1)
        op = O.Params(model="binary", **kw)
        return op, C.params("binary", **kw), O.state_from_real(op, None, *O.init_noise_real(op, 1, 0.5, 0.1))
    if N < 2048:
        kw = dict(Nx=N, Ny=N, lx=L, ly=L, nu=5e-3, dt=5e-4 * (256 / N) ** 3, lam=1e-3, epsilon=1.5 * L / N,
                  sigma_12=0.1, sigma_13=0.1, sigma_23=0.1, Xlambda=2.0, f_amp=0.1, f_k=4)
        op = O.Params(model="ternary", **kw)
        return op, C.params("ternary", **kw), O.state_from_real(op, None, *O.init_noise_real(op, 2, 1 / 3, 0.05))
    kw = bench.workload(N)
    op = O.Params(model="ternary", **kw)
    return op, C.params("ternary", **kw), O.state_from_real(op, O.random_vorticity(op, 1), *O.init_lens_real(op))


@pytest.mark.parametrize("model,N", [("ns", 1024), ("binary", 1024), ("ternary", 1024), ("ternary", 2048)])
def test_100_steps_full_size(model, N):
    """100 steps at the benchmark sizes: rel-L2 <= 1e-10 of the state and 1e-8
    of its change over the run, mass to 1e-13, against the C/OpenMP
    restatement of the oracle (oracle/chns_step.c, pinned to the numpy oracle
    to round-off in tests/test_c_oracle.py), the spectrum truncation every 25
    steps."""
    from oracle import c_oracle as CO

    op, cp, s0 = full_size_case(model, N)
    CO.threads(16)
    print(f"100 steps {model} {N}^2 (C oracle, 16 threads)", flush=True)
    ref = CO.run(op, s0, 100, spectrum_every=25)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        for _ in range(4):
            ctx.step(25)
            ctx.truncate_state()
        g = ctx.get_state()
        assert ctx.status() == 0
    for a, b, a0 in zip(g, ref.fields(), s0.fields()):
        assert rel(a, b) <= TOL_TRAJ
        assert rel(a - a0, b - a0) <= 1e-8
    for a, b in zip(g[1:], s0.fields()[1:]):
        assert abs(a[0, 0] - b[0, 0]) <= TOL_MASS * abs(b[0, 0])
