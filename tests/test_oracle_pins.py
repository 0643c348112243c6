"""Pins of the oracle restatement, independent of how the oracle is written.

Each check compares an oracle function with something the mathematics fixes:
a direct O(N^4) DFT, integer lattice counts, sympy-derived real-space closed
forms for band-limited fields (which the pseudo-spectral method must reproduce
to round-off because nothing aliases or is truncated), exact rational
arithmetic, Parseval, analytic ODE solutions, and published constants.
"""
import math
from fractions import Fraction

import numpy as np
import pytest
import sympy as sp

import oracle as O

TWO_PI = 2.0 * math.pi


def grid(p):
    x = (np.arange(p.Nx) * p.dx)[:, None] * np.ones((1, p.Ny))
    y = (np.arange(p.Ny) * p.dy)[None, :] * np.ones((p.Nx, 1))
    return x, y


# ---------------------------------------------------------------- transforms
def test_dfft_matches_direct_dft():
    p = O.Params(Nx=8, Ny=8)
    u = np.random.default_rng(0).standard_normal((8, 8))
    i = np.arange(8)
    E = np.exp(-2j * np.pi * np.outer(i, i) / 8)
    full = E @ u @ E.T
    assert np.allclose(O.dfft(u), full[:, :5], atol=1e-12, rtol=0)
    assert np.allclose(O.ifft(p, O.dfft(u)), u, atol=1e-14)


def test_single_mode_and_parseval():
    p = O.Params(Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI)
    x, y = grid(p)
    U = O.dfft(np.cos(3 * x + 5 * y))
    expect = np.zeros_like(U)
    expect[3, 5] = 0.5 * 32 * 32
    assert np.allclose(U, expect, atol=1e-10)
    u = np.random.default_rng(1).standard_normal((32, 32))
    Uh = O.dfft(u)
    w = np.full(17, 2.0)
    w[0] = w[16] = 1.0
    assert math.isclose(np.sum(u * u), np.sum(w * np.abs(Uh) ** 2) / 32 ** 2, rel_tol=1e-13)


# ---------------------------------------------------------------- kxky / truncation
def test_kxky_nyquist_positive():
    p = O.Params(Nx=16, Ny=16, lx=4 * math.pi, ly=4 * math.pi)
    kx, ky = O.kxky(p)
    assert kx[8] == 4.0 and kx[9] == -3.5 and ky[-1] == 4.0


@pytest.mark.parametrize("N,K", [(128, 1637), (2048, 412249)])
def test_truncation_keep_count_is_lattice_count(N, K):
    p = O.Params(Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI)
    R2 = (N // 4) ** 2
    count = 0
    for i in range(-N // 2 + 1, N // 2 + 1):
        jmax = math.isqrt(max(R2 - i * i, -1)) if R2 - i * i >= 0 else -1
        count += max(0, min(jmax, N // 2) + 1)
    assert count == K
    assert int(O.dealias_mask(p).sum()) == K


@pytest.mark.parametrize("N", [16, 64, 2048])
def test_two_thirds_rule_keep_count_is_lattice_count(N):
    """dealias=1: modes with |kx index| <= N/3 and ky index <= N/3 are kept;
    dealias=2: the circle 9 (i^2 + j^2) <= N^2 (integer arithmetic here)."""
    p1 = O.Params(Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI, dealias=1)
    p2 = O.Params(Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI, dealias=2)
    third = N // 3                      # N is not a multiple of 3: floor(N/3) < N/3
    square = (2 * third + 1) * (third + 1)
    circle = 0
    for i in range(-N // 2 + 1, N // 2 + 1):
        for j in range(0, N // 2 + 1):
            circle += 9 * (i * i + j * j) <= N * N
    m1, m2 = O.dealias_mask(p1), O.dealias_mask(p2)
    assert int(m1.sum()) == square
    assert int(m2.sum()) == circle
    # the reference rule keeps a subset of both
    m0 = O.dealias_mask(O.Params(Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI))
    assert not np.any(m0 & ~m1) and not np.any(m0 & ~m2)
    # kx index N/3 + 1 is cut on the ky = 0 row, N/3 kept (x index i, its negative N - i)
    assert m1[third, 0] and m1[N - third, 0] and not m1[third + 1, 0] and not m1[N - third - 1, 0]
    assert m1[0, third] and not m1[0, third + 1]


# ---------------------------------------------------------------- velocity / curl
def test_curl_of_velocity_is_vorticity_and_divergence_free():
    p = O.Params(Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI)
    w = O.truncation(p, O.dfft(O.random_vorticity(p, 3)))
    w[0, 0] = 0.0
    ux, uy = O.get_velocity(p, w)
    assert O.rel_l2(O.curl(p, ux, uy), w) < 1e-15
    assert np.max(np.abs(O.divergence(p, ux, uy))) < 1e-15 * np.max(np.abs(w))


def test_single_mode_velocity():
    p = O.Params(Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI)
    x, y = grid(p)
    ux, uy = O.get_velocity(p, O.dfft(2.0 * np.sin(x) + 3.0 * np.cos(2 * y)))
    # psi = 2 sin x + (3/4) cos 2y ; u = dpsi/dy, v = -dpsi/dx
    assert np.allclose(O.ifft(p, ux), -1.5 * np.sin(2 * y), atol=1e-14)
    assert np.allclose(O.ifft(p, uy), -2.0 * np.cos(x), atol=1e-14)


# ---------------------------------------------------------------- chemical potential
X1, X2 = sp.symbols("c1 c2")


def sym_muF(g1, g2, g3, dl, Xl, eps):
    """muF from the ternary free energy, differentiated by sympy.

    F0 = sum_i (g_i/2) c_i^2 (1-c_i)^2 + 3 Xl (c1 c2 c3)^2 with c1, c2, c3
    independent; G_i = dF0/dc_i; muF1 = [(G1-G2)/g2 + (G1-G3)/g3] 2 del/eps.
    """
    a, b, c = sp.symbols("a b c")
    F0 = (g1 / 2) * a**2 * (1 - a)**2 + (g2 / 2) * b**2 * (1 - b)**2 + (g3 / 2) * c**2 * (1 - c)**2 \
        + 3 * Xl * (a * b * c)**2
    G = [sp.diff(F0, v) for v in (a, b, c)]
    sub = {a: X1, b: X2, c: 1 - X1 - X2}
    G = [g.subs(sub) for g in G]
    m1 = ((G[0] - G[1]) / g2 + (G[0] - G[2]) / g3) * 2 * dl / eps
    m2 = ((G[1] - G[0]) / g1 + (G[1] - G[2]) / g3) * 2 * dl / eps
    return sp.expand(m1), sp.expand(m2)


def test_potential_pointwise_exact_rational():
    p = O.Params(Nx=16, Ny=16, lx=TWO_PI, ly=TWO_PI, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06,
                 epsilon=0.0625)
    g1, g2, g3, dl = p.gammas
    m1s, m2s = sym_muF(*[sp.Rational(Fraction(v)) for v in (g1, g2, g3, dl, p.Xlambda, p.epsilon)])
    rng = np.random.default_rng(5)
    c1 = rng.integers(-256, 1280, size=50) / 1024.0
    c2 = rng.integers(-256, 1280, size=50) / 1024.0
    o1, o2 = O.potential_pointwise(p, c1, c2)
    for k in range(50):
        s = {X1: sp.Rational(Fraction(c1[k])), X2: sp.Rational(Fraction(c2[k]))}
        e1, e2 = float(m1s.subs(s)), float(m2s.subs(s))
        scale = max(1.0, abs(e1), abs(e2))
        assert abs(o1[k] - e1) <= 1e-13 * scale
        assert abs(o2[k] - e2) <= 1e-13 * scale


def test_potential_zeros():
    p = O.Params(Nx=16, Ny=16, sigma_12=0.1, sigma_13=0.1, sigma_23=0.1)
    one3 = np.array([1.0 / 3.0])
    m1, m2 = O.potential_pointwise(p, one3, one3)
    assert abs(m1[0]) < 1e-15 and abs(m2[0]) < 1e-15
    for a, b in ((1.0, 0.0), (0.0, 1.0), (0.0, 0.0)):
        m1, m2 = O.potential_pointwise(p, np.array([a]), np.array([b]))
        assert m1[0] == 0.0 and m2[0] == 0.0


# ---------------------------------------------------------------- surface tension and nonlin (closed forms)
xs, ys = sp.symbols("x y")


def band_limited_case(model):
    """Fields with Fourier modes <= 2 at N = 64: no product aliases, nothing truncated."""
    p = O.Params(model=model, Nx=64, Ny=64, lx=TWO_PI, ly=TWO_PI, sigma=0.1, epsilon=0.2, lam=1e-3,
                 sigma_12=0.1, sigma_13=0.08, sigma_23=0.06)
    # mixed |k| so that grad(mu) is not parallel to grad(c); muF (degree 5) stays
    # inside the dealias disk (|k| <= 10*sqrt(2) < 16) and products resolve on 64^2
    c1 = sp.Rational(1, 3) + sp.Rational(1, 10) * sp.cos(xs) + sp.Rational(1, 20) * sp.sin(2 * ys)
    c2 = sp.Rational(1, 3) - sp.Rational(1, 20) * sp.cos(ys) + sp.Rational(1, 25) * sp.sin(2 * xs)
    if model == "binary":
        c1 = sp.Rational(1, 2) + sp.Rational(1, 10) * sp.cos(xs) + sp.Rational(1, 20) * sp.sin(2 * ys)
    w = sp.Rational(3, 10) * sp.sin(xs + ys) + sp.Rational(1, 5) * sp.cos(2 * ys)
    return p, c1, c2, w


def lap(f):
    return sp.diff(f, xs, 2) + sp.diff(f, ys, 2)


def velocity_of(w):
    # psi with -lap(psi) = w for the modes used: sin(x+y) (k^2 = 2), cos(2y) (k^2 = 4)
    psi = sp.Rational(3, 20) * sp.sin(xs + ys) + sp.Rational(1, 20) * sp.cos(2 * ys)
    assert sp.simplify(-lap(psi) - w) == 0
    return sp.diff(psi, ys), -sp.diff(psi, xs)


def on_grid(expr, p):
    f = sp.lambdify((xs, ys), expr, "numpy")
    x, y = grid(p)
    return np.broadcast_to(f(x, y), x.shape).astype(float)


def analytic_rhs(model, p, c1, c2, w):
    u, v = velocity_of(w)
    eps = sp.Rational(Fraction(p.epsilon))
    nw = -(u * sp.diff(w, xs) + v * sp.diff(w, ys))
    if model == "ternary":
        g1, g2, g3, dl = [sp.Rational(Fraction(q)) for q in p.gammas]
        m1, m2 = sym_muF(g1, g2, g3, dl, sp.Rational(Fraction(p.Xlambda)), eps)
        mf1 = m1.subs({X1: c1, X2: c2})
        mf2 = m2.subs({X1: c1, X2: c2})
        mu1 = mf1 - sp.Rational(3, 4) * g1 * eps * lap(c1)
        mu2 = mf2 - sp.Rational(3, 4) * g2 * eps * lap(c2)
        A1 = (1 + g3 / g1) * mu1 + (g3 / g2) * mu2
        A2 = (g3 / g1) * mu1 + (1 + g3 / g2) * mu2
        Fx = A1 * sp.diff(c1, xs) + A2 * sp.diff(c2, xs)
        Fy = A1 * sp.diff(c1, ys) + A2 * sp.diff(c2, ys)
        st = sp.diff(Fy, xs) - sp.diff(Fx, ys)
        lam = sp.Rational(Fraction(p.lam))
        n1 = -(u * sp.diff(c1, xs) + v * sp.diff(c1, ys)) + (lam / g1) * lap(mf1)
        n2 = -(u * sp.diff(c2, xs) + v * sp.diff(c2, ys)) + (lam / g2) * lap(mf2)
        return st, nw + st, n1, n2
    sig = sp.Rational(Fraction(p.sigma))
    P = c1 * (1 - c1) * (1 - 2 * c1)
    mu = sp.Rational(3, 2) * sig * eps * (-lap(c1)) + 24 * (sig / eps) * P
    st = sp.diff(mu * sp.diff(c1, ys), xs) - sp.diff(mu * sp.diff(c1, xs), ys)
    lam = sp.Rational(Fraction(p.lam))
    n1 = -(u * sp.diff(c1, xs) + v * sp.diff(c1, ys)) + lam * 24 * (sig / eps) * lap(P)
    return st, nw + st, n1, None


@pytest.mark.parametrize("model", ["binary", "ternary"])
def test_surface_tension_and_nonlin_closed_form(model):
    p, c1, c2, w = band_limited_case(model)
    st, nw, n1, n2 = analytic_rhs(model, p, c1, c2, w)
    C1 = O.dfft(on_grid(c1, p))
    C2 = O.dfft(on_grid(c2, p))
    W = O.dfft(on_grid(w, p))
    if model == "ternary":
        M1, M2 = O.potential(p, C1, C2)
        got_st = O.surface_tension(p, C1, C2, M1, M2)
        got = O.nonlin(p, W, C1, C2, M1, M2)
    else:
        got_st = O.surface_tension(p, C1)
        got = O.nonlin(p, W, C1)
    ref_st = on_grid(st, p)
    assert np.linalg.norm(ref_st) > 1e-3
    assert O.rel_l2(O.ifft(p, got_st), ref_st) < 1e-10   # the source is a small difference of O(1) terms
    assert O.rel_l2(O.ifft(p, got[0]), on_grid(nw, p)) < 1e-12
    assert O.rel_l2(O.ifft(p, got[1]), on_grid(n1, p)) < 1e-12
    if n2 is not None:
        assert O.rel_l2(O.ifft(p, got[2]), on_grid(n2, p)) < 1e-12


def test_surface_tension_vanishes_for_one_dimensional_fields():
    p = O.Params(model="ternary", Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI, sigma_12=0.1, sigma_13=0.08,
                 sigma_23=0.06)
    x, y = grid(p)
    C1 = O.truncation(p, O.dfft(0.3 + 0.2 * np.tanh(3 * np.sin(y))))
    C2 = O.truncation(p, O.dfft(0.3 - 0.1 * np.cos(2 * y) + 0.05 * np.sin(5 * y)))
    M1, M2 = O.potential(p, C1, C2)
    assert np.max(np.abs(O.surface_tension(p, C1, C2, M1, M2))) == 0.0


def test_taylor_green_has_no_advection():
    p = O.Params(model="ns", Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI)
    x, y = grid(p)
    N = O.nonlin(p, O.dfft(2.0 * np.sin(x) * np.sin(y)))
    assert np.max(np.abs(N)) < 1e-10


# ---------------------------------------------------------------- ETD2RK
def test_etd_coefficients_match_closed_form_where_well_conditioned():
    p = O.Params(model="ternary", Nx=64, Ny=64, lx=TWO_PI, ly=TWO_PI, nu=0.3, dt=0.01, lam=0.5, epsilon=0.5)
    tb = O.etd_tables(p)
    ks = O.ksqr_grid(p)
    c = -p.nu * ks
    good = np.abs(c * p.dt) > 1e-2
    ex = np.expm1(c[good] * p.dt)
    assert np.allclose(tb["eta"][good], np.exp(c[good] * p.dt), rtol=1e-15, atol=0)
    assert np.allclose(tb["z1w"][good], ex / c[good], rtol=1e-13, atol=0)
    phi2 = (ex - c[good] * p.dt) / (p.dt * c[good] ** 2)
    assert np.allclose(tb["z2w"][good], phi2, rtol=1e-9, atol=0)
    assert tb["z1w"][0, 0] == p.dt and tb["z2w"][0, 0] == 0.5 * p.dt


def test_ternary_etd_coefficients_match_closed_form():
    """Ternary phase-field coefficients: the linear operator of
    dc_i/dt = -(lambda/gamma_i) k^2 (M_i + 0.75 gamma_i eps k^2 c_i) is
    L_c = -0.75 lambda eps k^4 (the gamma_i cancel), so beta = exp(L_c dt),
    z1c = expm1(L_c dt)/L_c and z2c = (expm1(L_c dt) - L_c dt)/(dt L_c^2) wherever
    |L_c dt| is large enough for the reference expressions to keep their digits;
    a wrong power of k or a missing factor changes them at O(1)."""
    p = O.Params(model="ternary", Nx=64, Ny=64, lx=TWO_PI, ly=TWO_PI, nu=0.3, dt=0.05, lam=0.5, epsilon=0.7)
    tb = O.etd_tables(p)
    kx = np.fft.fftfreq(64, 1.0 / 64)[:, None]
    ky = np.arange(33)[None, :]
    k4 = (kx * kx + ky * ky) ** 2
    Lc = -0.75 * p.lam * p.epsilon * k4
    good = np.abs(Lc * p.dt) > 1e-2
    assert good.sum() > 1000
    ex = np.expm1(Lc[good] * p.dt)
    assert np.allclose(tb["beta"][good], np.exp(Lc[good] * p.dt), rtol=1e-14, atol=0)
    assert np.allclose(tb["z1c"][good], ex / Lc[good], rtol=1e-13, atol=0)
    assert np.allclose(tb["z2c"][good], (ex - Lc[good] * p.dt) / (p.dt * Lc[good] ** 2), rtol=1e-9, atol=0)
    assert tb["beta"][0, 0] == 1.0 and tb["z1c"][0, 0] == p.dt and tb["z2c"][0, 0] == 0.5 * p.dt


def ternary_linear_map(p, c10, c20, kx, ky):
    """Exact one-step map exp(A dt) of a small mode d_i cos(kx x + ky y) on a
    uniform ternary state at rest, from the linearised chemical potential.

    dc_i/dt = -(lambda/gamma_i) k^2 mu_i, mu_i = muF_i(c) + 0.75 gamma_i eps k^2 c_i
    (chem_potential.c:74-86, surface_tension.c:116-124, nonlin.c:369-370), so
    A = -0.75 lambda eps k^4 I - lambda k^2 diag(1/gamma_i) J with J the Jacobian
    of (muF_1, muF_2) at (c10, c20), derived here with sympy from the
    chemical potential's definition."""
    import scipy.linalg

    g1, g2, g3, dl = p.gammas
    a, b = sp.symbols("a b")
    c3 = 1 - a - b
    fc = lambda c: c * (1 - c) * (1 - 2 * c)
    X6 = 6 * sp.nsimplify(p.Xlambda)
    G1 = g1 * fc(a) + X6 * a * b * c3 * b * c3
    G2 = g2 * fc(b) + X6 * a * b * c3 * a * c3
    G3 = g3 * fc(c3) + X6 * a * b * c3 * a * b
    fac = 2 * dl / p.epsilon
    mu1 = ((G1 - G2) / g2 + (G1 - G3) / g3) * fac
    mu2 = ((G2 - G1) / g1 + (G2 - G3) / g3) * fac
    J = np.array([[float(sp.diff(m, v).subs({a: c10, b: c20})) for v in (a, b)] for m in (mu1, mu2)])
    k2 = float(kx * kx + ky * ky)
    A = -0.75 * p.lam * p.epsilon * k2 * k2 * np.eye(2) - p.lam * k2 * np.diag([1.0 / g1, 1.0 / g2]) @ J
    return scipy.linalg.expm(A * p.dt), A


def ternary_growth_case():
    N = 32
    p = O.Params(model="ternary", Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI, nu=5e-3, dt=1e-3, lam=1e-3,
                 epsilon=1.5 * TWO_PI / N, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06)
    return p, 0.3, 0.25, 3, 2


def ternary_growth_states(p, c10, c20, kx, ky, d=1e-9):
    """Two initial states: the mode on c1 only, then on c2 only."""
    x, y = grid(p)
    m = d * np.cos(kx * x + ky * y)
    w = np.zeros((p.Nx, p.Ny2), np.complex128)
    return [O.State(w, O.dfft(c10 + m), O.dfft(c20 + 0 * m)), O.State(w, O.dfft(c10 + 0 * m), O.dfft(c20 + m))]


def check_ternary_growth(G, expA, A, dt):
    """One ETD2RK step maps the mode by G (2 x 2); it must equal exp(A dt) to the
    scheme's O((|A| dt)^3) local error.  Compared as G - I against expm(A dt) - I
    (|A| dt ~ 1e-3, where the scheme's error is ~2e-7 of |expm(A dt) - I|, under
    the bar of 1e-6): a dropped z2 correction misses an O(|A| dt / 2) part of
    it, a wrong power of k in L_c or a wrong lambda/gamma_i an O(1) part."""
    D = expA - np.eye(2)
    assert np.max(np.abs(A)) * dt > 1e-4
    assert np.max(np.abs((G - np.eye(2)) - D)) < 1e-6 * np.max(np.abs(D))


def test_ternary_linear_growth_rate():
    p, c10, c20, kx, ky = ternary_growth_case()
    expA, A = ternary_linear_map(p, c10, c20, kx, ky)
    # unequal gammas: the two fields grow differently and couple
    assert abs(A[0, 1]) > 0.1 * abs(A[0, 0]) and abs(A[0, 0] - A[1, 1]) > 0.1 * abs(A[0, 0])
    G = np.zeros((2, 2))
    for j, s0 in enumerate(ternary_growth_states(p, c10, c20, kx, ky)):
        s1 = O.etd2rk(p, s0)
        d0 = (s0.cphi1, s0.cphi2)[j][kx, ky]
        for i, f in enumerate((s1.cphi1, s1.cphi2)):
            r = f[kx, ky] / d0
            assert abs(r.imag) < 1e-14
            G[i, j] = r.real
        # a single mode drives no flow at first order
        assert np.max(np.abs(s1.cwz)) < 1e-12 * abs(d0)
    check_ternary_growth(G, expA, A, p.dt)


def test_ternary_growth_pin_catches_planted_errors():
    """The growth-rate check fails for the mistakes it is meant to catch,
    planted in copies of the oracle's coefficient tables and parameters."""
    p, c10, c20, kx, ky = ternary_growth_case()
    expA, A = ternary_linear_map(p, c10, c20, kx, ky)
    states = ternary_growth_states(p, c10, c20, kx, ky)

    def response(tables=None, q=None):
        q = q or p
        G = np.zeros((2, 2))
        for j, s0 in enumerate(states):
            s1 = O.etd2rk(q, s0, tables)
            d0 = (s0.cphi1, s0.cphi2)[j][kx, ky]
            G[:, j] = [(s1.cphi1[kx, ky] / d0).real, (s1.cphi2[kx, ky] / d0).real]
        return G

    tb = O.etd_tables(p)
    ks = O.ksqr_grid(p)
    bad = []
    t = dict(tb)
    t["z2c"] = np.zeros_like(tb["z2c"])                       # dropped z2 term
    bad.append(response(t))
    t = dict(tb)
    Lc = -0.75 * p.lam * p.epsilon * ks ** 3                  # k^6 instead of k^4
    t["beta"] = np.exp(Lc * p.dt)
    t["z1c"] = np.where(Lc != 0, np.expm1(Lc * p.dt) / np.where(Lc != 0, Lc, 1), p.dt)
    bad.append(response(t))
    bad.append(response(q=O.Params(**{**p.__dict__, "sigma_13": 0.06, "sigma_23": 0.08})))   # gammas swapped
    for G in bad:
        with pytest.raises(AssertionError):
            check_ternary_growth(G, expA, A, p.dt)


def test_stokes_decay_single_mode():
    p = O.Params(model="ns", Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI, nu=0.05, dt=0.01)
    x, y = grid(p)
    s = O.State(O.dfft(np.cos(3 * x + 4 * y)))
    s = O.run(p, s, 50)
    expect = O.dfft(np.cos(3 * x + 4 * y)) * math.exp(-p.nu * 25 * 50 * p.dt)
    assert O.rel_l2(s.cwz, expect) < 1e-13


def test_forced_kolmogorov_steady_state():
    p = O.Params(model="ns", Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI, nu=0.05, dt=0.5, f_amp=0.1, f_k=4)
    s = O.run(p, O.State(np.zeros((32, 17), np.complex128)), 400)
    x, y = grid(p)
    amp = 0.1 / (0.05 * 16) * (1 - math.exp(-0.05 * 16 * 200.0))
    assert O.rel_l2(O.ifft(p, s.cwz), amp * np.sin(4 * y)) < 1e-12


@pytest.mark.parametrize("c0", [0.3, 0.5, 0.85])
def test_binary_linear_growth_rate(c0):
    """Linear stability of binary Cahn-Hilliard at rest: a small mode
    c = c0 + d cos(k.x) evolves as exp(s dt) with
    s = -lambda k^2 (1.5 sigma eps k^2 + 24 (sigma/eps) P'(c0)), P(c) = c(1-c)(1-2c).
    The quadratic term of P feeds only modes 0 and 2k and the cubic one is
    O(d^2) relative, so one ETD2RK step of the oracle must give that factor to
    the scheme's O((s dt)^3) error; a wrong weight, sign or power of k in the
    interface or bulk term changes s at O(1)."""
    p = O.Params(model="binary", Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI, nu=5e-3, dt=1e-3, lam=1e-3,
                 epsilon=1.5 * TWO_PI / 32, sigma=0.1)
    x, y = grid(p)
    kx, ky = 3, 2
    d = 1e-9
    s0 = O.State(np.zeros((32, 17), np.complex128), O.dfft(c0 + d * np.cos(kx * x + ky * y)))
    s1 = O.etd2rk(p, s0)
    k2 = kx * kx + ky * ky
    dp = 1.0 - 6.0 * c0 + 6.0 * c0 * c0
    rate = -p.lam * k2 * (1.5 * p.sigma * p.epsilon * k2 + 24.0 * (p.sigma / p.epsilon) * dp)
    got = s1.cphi1[kx, ky] / s0.cphi1[kx, ky]
    assert abs(got.imag) < 1e-15
    # (got - 1) is about rate*dt ~ 1e-5..1e-4; compare it to 1e-6 of itself
    assert abs((got.real - 1.0) - math.expm1(rate * p.dt)) < 1e-6 * abs(math.expm1(rate * p.dt))
    # the mean stays; a single mode drives no flow (grad mu x grad c = 0), up to round-off
    assert s1.cphi1[0, 0] == s0.cphi1[0, 0]
    assert np.max(np.abs(s1.cwz)) < 1e-12 * abs(s0.cphi1[kx, ky])


@pytest.mark.parametrize("model", ["binary", "ternary"])
def test_mass_conservation(model):
    p = O.Params(model=model, Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI, nu=5e-3, dt=5e-4, lam=1e-3,
                 epsilon=0.3, sigma=0.1, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06)
    if model == "ternary":
        c = O.init_noise_real(p, 4, 1 / 3, 0.05)
    else:
        c = O.init_noise_real(p, 4, 0.5, 0.1)
    s0 = O.state_from_real(p, O.random_vorticity(p, 2), *c)
    s = O.run(p, s0, 50)
    for a, b in zip(s.fields()[1:], s0.fields()[1:]):
        assert abs(a[0, 0] - b[0, 0]) <= 1e-13 * abs(b[0, 0])


# ---------------------------------------------------------------- diagnostics / ICs
def test_spectrum_sums_to_kinetic_energy():
    p = O.Params(model="ternary", Nx=64, Ny=64, lx=TWO_PI, ly=TWO_PI)
    c = O.init_noise_real(p, 4, 1 / 3, 0.05)
    s = O.state_from_real(p, O.random_vorticity(p, 2), *c)
    spec, _ = O.spectrum(p, s)
    eu = O.kinetic_energy(p, s)[0]
    assert math.isclose(float(np.sum(spec["ek"])), eu, rel_tol=1e-12)
    mean_c1_sq = float(np.mean(O.ifft(p, O.truncation(p, s.cphi1)) ** 2))
    assert math.isclose(float(np.sum(spec["sph1"])), mean_c1_sq, rel_tol=1e-12)


def test_glibc_rand_known_values():
    # glibc TYPE_3 generator: srand(1) -> 1804289383, 846930886 (widely published)
    v = O.glibc_rand_stream(1, 2) * 2147483647.0
    assert v[0] == 1804289383.0 and v[1] == 846930886.0


def test_splitmix64_known_value():
    # SplitMix64 with seed 0: first output 0xE220A8397B1DCDAF
    u = O.splitmix64_uniform(0, 1)[0]
    assert u == (0xE220A8397B1DCDAF >> 11) / 2.0 ** 53


def test_kinetic_energy_single_mode_closed_form():
    # omega = A cos(3x): v = (A/3) sin(3x), u = 0, so <|u|^2/2> = A^2/36 everywhere;
    # the phase masks depend on y only, so each per-phase mean equals the global one
    p = O.Params(model="ternary", Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI)
    x, y = grid(p)
    A = 2.0
    s = O.state_from_real(p, A * np.cos(3 * x), 0.5 + 0.4 * np.cos(y - 0.1), 0.5 + 0.3 * np.sin(2 * y + 0.2))
    eu, e1, e2 = O.kinetic_energy(p, s)
    for v in (eu, e1, e2):
        assert math.isclose(v, A * A / 36.0, rel_tol=1e-13)


def _binary_mu(p, c):
    """Variational derivative of the binary free energy density:
    24 (sigma/eps) c(1-c)(1-2c) - 1.5 sigma eps lap(c), lap by exact spectral derivative."""
    k2 = O.ksqr_grid(p)
    lap = np.fft.irfft2(-k2 * np.fft.rfft2(c), s=c.shape)
    return 24.0 * (p.sigma / p.epsilon) * c * (1 - c) * (1 - 2 * c) - 1.5 * p.sigma * p.epsilon * lap


def test_binary_free_energy_variation_is_chemical_potential():
    # d/dh F[c + h dc] = <mu, dc>: ties free_energy.c to the chemical potential of
    # surface_tension.c:20-46 (a dropped or mis-scaled term in either breaks it)
    p = O.Params(model="binary", Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI, sigma=0.3, epsilon=0.4)
    x, y = grid(p)
    c = 0.5 + 0.3 * np.cos(x) * np.sin(2 * y) + 0.1 * np.sin(3 * x)
    dc = 0.2 * np.cos(x) * np.sin(2 * y) + 0.1 * np.sin(3 * x) + 0.05 * np.cos(2 * x + y)

    def F(h):
        s = O.State(np.zeros((32, 17), complex), O.dfft(c + h * dc))
        return sum(O.free_energy(p, s))

    h = 1e-4
    dF = (8 * (F(h) - F(-h)) - (F(2 * h) - F(-2 * h))) / (12 * h)   # 4th-order central difference
    assert math.isclose(dF, float(np.mean(_binary_mu(p, c) * dc)), rel_tol=1e-9)


def test_ternary_free_energy_reduces_to_binary_on_a_two_phase_state():
    # with c2 = 0 the ternary energy is the binary one with sigma = sigma_13 (gamma1+gamma3
    # = 2 sigma_13); with c1 = 0 it is the binary one with sigma = sigma_23
    N, L = 32, TWO_PI
    pt = O.Params(model="ternary", Nx=N, Ny=N, lx=L, ly=L, sigma_12=0.3, sigma_13=0.2, sigma_23=0.45, epsilon=0.3)
    x, y = grid(pt)
    c = 0.5 + 0.3 * np.cos(x) * np.sin(2 * y) + 0.1 * np.sin(3 * x)
    z = np.zeros_like(c)
    cw = np.zeros((N, N // 2 + 1), complex)
    for sig, c1, c2 in ((pt.sigma_13, c, z), (pt.sigma_23, z, c)):
        pb = O.Params(model="binary", Nx=N, Ny=N, lx=L, ly=L, sigma=sig, epsilon=0.3)
        et = O.free_energy(pt, O.State(cw, O.dfft(c1), O.dfft(c2)))
        eb = O.free_energy(pb, O.State(cw, O.dfft(c)))
        assert math.isclose(et[0], eb[0], rel_tol=1e-12) and math.isclose(et[1], eb[1], rel_tol=1e-12)


def test_binary_free_energy_closed_form():
    # c = a + b cos(2y): <c^2 (1-c)^2> from an exact integral, <|grad c|^2> = b^2 m^2 / 2
    p = O.Params(model="binary", Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI, sigma=0.1)
    a, b, m = sp.Rational(1, 2), sp.Rational(3, 10), 2
    t = sp.symbols("t")
    cc = a + b * sp.cos(m * t)
    bulk = float(sp.integrate(sp.expand(cc ** 2 * (1 - cc) ** 2), (t, 0, 2 * sp.pi)) / (2 * sp.pi))
    x, y = grid(p)
    s = O.State(np.zeros((32, 17), complex), O.dfft(0.5 + 0.3 * np.cos(2 * y)))
    ephi, egphi = O.free_energy(p, s)
    assert math.isclose(ephi, 12 * (p.sigma / p.epsilon) * bulk, rel_tol=1e-13)
    assert math.isclose(egphi, 0.75 * p.sigma * p.epsilon * 0.09 * 4 / 2, rel_tol=1e-13)


def test_perimeter_of_a_band_is_twice_the_box_length():
    # c >= 0.5 on a band in y: two straight interfaces of length L each, so the
    # printed Sp_t = count * dx equals 2 L and half the points lie inside
    for N in (32, 64):
        p = O.Params(model="binary", Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI)
        x, y = grid(p)
        s = O.State(np.zeros((N, N // 2 + 1), complex), O.dfft(0.5 + 0.4 * np.cos(y - 0.1)))
        ((edge, inside),) = O.perimeter(p, s)
        assert math.isclose(edge * p.dx, 2 * TWO_PI, rel_tol=1e-14)
        assert inside == N * N // 2


def test_phase_spectrum_of_a_single_mode():
    # c = a + b cos(3y): sph holds a^2 at shell 0 and b^2/2 (the mean of (b cos)^2) at shell 3
    p = O.Params(model="binary", Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI)
    x, y = grid(p)
    s = O.State(np.zeros((32, 17), complex), O.dfft(0.5 + 0.3 * np.cos(3 * y)))
    spec, _ = O.spectrum(p, s)
    expect = np.zeros(p.nshell)
    expect[0], expect[3] = 0.25, 0.045
    np.testing.assert_allclose(spec["sph"], expect, rtol=1e-13, atol=1e-18)


def test_binary_droplet_half_level_is_the_circle():
    p = O.Params(model="binary", Nx=64, Ny=64)
    c = O.init_binary_droplet_real(p)
    x, y = grid(p)
    r = np.sqrt((x - p.lx / 2) ** 2 + (y - p.ly / 2) ** 2)
    assert np.all((c > 0.5) == (r < math.pi / 4.5))
    assert c.max() <= 1.0 and c.min() >= 0.0
