"""Oracle: a clean-room numpy FP64 restatement of the reference CHNS-2D step.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP product
path.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it; the product library never does.

Every function restates one reference function (C source read as text) and
cites the file:line it follows.  Array conventions follow the reference:

* complex half-spectra are ``[Nx][Ny/2+1]`` (``IDXC(i,j) = j + Ny2*i``,
  global.h:13), x (index i) slow, y (index j) fast;
* real fields are ``[Nx][Ny]`` (the reference pads rows to ``Ny+2`` for the
  in-place transform, global.h:14; the pad is never read by the physics);
* ``dfft`` is the unnormalised r2c, ``ifft`` the c2r scaled by ``1/(Nx*Ny)``
  (fft.c:11-31).

Scalar expressions whose rounding matters (the ETD2RK coefficient tables,
which cancel catastrophically at low k) are evaluated in exactly the
reference's operation order with IEEE double arithmetic and the C library's
``exp`` (``math.exp`` calls libm's ``exp``).  Everything else is reproduced to
round-off (numpy/pocketfft FFTs instead of FFTW).

Pinning (what checks each function independently of this restatement) is
listed in ``PINS`` at the bottom of the file; ``tests/test_oracle_pins.py``
implements most of them.  Frozen outputs at N = 16 and 32 are committed under
``tests/golden/`` (written by ``tests/golden/make_golden.py``).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field, replace

import numpy as np

PI = math.pi  # global.h:18 ``#define pi (M_PI)``

MODEL_NS, MODEL_BINARY, MODEL_TERNARY = 0, 1, 2
_MODEL_NAMES = {"ns": MODEL_NS, "binary": MODEL_BINARY, "ternary": MODEL_TERNARY}


# ---------------------------------------------------------------------------
# Parameters (global.h:13-134, macro.h:1-15, main.c:52-63,166-170)
# ---------------------------------------------------------------------------
@dataclass
class Params:
    """Runtime form of the reference's compile-time parameters.

    Defaults are the reference defaults (global.h:20-33,127-134, macro.h:8):
    TERNARY, N = 1024, L = 4*pi, nu = 0.01, dt = 1e-4, epsilon = 3*dx,
    lambda = 1e-4, Xlambda = 2, sigma_12/23/13 = 0.2/0.45/0.2.
    ``lam`` is the reference's ``lambda`` (a Python keyword).
    """

    Nx: int = 1024
    Ny: int = 1024
    lx: float = 4.0 * PI
    ly: float = 4.0 * PI
    nu: float = 0.01
    dt: float = 1e-4
    model: int = MODEL_TERNARY
    sigma: float = 2.0            # binary (global.h:102)
    epsilon: float | None = None  # None -> 3.0*dx (global.h:103,131)
    lam: float | None = None      # None -> 1e-6 binary / 1e-4 ternary
    sigma_12: float = 0.2
    sigma_23: float = 0.45
    sigma_13: float = 0.2
    Xlambda: float = 2.0
    smooth_filter: bool = False   # macro.h:15 SMOOTH_FILTER
    dealias: int = 0              # truncation rule: 0 the reference's N/4 circle, 1 the 2/3 rule, 2 an N/3 circle
    f_amp: float = 0.0            # Kolmogorov forcing amplitude on omega
    f_k: int = 0                  # forcing wavenumber index along y
    gravity: bool = False         # macro.h GRAVITY (nonlin.c:107-171)
    rho1: float | None = None     # None -> global.h:108-110 / 137-140 by model
    rho2: float | None = None
    rho3: float = 1.0
    g: float = 3.0
    vpm: bool = False             # macro.h VPM: penalised walls chi(x) (init.c:184-206)
    sponge: bool = False          # macro.h SPONGE: sponge(y) on omega (init.c:207-211), needs VPM
    chi_override: object = None   # test hook: replace chi(x) [Nx] (None: the init.c profile)
    sponge_override: object = None  # test hook: replace sponge(y) [Ny] (None: the init.c profile)

    def __post_init__(self):
        if isinstance(self.model, str):
            self.model = _MODEL_NAMES[self.model]
        if self.epsilon is None:
            self.epsilon = 3.0 * self.dx
        if self.lam is None:
            self.lam = 1e-6 if self.model == MODEL_BINARY else 1e-4
        if self.rho1 is None:
            self.rho1 = 1.044 if self.model == MODEL_BINARY else 0.957   # global.h:108,137
        if self.rho2 is None:
            self.rho2 = 1.0 if self.model == MODEL_BINARY else 1.044     # global.h:109,138

    # derived quantities --------------------------------------------------
    @property
    def Ny2(self):
        return self.Ny // 2 + 1          # global.h:25

    @property
    def dx(self):
        return self.lx / self.Nx         # global.h:27

    @property
    def dy(self):
        return self.ly / self.Ny         # global.h:28

    @property
    def factor_x(self):
        return (2 * PI) / self.lx        # main.c:54

    @property
    def factor_y(self):
        return (2 * PI) / self.ly        # main.c:55

    @property
    def kxmax(self):
        return (self.Nx / 2.0) * self.factor_x   # main.c:57

    @property
    def kymax(self):
        return (self.Ny / 2.0) * self.factor_y   # main.c:58

    @property
    def gammas(self):
        """gamma1..3 and del exactly as main.c:166-170."""
        s12, s23, s13 = self.sigma_12, self.sigma_23, self.sigma_13
        g1 = (s12 - s23 + s13)
        g2 = (s12 - s13 + s23)
        g3 = (s23 - s12 + s13)
        d = (6.0 * g1 * g2 * g3)
        d /= (g1 * g2 + g1 * g3 + g2 * g3)
        return g1, g2, g3, d

    @property
    def nshell(self):
        return int(round((float(self.Nx) / math.sqrt(2.0)) + 1.0))  # main.c:60


# ---------------------------------------------------------------------------
# L2/L3: transforms and spectral utilities
# ---------------------------------------------------------------------------
def kxky(p: Params):
    """kx[Nx], ky[Ny2] as functions.c:11-32 (Nyquist x-mode positive)."""
    i1 = np.arange(p.Nx)
    kx = np.where(i1 <= p.Nx // 2, p.factor_x * i1, p.factor_x * (i1 - p.Nx)).astype(np.float64)
    ky = (p.factor_y * np.arange(p.Ny2)).astype(np.float64)
    return kx, ky


def ksqr_grid(p: Params):
    """ksqr = kx[i]*kx[i] + ky[j]*ky[j] (e.g. functions.c:82, etd2rk.c:88)."""
    kx, ky = kxky(p)
    return (kx * kx)[:, None] + (ky * ky)[None, :]


def dealias_mask(p: Params):
    """Keep-mask of truncation() (functions.c:47-83): zero where ksqr > dax*dax.

    ``p.dealias`` selects the rule.  0 is the reference's circle of radius
    Nx/4 (functions.c:56-59).  1 is the 2/3 rule: zero where
    |kx| > (Nx/3) factor_x or |ky| > (Ny/3) factor_y.  2 is the reference
    predicate with radius Nx/3 (the Nx/3 of energy.c:13-15) instead of Nx/4.
    """
    if p.dealias == 1:
        kx, ky = kxky(p)
        dax = (p.Nx / 3.0) * p.factor_x
        day = (p.Ny / 3.0) * p.factor_y
        return ~(np.abs(kx)[:, None] > dax) & ~(ky[None, :] > day)
    dax = (p.Nx / (3.0 if p.dealias == 2 else 4.0)) * p.factor_x   # functions.c:56
    dealias = dax * dax                                # functions.c:59
    return ~(ksqr_grid(p) > dealias)                   # functions.c:75-78


_FILTER_CACHE = {}


def smooth_filter_factor(p: Params):
    """SMOOTH_FILTER multiplier exp(-36*pp^18), applied twice (functions.c:70-73).

    Evaluated with the C library's exp and pow (math.exp, math.pow) per mode,
    cached per grid (read-only array).
    """
    key = (p.Nx, p.Ny, p.lx, p.ly)
    f = _FILTER_CACHE.get(key)
    if f is None:
        ksqrmax = float(p.kxmax * p.kxmax + p.kymax * p.kymax)   # functions.c:54
        pp = ksqr_grid(p) / ksqrmax
        f = np.array([math.exp(-36.0 * math.pow(v, 18)) for v in pp.ravel()]).reshape(pp.shape)
        f.setflags(write=False)
        _FILTER_CACHE[key] = f
    return f


def truncation(p: Params, cv):
    """Return truncation(cv) (functions.c:47-83) as a new array."""
    if p.smooth_filter:
        f = smooth_filter_factor(p)
        return (cv * f) * f
    return np.where(dealias_mask(p), cv, 0.0 + 0.0j)


def dfft(u):
    """Unnormalised 2-D r2c [Nx][Ny] -> [Nx][Ny/2+1] (fft.c:11-14)."""
    return np.fft.rfft2(np.asarray(u, dtype=np.float64))


def ifft(p: Params, cu):
    """2-D c2r scaled by 1/(Nx*Ny) (fft.c:16-31).

    numpy's inverse already applies the factor 1/(Nx*Ny); for power-of-two
    sizes that factor is a power of two and the scaling is exact either way.
    Imaginary parts at ky = 0 and ky = Ny/2 are ignored, as a c2r does.
    """
    return np.fft.irfft2(np.asarray(cu, dtype=np.complex128), s=(p.Nx, p.Ny))


def _i_times(k, z):
    """(I*k)*z with C99 complex semantics: real -k*Im z, imag k*Re z."""
    return (-k * z.imag) + 1j * (k * z.real)


def _cscale(z, r):
    """complex * real, componentwise (C99 ``z * r``)."""
    return (z.real * r) + 1j * (z.imag * r)


def get_velocity(p: Params, cwz):
    """psi = w/k^2 (0 where k^2 <= 1e-12); u = i ky psi, v = -i kx psi (velocity.c:3-44)."""
    kx, ky = kxky(p)
    ks = ksqr_grid(p)
    safe = np.where(ks > 1e-12, ks, 1.0)
    tcwz = np.where(ks > 1e-12, (cwz.real / safe) + 1j * (cwz.imag / safe), 0.0 + 0.0j)  # velocity.c:20-25
    cux = _i_times(ky[None, :], tcwz)            # velocity.c:38
    cuy = _i_times(-kx[:, None], tcwz)           # velocity.c:39  (-I*kx)*psi
    return cux, cuy


def curl(p: Params, cux, cuy):
    """w = i (kx uy - ky ux) (curl.c:16-31)."""
    kx, ky = kxky(p)
    return 1j * (kx[:, None] * cuy - ky[None, :] * cux)


def divergence(p: Params, cux, cuy):
    """div = i (kx ux + ky uy) (functions.c:87-106)."""
    kx, ky = kxky(p)
    return 1j * (kx[:, None] * cux + ky[None, :] * cuy)


# ---------------------------------------------------------------------------
# L4: physics right-hand side
# ---------------------------------------------------------------------------
def potential(p: Params, cphi1, cphi2):
    """Ternary nonlinear chemical potentials muF1, muF2 and their r2c (chem_potential.c:12-94).

    Returns (cmuF1, cmuF2); the r2c outputs are *not* truncated (:90-91).
    """
    g1, g2, g3, dl = p.gammas
    g1i, g2i, g3i = 1.0 / g1, 1.0 / g2, 1.0 / g3                   # :21-23
    c1 = ifft(p, truncation(p, cphi1))                              # :27-32
    c2 = ifft(p, truncation(p, cphi2))
    muF1, muF2 = potential_pointwise(p, c1, c2)
    return dfft(muF1), dfft(muF2)                                   # :90-91


def potential_pointwise(p: Params, c1, c2):
    """Pointwise muF1, muF2 from real c1, c2 (chem_potential.c:40-41,74-86)."""
    g1, g2, g3, dl = p.gammas
    g1i, g2i, g3i = 1.0 / g1, 1.0 / g2, 1.0 / g3
    c3 = (1.0 - c1 - c2)
    c1c2c3 = c1 * c2 * c3
    fc1 = c1 * (1.0 - c1) * (1.0 - 2.0 * c1)
    fc2 = c2 * (1.0 - c2) * (1.0 - 2.0 * c2)
    fc3 = c3 * (1.0 - c3) * (1.0 - 2.0 * c3)
    X6 = 6.0 * p.Xlambda
    G1 = g1 * fc1 + X6 * c1c2c3 * c2 * c3
    G2 = g2 * fc2 + X6 * c1c2c3 * c1 * c3
    G3 = g3 * fc3 + X6 * c1c2c3 * c1 * c2
    muF1 = (G1 - G2) * g2i + (G1 - G3) * g3i
    muF2 = (G2 - G1) * g1i + (G2 - G3) * g3i
    fac = 2.0 * dl / p.epsilon
    return muF1 * fac, muF2 * fac


def binary_poly(c):
    """P(c) = c(1-c)(1-2c) (surface_tension.c:126-127, nonlin.c:225)."""
    return c * (1.0 - c) * (1.0 - 2.0 * c)


def surface_tension(p: Params, cphi1, cphi2=None, cmuF1=None, cmuF2=None):
    """Surface-tension vorticity source added to cnlz (surface_tension.c:13-227).

    Binary: curl(F(mu grad c)), mu = T(1.5 sigma eps k^2 c + 24 sigma/eps F(P(Tc))) (:19-112).
    Ternary: curl(F(A1 grad c1)) + curl(F(A2 grad c2)) (:114-225).
    Returns the contribution (to be added to cnlz).
    """
    kx, ky = kxky(p)
    ks = ksqr_grid(p)
    eps = p.epsilon
    if p.model == MODEL_BINARY:
        cphi = cphi1
        nphi = ifft(p, truncation(p, cphi))                              # :20-22
        ncphi = dfft(binary_poly(nphi))                                  # :24-35
        ncphi = _cscale(cphi, 1.5 * p.sigma * eps * ks) + _cscale(ncphi, 24.0 * (p.sigma / eps))  # :41-42
        tcux = _i_times(kx[:, None], cphi)                               # :53
        tcuy = _i_times(ky[None, :], cphi)                               # :54
        tux = ifft(p, truncation(p, tcux))                               # :66-71
        tuy = ifft(p, truncation(p, tcuy))
        nphi = ifft(p, truncation(p, ncphi))
        wall = (1.0 - vpm_fields(p)[0])[:, None] if p.vpm else 1.0
        tux = tux * (nphi * wall)                                         # :83-89
        tuy = tuy * (nphi * wall)
        return curl(p, dfft(tux), dfft(tuy))                             # :94-97
    g1, g2, g3, _ = p.gammas
    tcphi1 = cmuF1 + _cscale(cphi1, 0.75 * g1 * eps * ks)                # :123
    tcphi2 = cmuF2 + _cscale(cphi2, 0.75 * g2 * eps * ks)                # :124
    tcux = _i_times(kx[:, None], cphi1)                                  # :126
    tcuy = _i_times(ky[None, :], cphi1)                                  # :127
    tux = ifft(p, truncation(p, tcux))                                   # :132-136
    tuy = ifft(p, truncation(p, tcuy))
    tphi1 = ifft(p, truncation(p, tcphi1))
    tphi2 = ifft(p, truncation(p, tcphi2))
    A1 = ((1.0 + (g3 / g1)) * tphi1 + (g3 / g2) * tphi2)                 # :144-145
    wall = (1.0 - vpm_fields(p)[0])[:, None] if p.vpm else 1.0
    out = curl(p, dfft(tux * A1 * wall), dfft(tuy * A1 * wall))          # :151-166
    tcux = _i_times(kx[:, None], cphi2)                                  # :178
    tcuy = _i_times(ky[None, :], cphi2)                                  # :179
    tux = ifft(p, truncation(p, tcux))                                   # :184-187
    tuy = ifft(p, truncation(p, tcuy))
    A2 = ((g3 / g1) * tphi1 + (1.0 + (g3 / g2)) * tphi2)                 # :195-196
    out = out + curl(p, dfft(tux * A2 * wall), dfft(tuy * A2 * wall))    # :202-216
    return out


def forcing_spectrum(p: Params):
    """Kolmogorov forcing f = f_amp*sin(f_k*ky1*y) on omega as an exact half-spectrum.

    The reference forcing (forcing.c:28-160, hook etd2rk.c:214-216) does not
    compile or is dead code; this additive term in N_omega is new.  Its r2c is
    exactly -i*f_amp*Nx*Ny/2 at (kx=0, ky index f_k) and zero elsewhere.
    """
    F = np.zeros((p.Nx, p.Ny2), dtype=np.complex128)
    if p.f_amp != 0.0:
        if not (0 < p.f_k < p.Ny // 2):
            raise ValueError("f_k must lie in (0, Ny/2)")
        F[0, p.f_k] = -0.5j * p.f_amp * p.Nx * p.Ny
    return F


def vpm_fields(p: Params):
    """chi(x) [Nx] and sponge(y) [Ny] of init.c:184-211 (both x- or y-only), and 1/eta.

    eta = 1.5 dt (global.h:48); eps = 3 dx, walls at xd = 6 dx, xu = lx - xd
    (likewise in y with dy); sponge includes the factor 1/eta (init.c:210).
    """
    etainv = 1.0 / (1.5 * p.dt)
    eps = 3.0 * p.dx
    xd, yd = 6.0 * p.dx, 6.0 * p.dy
    xu, yu = p.lx - xd, p.ly - yd
    x = np.arange(p.Nx) * p.dx
    y = np.arange(p.Ny) * p.dy
    chi = 0.5 * (1.0 - np.tanh(2.0 * (x - xd) / eps)) + 0.5 * (1.0 + np.tanh(2.0 * (x - xu) / eps))
    sp = (0.5 * (1.0 - np.tanh(2.0 * (y - yd) / eps)) + 0.5 * (1.0 + np.tanh(2.0 * (y - yu) / eps))) * etainv
    if p.chi_override is not None:
        chi = np.asarray(p.chi_override, dtype=np.float64)
    if p.sponge_override is not None:
        sp = np.asarray(p.sponge_override, dtype=np.float64)
    return chi, sp, etainv


def gravity_term(p: Params, cphi1, cphi2=None):
    """Buoyancy torque added to cnlz (nonlin.c:107-171): i kx F(-g (rho(c) - 1)) with
    rho = rho2 + (rho1 - rho2) c (binary) or rho3 + (rho1 - rho3) c1 + (rho2 - rho3) c2
    (ternary), c from the untruncated state."""
    kx, _ = kxky(p)
    if p.model == MODEL_BINARY:
        t = ifft(p, cphi1)
        t = (p.rho2 + (p.rho1 - p.rho2) * t) - 1.0
    else:
        c1 = ifft(p, cphi1)
        c2 = ifft(p, cphi2)
        t = (p.rho3 + (p.rho1 - p.rho3) * c1 + (p.rho2 - p.rho3) * c2) - 1.0
    t = t * ((-1.0) * p.g)
    return _i_times(kx[:, None], dfft(t))


def nonlin(p: Params, cwz, cphi1=None, cphi2=None, cmuF1=None, cmuF2=None):
    """All nonlinear terms (nonlin.c:13-381).

    Returns cnlz (NS) / (cnlz, cnlphi) (binary) / (cnlz, cnlphi1, cnlphi2) (ternary).
    """
    kx, ky = kxky(p)
    ks = ksqr_grid(p)
    tcux = _i_times(kx[:, None], cwz)                                    # :36
    tcuy = _i_times(ky[None, :], cwz)                                    # :37
    cux, cuy = get_velocity(p, cwz)                                      # :42
    ux = ifft(p, truncation(p, cux))                                     # :44-48
    uy = ifft(p, truncation(p, cuy))
    tux = ifft(p, truncation(p, tcux))
    tuy = ifft(p, truncation(p, tcuy))
    nlz = ux * tux + uy * tuy                                            # :62
    cnlz = -dfft(nlz)                                                    # :74,92
    if p.vpm:
        chi, sp, etainv = vpm_fields(p)
        wall = chi[:, None] * etainv
        cnlz = cnlz - curl(p, dfft(ux * wall), dfft(uy * wall))          # :63-66,76-80,93-95
        if p.sponge:
            twz = ifft(p, truncation(p, cwz)) * sp[None, :]              # :50-54,68-70,82-84
            cnlz = cnlz - dfft(twz)                                      # :97-99
    if p.model == MODEL_NS:
        return cnlz + forcing_spectrum(p)
    if p.model == MODEL_BINARY:
        cphi = cphi1
        cnlz = cnlz + surface_tension(p, cphi)                           # :104
        if p.gravity:
            cnlz = cnlz + gravity_term(p, cphi)                          # :108-131
        cnlz = cnlz + forcing_spectrum(p)
        cux, cuy = get_velocity(p, cwz)                                  # :177-179
        tcphi = _i_times(kx[:, None], cphi)                              # :187
        ncphi = _i_times(ky[None, :], cphi)                              # :188
        tphi = ifft(p, truncation(p, tcphi))                             # :194-199
        nphi = ifft(p, truncation(p, ncphi))
        tux = ifft(p, truncation(p, cux))
        tuy = ifft(p, truncation(p, cuy))
        nlphi = tux * tphi + tuy * nphi                                  # :207
        if p.vpm:
            nlphi = nlphi * (1.0 - vpm_fields(p)[0])[:, None]            # :209-211
        cnlphi = dfft(nlphi)                                             # :215
        tphi = binary_poly(ifft(p, truncation(p, cphi)))                 # :217-228
        tcphi = dfft(tphi)                                               # :230
        tcphi = _cscale(tcphi, 24.0 * (p.sigma / p.epsilon) * ks)       # :240
        cnlphi = -cnlphi - _cscale(tcphi, p.lam)                         # :251
        return cnlz, cnlphi
    g1, g2, g3, _ = p.gammas
    cnlz = cnlz + surface_tension(p, cphi1, cphi2, cmuF1, cmuF2)         # :104
    if p.gravity:
        cnlz = cnlz + gravity_term(p, cphi1, cphi2)                      # :132-170
    cnlz = cnlz + forcing_spectrum(p)
    cux, cuy = get_velocity(p, cwz)                                      # :258-260
    tux = ifft(p, truncation(p, cux))                                    # :262-264
    tuy = ifft(p, truncation(p, cuy))
    tphi1 = ifft(p, truncation(p, _i_times(kx[:, None], cphi1)))         # :272-284
    tphi2 = ifft(p, truncation(p, _i_times(ky[None, :], cphi1)))
    wall = (1.0 - vpm_fields(p)[0])[:, None] if p.vpm else 1.0
    cnlphi1 = dfft((tphi1 * tux + tphi2 * tuy) * wall)                   # :292-300
    tphi1 = ifft(p, truncation(p, _i_times(kx[:, None], cphi2)))         # :314-326
    tphi2 = ifft(p, truncation(p, _i_times(ky[None, :], cphi2)))
    cnlphi2 = dfft((tphi1 * tux + tphi2 * tuy) * wall)                   # :334-348
    cnlphi1 = -cnlphi1 - _cscale(cmuF1, (p.lam / g1) * ks)              # :369
    cnlphi2 = -cnlphi2 - _cscale(cmuF2, (p.lam / g2) * ks)              # :370
    return cnlz, cnlphi1, cnlphi2


# ---------------------------------------------------------------------------
# L5: ETD2RK (etd2rk.c:14-224)
# ---------------------------------------------------------------------------
_EXP = np.frompyfunc(math.exp, 1, 1)


def _libm_exp(a):
    """Elementwise libm exp (bit-identical to C's exp on this host)."""
    return np.asarray(_EXP(np.asarray(a, dtype=np.float64)), dtype=np.float64)


def etd_tables(p: Params):
    """Per-mode ETD2RK coefficients in the reference's exact operation order.

    Returns dict with, for omega: eta, z1w (stage 1), z2w (stage 2); for the
    phase field(s): beta, z1c, z2c.  Follows etd2rk.c:88-127 (stage 1) and
    :162-203 (stage 2): thresholds |c| > 1e-12 (binary stage 2: c != 0.0),
    fallbacks dt and 0.5*dt.
    """
    ks = ksqr_grid(p)
    nu, dt = p.nu, p.dt
    c = -nu * ks                                                     # :90,164
    eta = _libm_exp(-nu * dt * ks)                                   # :91,165
    big = np.abs(c) > 1e-12                                          # :92,167
    cs = np.where(big, c, 1.0)
    z1w = np.where(big, (eta - 1.0) / cs, dt)                        # :94,96
    z2w = np.where(big, (eta - 1.0 - cs * dt) / (dt * cs * cs), 0.5 * dt)  # :169,171
    out = dict(eta=eta, z1w=z1w, z2w=z2w)
    if p.model == MODEL_NS:
        return out
    if p.model == MODEL_BINARY:
        cc = -1.5 * p.lam * p.sigma * p.epsilon * ks * ks            # :102,177
        big2 = np.abs(cc) != 0.0                                     # :181
    else:
        cc = -0.75 * p.lam * p.epsilon                               # :117,191
        cc = cc * (ks * ks)                                          # :118,192
        big2 = np.abs(cc) > 1e-12                                    # :195
    beta = _libm_exp(dt * cc)                                        # :104,119,179,193
    big1 = np.abs(cc) > 1e-12                                        # :106,121
    ccs1 = np.where(big1, cc, 1.0)
    z1c = np.where(big1, (beta - 1.0) / ccs1, dt)                    # :108,123
    ccs2 = np.where(big2, cc, 1.0)
    num = (beta - 1.0 - ccs2 * dt)                                   # :183,197
    z2c = np.where(big2, num / (dt * ccs2 * ccs2), 0.5 * dt)        # :183,198
    out.update(beta=beta, z1c=z1c, z2c=z2c)
    return out


@dataclass
class State:
    """Spectral state: omega-hat and the phase field(s) (IDXC [Nx][Ny2])."""

    cwz: np.ndarray
    cphi1: np.ndarray | None = None
    cphi2: np.ndarray | None = None

    def copy(self):
        return State(self.cwz.copy(),
                     None if self.cphi1 is None else self.cphi1.copy(),
                     None if self.cphi2 is None else self.cphi2.copy())

    def fields(self):
        return [f for f in (self.cwz, self.cphi1, self.cphi2) if f is not None]


def rhs(p: Params, s: State):
    """N(u): potential() then nonlin() as etd2rk.c:28,40,51,75 call them."""
    if p.model == MODEL_NS:
        return [nonlin(p, s.cwz)]
    if p.model == MODEL_BINARY:
        return list(nonlin(p, s.cwz, s.cphi1))
    cmuF1, cmuF2 = potential(p, s.cphi1, s.cphi2)                    # :40,149
    return list(nonlin(p, s.cwz, s.cphi1, s.cphi2, cmuF1, cmuF2))    # :51,150


def etd2rk(p: Params, s: State, tables=None):
    """One ETD2RK step (etd2rk.c:14-224); returns a new State."""
    tb = etd_tables(p) if tables is None else tables
    n1 = rhs(p, s)                                                   # stage-1 RHS
    nn = [x.copy() for x in n1]                                      # :62-65 save N_n
    f = s.fields()
    a = [_cscale(f[0], tb["eta"]) + _cscale(n1[0], tb["z1w"])]       # :99
    for k in range(1, len(f)):
        a.append(_cscale(f[k], tb["beta"]) + _cscale(n1[k], tb["z1c"]))  # :113,129-130
    sa = State(*a)
    n2 = rhs(p, sa)                                                  # :146-153
    out = [a[0] + _cscale(n2[0] - nn[0], tb["z2w"])]                 # :174
    for k in range(1, len(f)):
        out.append(a[k] + _cscale(n2[k] - nn[k], tb["z2c"]))         # :188,205-206
    return State(*out)


def run(p: Params, s: State, nsteps: int, spectrum_every: int = 0):
    """Advance nsteps with the main.c:260-368 order.

    ``spectrum_every`` > 0 reproduces the in-place truncation that
    ``spectrum()`` applies to the phase fields every navg/2 steps
    (main.c:281-284, spectrum.c:24-26).
    """
    tb = etd_tables(p)
    for it in range(1, nsteps + 1):
        s = etd2rk(p, s, tb)
        if spectrum_every and it % spectrum_every == 0:
            s = truncate_state(p, s)
    return s


def main_loop(p: Params, s: State, maxiter: int, nfile: int, diag_every: int = 1):
    """The time loop of main.c:214-367 with its diagnostics and dumps.

    Returns the records the reference writes, in order: ``ken`` rows
    (t, eu[, eph1[, eph2]]) of kinetic_energy.c:109,135,141; ``fe`` rows
    (t, ephi, egphi) of free_energy.c:149; ``gamma`` rows per phase
    (t, Sp_0, Sp_t, def) of perimeter.c:117-123,253-265; ``spectra`` (one dict
    per spectrum() call, spectrum.c:150-187) and ``dumps`` (real fields
    [wz, phi1, phi2] per dump, main.c:230-250 at t = 0 without truncation,
    main.c:330-357 every navg steps with truncation).  ``diag_every`` > 1
    thins the per-step diagnostics (1 = the reference).
    """
    navg = maxiter // nfile                                          # main.c:52
    tb = etd_tables(p)
    nph = {MODEL_NS: 0, MODEL_BINARY: 1, MODEL_TERNARY: 2}[p.model]
    rec = {"ken": [], "fe": [], "gamma": [[] for _ in range(nph)], "spectra": [], "dumps": []}
    t = 0.0

    def diagnostics(s):
        if nph:
            rec["fe"].append((t,) + tuple(free_energy(p, s)))
        ke, cnts = kinetic_energy(p, s, counts=True)
        if nph != 1 or cnts[0] != 0:                                 # binary: kinetic_energy.c:103-108
            rec["ken"].append((t,) + tuple(ke))
        if nph:
            for ph, (c1, c2) in enumerate(perimeter(p, s)):
                spt = np.float64(c1 * p.dx)
                sp0 = np.float64(2.0 * math.sqrt(PI * c2 * p.dx * p.dy))
                with np.errstate(divide="ignore", invalid="ignore"):   # no phase-1 points: 0/0 as in C
                    rec["gamma"][ph].append((t, float(sp0), float(spt), float((spt / sp0) - 1.0)))

    def dump(s, trunc):
        f = [s.cwz] + ([s.cphi1] if nph >= 1 else []) + ([s.cphi2] if nph == 2 else [])
        rec["dumps"].append([ifft(p, truncation(p, c) if trunc else c) for c in f])

    diagnostics(s)
    sp, s = spectrum(p, s)
    rec["spectra"].append(sp)
    dump(s, False)
    for it in range(1, maxiter + 1):                                 # main.c:260
        t += p.dt                                                    # :263
        s = etd2rk(p, s, tb)
        if it % diag_every == 0:
            diagnostics(s)
        if it % (navg // 2) == 0:                                    # :281
            sp, s = spectrum(p, s)
            rec["spectra"].append(sp)
        if it % navg == 0:                                           # :301
            dump(s, True)
    return rec, s


def truncate_state(p: Params, s: State):
    """State side effect of spectrum(): truncation of the phase fields (spectrum.c:22-27)."""
    return State(s.cwz,
                 None if s.cphi1 is None else truncation(p, s.cphi1),
                 None if s.cphi2 is None else truncation(p, s.cphi2))


# ---------------------------------------------------------------------------
# L6: diagnostics (free_energy.c, kinetic_energy.c, spectrum.c, perimeter.c)
# ---------------------------------------------------------------------------
def free_energy(p: Params, s: State):
    """(ephi, egphi) from *untruncated* fields (free_energy.c:13-157)."""
    kx, ky = kxky(p)
    scale = 1.0 / (p.Nx * p.Ny)
    eps = p.epsilon
    if p.model == MODEL_BINARY:
        c = ifft(p, s.cphi1)
        t = c * c * (1.0 - c) * (1.0 - c)
        t = t * (12.0 * (p.sigma / eps))
        ephi = float(np.sum(t * scale))
        tux = ifft(p, _i_times(kx[:, None], s.cphi1))
        tuy = ifft(p, _i_times(ky[None, :], s.cphi1))
        egphi = float(np.sum(tux * tux + tuy * tuy)) * ((0.75 * p.sigma * eps) * scale)
        return ephi, egphi
    g1, g2, g3, _ = p.gammas
    c1 = ifft(p, s.cphi1)
    c2 = ifft(p, s.cphi2)
    c3 = 1.0 - c1 - c2
    t = (g1 * c1 * c1 * (1.0 - c1) * (1.0 - c1)
         + g2 * c2 * c2 * (1.0 - c2) * (1.0 - c2)
         + g3 * c3 * c3 * (c1 + c2) * (c1 + c2))
    t = t * (6.0 / eps)
    ephi = float(np.sum(t * scale))
    c1x = ifft(p, _i_times(kx[:, None], s.cphi1))
    c1y = ifft(p, _i_times(ky[None, :], s.cphi1))
    c2x = ifft(p, _i_times(kx[:, None], s.cphi2))
    c2y = ifft(p, _i_times(ky[None, :], s.cphi2))
    eg = (g1 * (c1x * c1x + c1y * c1y) + g2 * (c2x * c2x + c2y * c2y)
          + g3 * (c1x + c2x) * (c1x + c2x) + g3 * (c1y + c2y) * (c1y + c2y))
    egphi = float(np.sum(eg)) * (0.375 * eps * scale)
    return ephi, egphi


def kinetic_energy(p: Params, s: State, counts: bool = False):
    """Mean KE and per-phase KE (kinetic_energy.c:3-144).

    Returns (eu,) for NS, (eu, eph) binary, (eu, eph1, eph2) ternary, with the
    per-phase values already divided by their counts (when non-zero); with
    ``counts``, also the list of those counts (points with c_i >= 0.5).
    """
    cux, cuy = get_velocity(p, s.cwz)
    ux = ifft(p, truncation(p, cux))
    uy = ifft(p, truncation(p, cuy))
    scale = 1.0 / (p.Nx * p.Ny)
    e = 0.5 * (ux * ux + uy * uy)
    eu = float(np.sum(e)) * scale
    if p.model == MODEL_NS:
        return ((eu,), []) if counts else (eu,)
    outs, cnts = [eu], []
    for cphi in ([s.cphi1] if p.model == MODEL_BINARY else [s.cphi1, s.cphi2]):
        ph = ifft(p, truncation(p, cphi))
        m = ph >= 0.5
        cnt = int(np.sum(m))
        eph = float(np.sum(np.where(m, e, 0.0)))
        outs.append(eph / cnt if cnt != 0 else eph)
        cnts.append(cnt)
    return (tuple(outs), cnts) if counts else tuple(outs)


def spectrum(p: Params, s: State):
    """Shell spectra (spectrum.c:3-194) and the truncated state (:21-27).

    Returns (spectra dict, new State).  Shell index ik = (int)(|k|/dk + 0.5);
    weight 1/2 (ek, ensk) or 1 (phase spectra) at ky = 0 or ky = kymax, else
    1 (ek) or 2 (phase).  The x/y spectra are sized N/2+1 here (the
    reference's kxm-sized arrays overflow for factor <= 1).
    """
    kx, ky = kxky(p)
    s2 = truncate_state(p, s)
    cux, cuy = get_velocity(p, s.cwz)
    cux = truncation(p, cux)
    cuy = truncation(p, cuy)
    dk = 2.0 * PI / p.lx
    scale = 1.0 / (p.Nx * p.Ny)
    kxs = (kx * kx)[:, None] + 0.0 * ky[None, :]
    kys = 0.0 * kx[:, None] + (ky * ky)[None, :]
    ks = kxs + kys
    ik = ((np.sqrt(ks) / dk) + 0.5).astype(np.int64)
    ikx = ((np.sqrt(kxs) / dk) + 0.5).astype(np.int64)
    iky = ((np.sqrt(kys) / dk) + 0.5).astype(np.int64)
    eu = (cux * np.conj(cux) + cuy * np.conj(cuy)).real * (scale * scale)
    ens = eu * ks
    edge = (ky == 0.0) | (ky == p.kymax)
    w_half = np.where(edge, 0.5, 1.0)[None, :] * np.ones((p.Nx, 1))
    w_two = np.where(edge, 1.0, 2.0)[None, :] * np.ones((p.Nx, 1))
    ns = p.nshell
    out = {}
    out["ek"] = np.bincount(ik.ravel(), (w_half * eu).ravel(), minlength=ns)[:ns]
    out["ensk"] = np.bincount(ik.ravel(), (w_half * ens).ravel(), minlength=ns)[:ns]
    nxy = p.Nx // 2 + 2
    out["ekx"] = np.bincount(ikx.ravel(), (w_half * eu).ravel(), minlength=nxy)
    out["eky"] = np.bincount(iky.ravel(), (w_half * eu).ravel(), minlength=nxy)
    if p.model == MODEL_BINARY:
        c = s2.cphi1
        eph = (c * np.conj(c)).real * (scale * scale)
        out["sph"] = np.bincount(ik.ravel(), (w_two * eph).ravel(), minlength=ns)[:ns]
    elif p.model == MODEL_TERNARY:
        c1, c2 = s2.cphi1, s2.cphi2
        e1 = (c1 * np.conj(c1)).real * (scale * scale)
        e2 = (c2 * np.conj(c2)).real * (scale * scale)
        e12 = np.abs((c1.real * c2.real + c1.imag * c2.imag) * (scale * scale))
        out["sph1"] = np.bincount(ik.ravel(), (w_two * e1).ravel(), minlength=ns)[:ns]
        out["sph2"] = np.bincount(ik.ravel(), (w_two * e2).ravel(), minlength=ns)[:ns]
        out["sp12"] = np.bincount(ik.ravel(), (w_two * e12).ravel(), minlength=ns)[:ns]
    return out, s2


def perimeter(p: Params, s: State):
    """Interface counts by sign change of thresholded c (perimeter.c:127-266).

    Returns a list of (count_interface, count_inside) per phase field, using
    the untruncated c (perimeter.c:128-131) thresholded at 0.5 and periodic
    neighbours in x and y.
    """
    outs = []
    fields = [s.cphi1] if p.model == MODEL_BINARY else [s.cphi1, s.cphi2]
    for cphi in fields:
        ph = np.where(ifft(p, cphi) >= 0.5, 1.0, -1.0)
        nx = np.roll(ph, -1, axis=0)
        ny = np.roll(ph, -1, axis=1)
        edge = ((ph * nx) < 0.0) | ((ph * ny) < 0.0)
        outs.append((int(np.sum(edge)), int(np.sum(ph > 0.0))))
    return outs


# ---------------------------------------------------------------------------
# L7: initial conditions (init.c) and synthetic-input generators
# ---------------------------------------------------------------------------
_libc = None


def _glibc():
    global _libc
    if _libc is None:
        _libc = ctypes.CDLL("libc.so.6")
        _libc.rand.restype = ctypes.c_int
        _libc.srand.argtypes = [ctypes.c_uint]
    return _libc


def glibc_rand_stream(seed: int, count: int):
    """``srand(seed)`` then ``count`` calls of ``rand()/(double)RAND_MAX``."""
    lc = _glibc()
    lc.srand(seed)
    RAND_MAX = 2147483647.0
    return np.array([lc.rand() for _ in range(count)], dtype=np.float64) / RAND_MAX


def init_ternary_noise_real(p: Params, nranks: int = 1):
    """Real c1, c2 of init.c:265-289: per-rank ``srand(10+2*rank)``, 1/4 + 0.001*rand.

    Rank r owns rows [r*Nx/nranks, (r+1)*Nx/nranks) (FFTW slab, Nx divisible).
    """
    if p.Nx % nranks:
        raise ValueError("nranks must divide Nx")
    ln = p.Nx // nranks
    c1 = np.empty((p.Nx, p.Ny))
    c2 = np.empty((p.Nx, p.Ny))
    for r in range(nranks):
        rn = glibc_rand_stream(10 + r * 2, 2 * ln * p.Ny).reshape(ln, p.Ny, 2)
        c1[r * ln:(r + 1) * ln] = (1.0 / 4.0) + 0.001 * rn[..., 0]
        c2[r * ln:(r + 1) * ln] = (1.0 / 4.0) + 0.001 * rn[..., 1]
    return c1, c2


def init_binary_droplet_real(p: Params):
    """Binary tanh droplet of init.c:227-246."""
    x = (np.arange(p.Nx) * p.dx)[:, None]
    y = (np.arange(p.Ny) * p.dy)[None, :]
    r = np.sqrt((x - p.lx / 2.0) * (x - p.lx / 2.0) + (y - p.ly / 2.0) * (y - p.ly / 2.0))
    phi = 1.0 - np.tanh((r - (PI / 4.5)) / (0.5 * p.epsilon))
    return phi * 0.5


def splitmix64_uniform(seed: int, count: int):
    """U[0,1) from the top 53 bits of SplitMix64 (Steele, Lea, Flood 2014)."""
    M = (1 << 64) - 1
    st = seed & M
    out = np.empty(count, dtype=np.float64)
    for n in range(count):
        st = (st + 0x9E3779B97F4A7C15) & M
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z = z ^ (z >> 31)
        out[n] = (z >> 11) * (1.0 / 9007199254740992.0)
    return out


def init_noise_real(p: Params, seed: int, mean: float, amp: float):
    """c = mean + U[-amp, amp] (binary) or c1, c2 interleaved per point (ternary).

    Points row-major (x slow); u -> mean + amp*(2u - 1).
    """
    n = p.Nx * p.Ny
    if p.model == MODEL_BINARY:
        u = splitmix64_uniform(seed, n).reshape(p.Nx, p.Ny)
        return (mean + amp * (2.0 * u - 1.0),)
    u = splitmix64_uniform(seed, 2 * n).reshape(p.Nx, p.Ny, 2)
    return mean + amp * (2.0 * u[..., 0] - 1.0), mean + amp * (2.0 * u[..., 1] - 1.0)


def init_lens_real(p: Params, radius: float = 0.4, sep: float = 1.2):
    """Ternary liquid-lens IC: c1 below / c2 above a band interface, two c3 lenses.

    H(y) = (tanh((y-L/4)/(eps/2)) - tanh((y-3L/4)/(eps/2)))/2; lenses
    c3 = max over two discs of (1 - tanh((r - radius)/(eps/2)))/2 centred at
    (L/2 +- sep/2, 3L/4); c1 = (1-c3)(1-H), c2 = (1-c3) H.
    """
    x = (np.arange(p.Nx) * p.dx)[:, None] * np.ones((1, p.Ny))
    y = (np.arange(p.Ny) * p.dy)[None, :] * np.ones((p.Nx, 1))
    w = 0.5 * p.epsilon
    H = 0.5 * (np.tanh((y - p.ly / 4.0) / w) - np.tanh((y - 3.0 * p.ly / 4.0) / w))
    c3 = np.zeros_like(x)
    for sgn in (-1.0, 1.0):
        xc = p.lx / 2.0 + sgn * sep / 2.0
        yc = 3.0 * p.ly / 4.0
        r = np.sqrt((x - xc) * (x - xc) + (y - yc) * (y - yc))
        c3 = np.maximum(c3, 0.5 * (1.0 - np.tanh((r - radius) / w)))
    c1 = (1.0 - c3) * (1.0 - H)
    c2 = (1.0 - c3) * H
    return c1, c2


def state_from_real(p: Params, wz=None, phi1=None, phi2=None):
    """dfft + truncation of real fields, as init.c:253-262 / 291-294 do."""
    cw = np.zeros((p.Nx, p.Ny2), dtype=np.complex128) if wz is None else truncation(p, dfft(wz))
    c1 = None if phi1 is None else truncation(p, dfft(phi1))
    c2 = None if phi2 is None else truncation(p, dfft(phi2))
    return State(cw, c1, c2)


def random_vorticity(p: Params, seed: int, amp: float = 1.0, kmax: float | None = None):
    """Smooth random real vorticity field (truncated), used for test inputs."""
    rng = np.random.default_rng(seed)
    kx, ky = kxky(p)
    ks = ksqr_grid(p)
    km = (p.Nx / 8.0) * p.factor_x if kmax is None else kmax
    spec = (rng.standard_normal((p.Nx, p.Ny2)) + 1j * rng.standard_normal((p.Nx, p.Ny2)))
    spec *= np.exp(-ks / (km * km)) * amp * p.Nx * p.Ny / 10.0
    spec[0, 0] = 0.0
    w = np.fft.irfft2(spec, s=(p.Nx, p.Ny))
    return w


def rel_l2(a, b):
    """Relative L2 distance ||a-b|| / ||b||."""
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


PINS = {
    # function: the independent check (tests/test_oracle_pins.py unless named)
    "dfft/ifft": "direct O(N^4) DFT, Parseval, single-mode transforms",
    "kxky/truncation": "Nyquist sign; kept-mode count equals the integer lattice count",
    "get_velocity/curl": "curl(get_velocity(w)) = w, div u = 0, single-mode closed form",
    "potential": "exact rational evaluation (fractions) + pure-phase / symmetric zeros",
    "surface_tension": "sympy real-space closed form for band-limited fields; 1-D fields give zero",
    "nonlin": "sympy closed form of u.grad w, u.grad c and lap(muF); Taylor-Green zero advection",
    "etd_tables": "closed form where well conditioned; bit-compared with the C tables (tests/test_abi.py)",
    "etd2rk": "Stokes decay, forced Kolmogorov steady state, mass conservation, "
              "binary Cahn-Hilliard linear growth rate of one step",
    "free_energy": "binary closed form; d/dh F = <mu, dc> against the chemical potential; "
                   "ternary reduces to binary on two-phase states",
    "kinetic_energy": "single-mode closed form A^2/(4k^2), per-phase means",
    "spectrum": "Parseval (sum ek = eu, sum sph1 = <c1^2>); single-mode phase spectrum",
    "perimeter": "band of two straight interfaces: Sp_t = 2L, half the points inside",
    "init_ternary_noise_real": "glibc rand known values; bit-compared with the C driver (tests/test_driver.py)",
    "init_binary_droplet_real": "0.5 level on the circle r = pi/4.5; compared with the C driver",
    "splitmix64_uniform": "published first output for seed 0",
    "main_loop": "cadence and record counts against the C driver's output files (tests/test_driver.py)",
}
