"""Oracle: discrete adjoint of the oracle's ternary ETD2RK step.

TEST INFRASTRUCTURE ONLY.  Like ``chns_oracle`` this module is the checker
for the HIP product path; only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import it.

The reference has no adjoint.  This module is the exact transpose (vector-
Jacobian product) of ``chns_oracle.etd2rk`` for the TERNARY model, i.e. of
the reference step etd2rk.c:14-224 with potential() (chem_potential.c:12-94),
nonlin() (nonlin.c:13-381) and surface_tension() (surface_tension.c:114-225)
as the oracle restates them.  Pins (``tests/test_adjoint_oracle.py``):
central finite differences of the oracle forward for the state and for
sigma_12/13/23, and the transpose identity <J v, w> = <v, J^T w> with J v
from finite differences.

Conventions.  A half-spectrum X ([Nx][Ny2], Hermitian-consistent) stands for
the real field x = irfft2(X).  Its adjoint Xbar is the unnormalised r2c of
the real-space gradient dL/dx, so that

    dL = sum_x xbar(x) dx(x) = (1/N^2) sum_half w_j Re(conj(Xbar) dX),

w_j = 1 at ky = 0 and ky = Ny/2, else 2 (``spec_dot``).  With it, the
transpose of a real Fourier multiplier m(k) is conj(m(k)), the transpose of
irfft2 (c2r) is rfft2 (r2c) and vice versa, and pointwise products transpose
pointwise.  Every read of the state in the right-hand side goes through
truncation(), so the modes outside the dealias disc never feed back and the
adjoint state lives on the kept modes only.
"""
from __future__ import annotations

import numpy as np

from . import chns_oracle as O

SIGMA_PARTIALS = ("fac", "G1", "G2", "G3", "g1i", "g2i", "g3i", "mc1", "mc2",
                  "A11", "A12", "A21", "A22", "lg1", "lg2")


def spec_dot(p: O.Params, A, B):
    """(1/N^2) sum over the half-spectrum of w_j Re(conj(A) B) = sum_x a(x) b(x)."""
    w = np.full(p.Ny2, 2.0)
    w[0] = 1.0
    if p.Ny % 2 == 0:
        w[-1] = 1.0
    return float(np.sum(w[None, :] * (A.real * B.real + A.imag * B.imag))) / (p.Nx * p.Ny)


def _zero_partials():
    return {k: 0.0 for k in SIGMA_PARTIALS}


def potential_vjp(p: O.Params, c1, c2, r1, r2):
    """Transpose of potential_pointwise (chem_potential.c:40-86) at (c1, c2).

    r1, r2: real-space adjoints of muF1, muF2.  Returns (c1bar, c2bar, partials)
    with the real-space sums of the coefficient partials ``fac``, ``G1..3``
    (the gamma_j multiplying fc_j inside G_j) and ``g1i..g3i`` (1/gamma_j).
    """
    g1, g2, g3, dl = p.gammas
    g1i, g2i, g3i = 1.0 / g1, 1.0 / g2, 1.0 / g3
    fac = 2.0 * dl / p.epsilon
    X6 = 6.0 * p.Xlambda
    c3 = 1.0 - c1 - c2
    c123 = c1 * c2 * c3
    fc = [c * (1.0 - c) * (1.0 - 2.0 * c) for c in (c1, c2, c3)]
    G1 = g1 * fc[0] + X6 * c123 * c2 * c3
    G2 = g2 * fc[1] + X6 * c123 * c1 * c3
    G3 = g3 * fc[2] + X6 * c123 * c1 * c2
    m1 = (G1 - G2) * g2i + (G1 - G3) * g3i
    m2 = (G2 - G1) * g1i + (G2 - G3) * g3i
    part = {
        "fac": float(np.sum(r1 * m1 + r2 * m2)),
        "g1i": float(np.sum(fac * r2 * (G2 - G1))),
        "g2i": float(np.sum(fac * r1 * (G1 - G2))),
        "g3i": float(np.sum(fac * (r1 * (G1 - G3) + r2 * (G2 - G3)))),
    }
    Gb1 = fac * (r1 * (g2i + g3i) - r2 * g1i)
    Gb2 = fac * (-r1 * g2i + r2 * (g1i + g3i))
    Gb3 = -fac * (r1 + r2) * g3i
    part["G1"] = float(np.sum(Gb1 * fc[0]))
    part["G2"] = float(np.sum(Gb2 * fc[1]))
    part["G3"] = float(np.sum(Gb3 * fc[2]))
    fcp = [1.0 - 6.0 * c + 6.0 * c * c for c in (c1, c2, c3)]
    # partial derivatives of G_j with (c1, c2, c3) independent
    d1 = Gb1 * (g1 * fcp[0] + X6 * c2 * c2 * c3 * c3) + Gb2 * (2.0 * X6 * c1 * c2 * c3 * c3) \
        + Gb3 * (2.0 * X6 * c1 * c2 * c2 * c3)
    d2 = Gb1 * (2.0 * X6 * c1 * c2 * c3 * c3) + Gb2 * (g2 * fcp[1] + X6 * c1 * c1 * c3 * c3) \
        + Gb3 * (2.0 * X6 * c1 * c1 * c2 * c3)
    d3 = Gb1 * (2.0 * X6 * c1 * c2 * c2 * c3) + Gb2 * (2.0 * X6 * c1 * c1 * c2 * c3) \
        + Gb3 * (g3 * fcp[2] + X6 * c1 * c1 * c2 * c2)
    return d1 - d3, d2 - d3, part   # c3 = 1 - c1 - c2


def rhs_vjp(p: O.Params, s: O.State, gw, g1, g2):
    """Transpose of the ternary right-hand side N (potential + nonlin) at s.

    gw, g1, g2: spectral adjoints of N_omega, N_c1, N_c2 (kept modes only).
    Returns ([wbar, c1bar, c2bar], partials): the spectral adjoints of the
    state (kept modes) and the sums of the coefficient partials.
    """
    if p.model != O.MODEL_TERNARY:
        raise ValueError("the oracle adjoint covers the ternary model")
    kx, ky = O.kxky(p)
    KX, KY = kx[:, None], ky[None, :]
    ks = O.ksqr_grid(p)
    mask = O.dealias_mask(p)
    T = lambda X: np.where(mask, X, 0.0 + 0.0j)   # noqa: E731
    I = lambda X: O.ifft(p, X)                     # noqa: E731
    R = O.dfft
    ik = O._i_times
    g1_, g2_, g3_, _ = p.gammas
    eps = p.epsilon
    mc1, mc2 = 0.75 * g1_ * eps, 0.75 * g2_ * eps
    a11, a12, a21, a22 = 1.0 + g3_ / g1_, g3_ / g2_, g3_ / g1_, 1.0 + g3_ / g2_
    lg1, lg2 = p.lam / g1_, p.lam / g2_
    Q = np.where(ks > 1e-12, 1.0 / np.where(ks > 1e-12, ks, 1.0), 0.0)
    cw, cp1, cp2 = s.cwz, s.cphi1, s.cphi2

    # forward intermediates at s (nonlin.c, surface_tension.c, chem_potential.c)
    c1, c2 = I(T(cp1)), I(T(cp2))
    cx1, cy1 = I(T(ik(KX, cp1))), I(T(ik(KY, cp1)))
    cx2, cy2 = I(T(ik(KX, cp2))), I(T(ik(KY, cp2)))
    cux, cuy = O.get_velocity(p, cw)
    u, v = I(T(cux)), I(T(cuy))
    wx, wy = I(T(ik(KX, cw))), I(T(ik(KY, cw)))
    muF1, muF2 = O.potential_pointwise(p, c1, c2)
    M1, M2 = R(muF1), R(muF2)
    mu1 = I(T(M1 + mc1 * ks * cp1))
    mu2 = I(T(M2 + mc2 * ks * cp2))
    A1 = a11 * mu1 + a12 * mu2
    A2 = a21 * mu1 + a22 * mu2

    # N_w = -R(u wx + v wy) + i kx R(Fy) - i ky R(Fx); N_i = -R(u cx_i + v cy_i) - lg_i k^2 M_i
    xi = I(gw)
    xix, xiy = I(ik(KX, gw)), I(ik(KY, gw))
    nlzb, Fxb, Fyb = -xi, xiy, -xix
    ad1, ad2 = -I(g1), -I(g2)
    ub = nlzb * wx + ad1 * cx1 + ad2 * cx2
    vb = nlzb * wy + ad1 * cy1 + ad2 * cy2
    wxb, wyb = nlzb * u, nlzb * v
    cx1b, cy1b = ad1 * u + Fxb * A1, ad1 * v + Fyb * A1
    cx2b, cy2b = ad2 * u + Fxb * A2, ad2 * v + Fyb * A2
    A1b = Fxb * cx1 + Fyb * cy1
    A2b = Fxb * cx2 + Fyb * cy2
    mu1b = a11 * A1b + a21 * A2b
    mu2b = a12 * A1b + a22 * A2b
    part = _zero_partials()
    part["A11"] = float(np.sum(A1b * mu1))
    part["A12"] = float(np.sum(A1b * mu2))
    part["A21"] = float(np.sum(A2b * mu1))
    part["A22"] = float(np.sum(A2b * mu2))
    Mu1, Mu2 = T(R(mu1b)), T(R(mu2b))
    part["mc1"] = spec_dot(p, Mu1, ks * cp1)
    part["mc2"] = spec_dot(p, Mu2, ks * cp2)
    part["lg1"] = -spec_dot(p, g1, ks * M1)
    part["lg2"] = -spec_dot(p, g2, ks * M2)
    r1 = I(Mu1 - lg1 * ks * g1)
    r2 = I(Mu2 - lg2 * ks * g2)
    c1b, c2b, pp = potential_vjp(p, c1, c2, r1, r2)
    for k, val in pp.items():
        part[k] += val
    wbar = T(Q * (ik(-KY, R(ub)) + ik(KX, R(vb))) + ik(-KX, R(wxb)) + ik(-KY, R(wyb)))
    c1bar = T(R(c1b) + ik(-KX, R(cx1b)) + ik(-KY, R(cy1b)) + mc1 * ks * Mu1)
    c2bar = T(R(c2b) + ik(-KX, R(cx2b)) + ik(-KY, R(cy2b)) + mc2 * ks * Mu2)
    return [wbar, c1bar, c2bar], part


def step_vjp(p: O.Params, s: O.State, ubar, tb=None):
    """Transpose of one ETD2RK step at s (etd2rk.c:14-224).

    Forward: a = E u + Z1 N(u); u' = a + Z2 (N(a) - N(u)).  Given ubar = dL/du'
    (kept modes), returns (dL/du, partials):
      abar = ubar + J_N(a)^T Z2 ubar,
      dL/du = E abar + J_N(u)^T (Z1 abar - Z2 ubar).
    """
    tb = O.etd_tables(p) if tb is None else tb
    n1 = O.rhs(p, s)
    f = s.fields()
    a = O.State(O._cscale(f[0], tb["eta"]) + O._cscale(n1[0], tb["z1w"]),
                *[O._cscale(f[k], tb["beta"]) + O._cscale(n1[k], tb["z1c"]) for k in (1, 2)])
    E = [tb["eta"], tb["beta"], tb["beta"]]
    Z1 = [tb["z1w"], tb["z1c"], tb["z1c"]]
    Z2 = [tb["z2w"], tb["z2c"], tb["z2c"]]
    g2 = [Z2[k] * ubar[k] for k in range(3)]
    J2, part2 = rhs_vjp(p, a, *g2)
    abar = [ubar[k] + J2[k] for k in range(3)]
    g1 = [Z1[k] * abar[k] - g2[k] for k in range(3)]
    J1, part1 = rhs_vjp(p, s, *g1)
    out = [E[k] * abar[k] + J1[k] for k in range(3)]
    return out, {k: part1[k] + part2[k] for k in SIGMA_PARTIALS}


def spectrum_loss(p: O.Params, s: O.State, target):
    """Loss sum_s (S11(s) - target(s))^2 on the c1 shell spectrum sph1 of the
    truncated state (spectrum.c:83-103) and its spectral adjoint.

    Returns (loss, S11, [wbar, c1bar, c2bar]) with
    c1bar = 4 (S11(s(k)) - target(s(k))) c1(k) / N^2 on the kept modes.
    """
    sp, s2 = O.spectrum(p, s)
    S = sp["sph1"]
    d = S - np.asarray(target, dtype=np.float64)
    loss = float(np.sum(d * d))
    kx, ky = O.kxky(p)
    ksq = (kx * kx)[:, None] + (ky * ky)[None, :]
    dk = 2.0 * O.PI / p.lx
    ik = ((np.sqrt(ksq) / dk) + 0.5).astype(np.int64)
    mask = O.dealias_mask(p)
    fac = np.where(mask, 4.0 * d[np.minimum(ik, len(d) - 1)] / (p.Nx * p.Ny), 0.0)
    z = np.zeros_like(s.cwz)
    return loss, S, [z, fac * s2.cphi1, z.copy()]


def free_energy_loss(p: O.Params, s: O.State):
    """Loss = ephi + egphi of free_energy.c:13-157 evaluated on the truncated
    phase fields T(c1), T(c2), its spectral adjoint, and its direct
    derivatives dL/dgamma_i (the gamma_i weights of the two energies).

    Returns (loss, ephi, egphi, [wbar, c1bar, c2bar], dgam[3]).  With
    s = c1 + c2 = 1 - c3, ephi = b sum f, f = g1 c1^2 (1-c1)^2 +
    g2 c2^2 (1-c2)^2 + g3 s^2 (1-s)^2, and egphi = q sum g1 |grad c1|^2 +
    g2 |grad c2|^2 + g3 |grad c1 + grad c2|^2 (b = 6 scale/eps,
    q = 0.375 eps scale); c_i bar = rfft2(df/dc_i) - i kx rfft2(dg/dc_i,x)
    - i ky rfft2(dg/dc_i,y) on the kept modes.
    """
    g1, g2, g3, _ = p.gammas
    kx, ky = O.kxky(p)
    mask = O.dealias_mask(p)
    scale = 1.0 / (p.Nx * p.Ny)
    b = (6.0 / p.epsilon) * scale
    q = 0.375 * p.epsilon * scale
    C1 = O.truncation(p, s.cphi1)
    C2 = O.truncation(p, s.cphi2)
    c1, c2 = O.ifft(p, C1), O.ifft(p, C2)
    c1x, c1y = O.ifft(p, 1j * kx[:, None] * C1), O.ifft(p, 1j * ky[None, :] * C1)
    c2x, c2y = O.ifft(p, 1j * kx[:, None] * C2), O.ifft(p, 1j * ky[None, :] * C2)
    sm = c1 + c2
    f1, f2, f3 = c1 * c1 * (1 - c1) ** 2, c2 * c2 * (1 - c2) ** 2, sm * sm * (1 - sm) ** 2
    e1, e2 = c1x * c1x + c1y * c1y, c2x * c2x + c2y * c2y
    e3 = (c1x + c2x) ** 2 + (c1y + c2y) ** 2
    ephi = b * float(np.sum(g1 * f1 + g2 * f2 + g3 * f3))
    egphi = q * float(np.sum(g1 * e1 + g2 * e2 + g3 * e3))
    dgam = [b * float(np.sum(f1)) + q * float(np.sum(e1)), b * float(np.sum(f2)) + q * float(np.sum(e2)),
            b * float(np.sum(f3)) + q * float(np.sum(e3))]
    d3 = g3 * 2 * sm * (1 - sm) * (1 - 2 * sm)
    x1 = b * (g1 * 2 * c1 * (1 - c1) * (1 - 2 * c1) + d3)
    x2 = b * (g2 * 2 * c2 * (1 - c2) * (1 - 2 * c2) + d3)
    y1x, y1y = q * 2 * (g1 * c1x + g3 * (c1x + c2x)), q * 2 * (g1 * c1y + g3 * (c1y + c2y))
    y2x, y2y = q * 2 * (g2 * c2x + g3 * (c1x + c2x)), q * 2 * (g2 * c2y + g3 * (c1y + c2y))

    def seed(x, yx, yy):
        z = O.dfft(x) - 1j * kx[:, None] * O.dfft(yx) - 1j * ky[None, :] * O.dfft(yy)
        return np.where(mask, z, 0.0)

    z = np.zeros_like(s.cwz)
    return ephi + egphi, ephi, egphi, [z, seed(x1, y1x, y1y), seed(x2, y2x, y2y)], dgam


def sigma_gradient(p: O.Params, part, dgam=(0.0, 0.0, 0.0)):
    """dL/d(sigma_12, sigma_13, sigma_23) from the coefficient partials.

    gamma_1 = s12 - s23 + s13, gamma_2 = s12 - s13 + s23, gamma_3 = s23 - s12 + s13,
    del = 6 g1 g2 g3 / (g1 g2 + g1 g3 + g2 g3) (main.c:166-170), fac = 2 del/eps,
    mc_i = 0.75 g_i eps, A11 = 1 + g3/g1, A12 = g3/g2, A21 = g3/g1, A22 = 1 + g3/g2,
    lg_i = lambda/g_i.  dgam: the loss's own dL/dgamma_i (free_energy_loss).
    """
    g1, g2, g3, _ = p.gammas
    eps, lam = p.epsilon, p.lam
    D = g1 * g2 + g1 * g3 + g2 * g3
    P3 = g1 * g2 * g3
    dd1 = 6.0 * g2 * g3 / D - 6.0 * P3 * (g2 + g3) / (D * D)
    dd2 = 6.0 * g1 * g3 / D - 6.0 * P3 * (g1 + g3) / (D * D)
    dd3 = 6.0 * g1 * g2 / D - 6.0 * P3 * (g1 + g2) / (D * D)
    rf = part["fac"] * 2.0 / eps
    dg1 = (part["G1"] - part["g1i"] / (g1 * g1) + 0.75 * eps * part["mc1"]
           - g3 / (g1 * g1) * (part["A11"] + part["A21"]) - lam / (g1 * g1) * part["lg1"] + rf * dd1)
    dg2 = (part["G2"] - part["g2i"] / (g2 * g2) + 0.75 * eps * part["mc2"]
           - g3 / (g2 * g2) * (part["A12"] + part["A22"]) - lam / (g2 * g2) * part["lg2"] + rf * dd2)
    dg3 = (part["G3"] - part["g3i"] / (g3 * g3) + (part["A11"] + part["A21"]) / g1
           + (part["A12"] + part["A22"]) / g2 + rf * dd3 + dgam[2])
    dg1 += dgam[0]
    dg2 += dgam[1]
    return np.array([dg1 + dg2 - dg3, dg1 - dg2 + dg3, -dg1 + dg2 + dg3])


def loss_and_grad(p: O.Params, s0: O.State, nsteps: int, target):
    """Forward nsteps from s0 (all states kept), the spectrum loss at the end
    (target None: the free-energy loss), and the reverse sweep.  Returns (loss,
    real-space gradients [gw, gc1, gc2] with respect to the real fields that
    set_real_state turns into s0, dL/dsigma [s12, s13, s23])."""
    tb = O.etd_tables(p)
    traj = [s0]
    for _ in range(nsteps):
        traj.append(O.etd2rk(p, traj[-1], tb))
    dgam = (0.0, 0.0, 0.0)
    if target is None:
        loss, _, _, ubar, dgam = free_energy_loss(p, traj[-1])
    else:
        loss, _, ubar = spectrum_loss(p, traj[-1], target)
    acc = _zero_partials()
    for n in range(nsteps - 1, -1, -1):
        ubar, part = step_vjp(p, traj[n], ubar, tb)
        for k in SIGMA_PARTIALS:
            acc[k] += part[k]
    grads = [O.ifft(p, x) for x in ubar]
    return loss, grads, sigma_gradient(p, acc, dgam), acc
