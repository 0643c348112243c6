"""GPU parity of the discrete adjoint (chns_adj_* through the C ABI).

Checks, in order of independence from the HIP code:
* the library gradient against central finite differences of the library's
  own forward step (1e-6 relative), including sigma;
* the library's reverse step and full reverse sweep against the oracle's
  adjoint (oracle/chns_adjoint.py, itself pinned by finite differences of the
  oracle forward) at 1e-10;
* bitwise repeatability, and the torch.autograd.Function wrapper a caller
  would write around the C ABI.
"""
import math

import numpy as np
import pytest

import chns_amd as C
import oracle as O
from oracle import chns_adjoint as OA

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi
TOL = 1e-10       # library vs oracle adjoint (FP64 round-off)
TOL_FD = 1e-6     # library gradient vs central differences of the library forward


def adj_case(N, dt=None, **over):
    kw = dict(Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI, nu=0.02, dt=2e-3 * (16.0 / N) ** 2 if dt is None else dt,
              lam=0.02, epsilon=1.5 * TWO_PI / N, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06, Xlambda=2.0)
    kw.update(over)
    op = O.Params(model="ternary", **kw)
    c1, c2 = O.init_noise_real(op, 5, 1.0 / 3.0, 0.1)
    reals = [O.random_vorticity(op, 6, amp=2.0), c1, c2]
    return op, C.params("ternary", **kw), kw, reals


def target_for(op, nsteps):
    tgt = O.state_from_real(op, None, *O.init_noise_real(op, 7, 1.0 / 3.0, 0.1))
    return O.spectrum(op, O.run(op, tgt, nsteps))[0]["sph1"]


def rel(a, b):
    return O.rel_l2(a, b)


@pytest.mark.parametrize("N", [16, 32, 64])
def test_reverse_step_matches_oracle(N):
    """One reverse step from arbitrary seeds: dL/du and the sigma partials."""
    op, cp, _, reals = adj_case(N)
    s0 = O.state_from_real(op, *reals)
    rng = np.random.default_rng(4)
    seeds = [O.truncation(op, O.dfft(rng.standard_normal((N, N)))) for _ in range(3)]
    ref, part = OA.step_vjp(op, s0, seeds)
    with C.Context(cp) as ctx:
        ctx.set_state(*s0.fields())
        ctx.adj_forward(1, 1)
        ctx.adj_seed(*seeds)
        ctx.adj_backward()
        got, gpart = ctx.adj_get()
    for a, b in zip(got, ref):
        assert rel(a, b) < TOL
    for k in OA.SIGMA_PARTIALS:
        assert abs(gpart[k] - part[k]) <= TOL * max(abs(part[k]), 1e-300) + 1e-14 * max(map(abs, part.values())), k


@pytest.mark.parametrize("N,nsteps,K", [(16, 5, 2), (32, 7, 3), (64, 4, 4)])
def test_loss_gradient_matches_oracle(N, nsteps, K):
    """Spectrum loss, gradients for the initial fields and sigma, with checkpoint segments."""
    op, cp, _, reals = adj_case(N)
    target = target_for(op, nsteps)
    loss, grads, gsig, _ = OA.loss_and_grad(op, O.state_from_real(op, *reals), nsteps, target)
    with C.Context(cp) as ctx:
        ctx.set_real_state(*reals)
        ctx.adj_forward(nsteps, K)
        gl, _ = ctx.adj_seed_spectrum(target)
        gw, g1, g2, gs = ctx.adj_backward()
    assert abs(gl - loss) <= 1e-10 * loss
    for a, b in zip((gw, g1, g2), grads):
        assert rel(a, b) < TOL
    assert np.all(np.abs(gs - gsig) <= 1e-9 * np.abs(gsig))


def gpu_loss(cp, reals, nsteps, target):
    with C.Context(cp) as ctx:
        ctx.set_real_state(*reals)
        ctx.step(nsteps)
        ctx.adj_forward(0, 1)
        loss, _ = ctx.adj_seed_spectrum(target)
    return loss


def richardson(f, h):
    """Central difference with the O(h^2) term removed: (4 D(h/2) - D(h)) / 3."""
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(h / 2) - f(-h / 2)) / h
    return (4.0 * d2 - d1) / 3.0


@pytest.mark.parametrize("N,nsteps,K", [(128, 12, 5), (512, 40, 10)])
def test_gradient_matches_central_differences(N, nsteps, K):
    """Library gradient against central differences of the library forward."""
    op, cp, kw, reals = adj_case(N)
    target = O.spectrum(op, O.state_from_real(op, None, *O.init_noise_real(op, 7, 1.0 / 3.0, 0.1)))[0]["sph1"]
    with C.Context(cp) as ctx:
        ctx.set_real_state(*reals)
        ctx.adj_forward(nsteps, K)
        ctx.adj_seed_spectrum(target)
        gw, g1, g2, gs = ctx.adj_backward()
    rng = np.random.default_rng(8)
    for k, g in ((1, g1), (2, g2)):
        d = rng.standard_normal((N, N))

        def f(h):
            moved = [x.copy() for x in reals]
            moved[k] += h * d
            return gpu_loss(cp, moved, nsteps, target)

        fd = richardson(f, 1e-5)   # larger steps leave the asymptotic range at 512^2
        ad = float(np.sum(g * d))
        assert abs(fd - ad) <= TOL_FD * abs(ad), (k, fd, ad)
    for k, name in enumerate(("sigma_12", "sigma_13", "sigma_23")):
        def f(h):
            return gpu_loss(C.params("ternary", **{**kw, name: kw[name] + h}), reals, nsteps, target)

        fd = richardson(f, 1e-5 * kw[name])
        assert abs(fd - gs[k]) <= TOL_FD * abs(gs[k]), (name, fd, gs[k])


def test_backward_is_repeatable_and_leaves_state():
    op, cp, _, reals = adj_case(128)
    target = target_for(op, 0)
    outs = []
    with C.Context(cp) as ctx:
        ctx.set_real_state(*reals)
        ctx.adj_forward(9, 4)
        final = ctx.get_state()
        for _ in range(2):
            ctx.adj_seed_spectrum(target)
            outs.append(ctx.adj_backward())
        after = ctx.get_state()
        ctx.step(1)   # the context still steps after a reverse sweep
        ctx.sync()
    for a, b in zip(outs[0][:3], outs[1][:3]):
        assert np.array_equal(a, b)
    assert np.array_equal(outs[0][3], outs[1][3])
    for a, b in zip(final, after):
        assert np.array_equal(a, b)


def test_rejects_non_ternary_and_out_of_order():
    kw = dict(Nx=32, Ny=32, lx=TWO_PI, ly=TWO_PI)
    with C.Context(C.params("binary", **kw)) as ctx:
        with pytest.raises(C.ChnsError):
            ctx.adj_forward(1, 1)
    with C.Context(C.params("ternary", **kw)) as ctx:
        with pytest.raises(C.ChnsError):
            ctx.adj_forward(1, 1)          # no state yet
        ctx.set_real_state(None, np.full((32, 32), 0.3), np.full((32, 32), 0.3))
        with pytest.raises(C.ChnsError):
            ctx.adj_backward()             # not seeded
        ctx.adj_forward(2, 1)
        with pytest.raises(ValueError):
            ctx.adj_seed_spectrum(np.zeros(3))   # wrong length (checked by the binding)


def test_torch_autograd_function():
    """The wrapper a training loop would write around the C ABI (test harness only)."""
    torch = pytest.importorskip("torch")
    op, cp, kw, reals = adj_case(64)
    nsteps, K = 6, 4
    target = target_for(op, nsteps)

    class SpectrumLoss(torch.autograd.Function):
        @staticmethod
        def forward(fctx, c1, c2, sig):
            p = C.params("ternary", **{**kw, "sigma_12": float(sig[0]), "sigma_13": float(sig[1]),
                                       "sigma_23": float(sig[2])})
            ctx = C.Context(p)
            ctx.set_real_state(reals[0], c1.detach().numpy(), c2.detach().numpy())
            ctx.adj_forward(nsteps, K)
            loss, _ = ctx.adj_seed_spectrum(target)
            fctx.lib_ctx = ctx
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(fctx, gout):
            _, g1, g2, gs = fctx.lib_ctx.adj_backward()
            fctx.lib_ctx.close()
            return (gout * torch.from_numpy(g1), gout * torch.from_numpy(g2), gout * torch.from_numpy(gs))

    c1 = torch.tensor(reals[1], requires_grad=True)
    c2 = torch.tensor(reals[2], requires_grad=True)
    sig = torch.tensor([kw["sigma_12"], kw["sigma_13"], kw["sigma_23"]], dtype=torch.float64, requires_grad=True)
    loss = 3.0 * SpectrumLoss.apply(c1, c2, sig)
    loss.backward()
    ref_loss, grads, gsig, _ = OA.loss_and_grad(op, O.state_from_real(op, *reals), nsteps, target)
    assert abs(loss.item() - 3.0 * ref_loss) <= 1e-10 * 3.0 * ref_loss
    assert rel(c1.grad.numpy(), 3.0 * grads[1]) < TOL
    assert rel(c2.grad.numpy(), 3.0 * grads[2]) < TOL
    assert np.all(np.abs(sig.grad.numpy() - 3.0 * gsig) <= 1e-9 * np.abs(3.0 * gsig))


def test_reverse_step_2048():
    """One reverse step at 2048^2 (LDS-staged adjoint row passes) against the oracle."""
    N = 2048
    kw = dict(Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI, nu=2e-3, dt=1e-6, lam=5e-4, epsilon=1.5 * TWO_PI / N,
              sigma_12=0.1, sigma_13=0.08, sigma_23=0.06, Xlambda=2.0)
    op = O.Params(model="ternary", **kw)
    s0 = O.state_from_real(op, O.random_vorticity(op, 2), *O.init_lens_real(op))
    rng = np.random.default_rng(5)
    seeds = [O.truncation(op, O.dfft(rng.standard_normal((N, N)))) for _ in range(3)]
    ref, part = OA.step_vjp(op, s0, seeds)
    with C.Context(C.params("ternary", **kw)) as ctx:
        ctx.set_state(*s0.fields())
        ctx.adj_forward(1, 1)
        ctx.adj_seed(*seeds)
        ctx.adj_backward()
        got, gpart = ctx.adj_get()
    for a, b in zip(got, ref):
        assert rel(a, b) < TOL
    for k in OA.SIGMA_PARTIALS:
        assert abs(gpart[k] - part[k]) <= 1e-9 * abs(part[k]) + 1e-13 * max(map(abs, part.values())), k


@pytest.mark.parametrize("N", [64, 2048])
def test_stored_intermediates_match_recompute(N):
    """Keeping the replay's forward intermediates gives the same sweep as recomputing
    them (both modes of chns_adj_set_store): bitwise below 256^2; from 256^2 the
    standalone column pass A of the recompute uses 16-point radix stages where the
    fused pass A of the forward uses 8, so the last bits differ."""
    kw = dict(Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI, nu=2e-3, dt=1e-6 if N > 64 else 2e-5, lam=5e-4,
              epsilon=1.5 * TWO_PI / N, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06, Xlambda=2.0)
    op = O.Params(model="ternary", **kw)
    c1, c2 = O.init_lens_real(op)
    target = O.spectrum(op, O.state_from_real(op, None, c2, c1))[0]["sph1"]
    outs = []
    for store in (True, False):
        with C.Context(C.params("ternary", **kw)) as ctx:
            ctx.adj_set_store(store)
            ctx.set_real_state(O.random_vorticity(op, 2), c1, c2)
            ctx.adj_forward(7, 3)
            ctx.adj_seed_spectrum(target)
            outs.append(ctx.adj_backward())
    if N < 256:
        for a, b in zip(outs[0][:3], outs[1][:3]):
            assert np.array_equal(a, b)
        assert np.array_equal(outs[0][3], outs[1][3])
    else:
        for a, b in zip(outs[0][1:3], outs[1][1:3]):
            assert rel(a, b) < TOL
        # the omega gradient is a cancelling sum far below the c gradients here:
        # compare it on their scale
        scale = np.linalg.norm(outs[1][1]) + np.linalg.norm(outs[1][2])
        assert np.linalg.norm(outs[0][0] - outs[1][0]) < TOL * scale
        assert np.all(np.abs(outs[0][3] - outs[1][3]) <= TOL * np.abs(outs[1][3]))


@pytest.mark.parametrize("N,nsteps,K", [(16, 5, 2), (32, 7, 3), (64, 4, 4)])
def test_free_energy_loss_gradient_matches_oracle(N, nsteps, K):
    """Final free energy (of the truncated phase fields) as the loss: value,
    gradients for the initial fields and sigma (with its direct part) against
    the oracle adjoint (oracle/chns_adjoint.py, pinned by finite differences of
    the oracle forward in tests/test_adjoint_oracle.py)."""
    op, cp, _, reals = adj_case(N)
    loss, grads, gsig, _ = OA.loss_and_grad(op, O.state_from_real(op, *reals), nsteps, None)
    with C.Context(cp) as ctx:
        ctx.set_real_state(*reals)
        ctx.adj_forward(nsteps, K)
        gl, e1, e2 = ctx.adj_seed_free_energy()
        gw, g1, g2, gs = ctx.adj_backward()
        state = ctx.get_state()
    assert abs(gl - loss) <= 1e-12 * loss and abs(e1 + e2 - gl) <= 1e-15 * gl
    for a, b in zip((gw, g1, g2), grads):
        assert rel(a, b) < TOL
    assert np.all(np.abs(gs - gsig) <= 1e-9 * np.abs(gsig))
    ref = O.run(op, O.state_from_real(op, *reals), nsteps)
    for a, b in zip(state, ref.fields()):
        assert rel(a, b) < 1e-12                      # the seed leaves the state as it was


def gpu_fe_loss(cp, reals, nsteps):
    with C.Context(cp) as ctx:
        ctx.set_real_state(*reals)
        ctx.step(nsteps)
        ctx.adj_forward(0, 1)
        return ctx.adj_seed_free_energy()[0]


@pytest.mark.parametrize("N,nsteps,K", [(128, 12, 5)])
def test_free_energy_gradient_matches_central_differences(N, nsteps, K):
    """The library's free-energy gradient against central differences of the
    library forward, for c1, c2 and sigma_12/13/23."""
    from dataclasses import replace

    op, cp, kw, reals = adj_case(N)
    with C.Context(cp) as ctx:
        ctx.set_real_state(*reals)
        ctx.adj_forward(nsteps, K)
        ctx.adj_seed_free_energy()
        gw, g1, g2, gs = ctx.adj_backward()
    rng = np.random.default_rng(9)
    for k, g in ((1, g1), (2, g2)):
        d = rng.standard_normal((N, N))

        def f(h):
            moved = [x.copy() for x in reals]
            moved[k] += h * d
            return gpu_fe_loss(cp, moved, nsteps)

        fd = richardson(f, 1e-5)
        ad = float(np.sum(g * d))
        assert abs(fd - ad) <= TOL_FD * abs(ad), (k, fd, ad)
    for i, name in enumerate(("sigma_12", "sigma_13", "sigma_23")):
        def f(h):
            return gpu_fe_loss(C.params("ternary", **{**kw, name: kw[name] * (1.0 + h)}), reals, nsteps)

        fd = richardson(f, 1e-5) / kw[name]
        assert abs(fd - gs[i]) <= TOL_FD * abs(gs[i]), (name, fd, gs[i])
