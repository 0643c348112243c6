"""chns_free_energy_noalloc against chns_free_energy: the same outputs bit for
bit, on two calls in a row on one GPU, the second on a smaller grid (the two
share the static scratch array), and no device memory taken by it."""
import ctypes
import math

import numpy as np
import pytest

import chns_amd as C
import oracle as O

pytestmark = pytest.mark.gpu

TWO_PI = 2.0 * math.pi


def context(model, N):
    # dt shrinks as N^-2 with the grid-scale noise (the explicit k^2 muF term)
    kw = dict(Nx=N, Ny=N, lx=TWO_PI, ly=TWO_PI, nu=5e-3, dt=5e-4 * (32.0 / N) ** 2, lam=1e-3, epsilon=1.5 * TWO_PI / N,
              sigma=0.1, sigma_12=0.1, sigma_13=0.08, sigma_23=0.06)
    op = O.Params(model=model, **kw)
    nph = 2 if model == "ternary" else 1
    c = O.init_noise_real(op, N, 1 / 3 if nph == 2 else 0.5, 0.05)[:nph]
    ctx = C.Context(C.params(model, **kw))
    ctx.set_real_state(O.random_vorticity(op, 1), *c)
    ctx.step(3)
    return ctx


def free_bytes():
    """Free device memory as the HIP runtime that libchns_amd.so uses sees it
    (torch, if another test imported it, brings a second runtime of its own)."""
    C.load_library()
    with open("/proc/self/maps") as f:
        paths = sorted({ln.split()[-1] for ln in f if "libamdhip64" in ln and "/torch/" not in ln})
    assert paths, "no HIP runtime mapped"
    hip = ctypes.CDLL(paths[0])
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
    return f.value


@pytest.mark.parametrize("model", ["ternary", "binary"])
def test_noalloc_free_energy_matches_and_allocates_nothing(model):
    big, small = context(model, 2048), context(model, 64)
    try:
        before = free_bytes()
        a_big = big.free_energy_noalloc()       # the first calls of either entry point on these contexts
        a_small = small.free_energy_noalloc()   # second call, smaller input, same scratch array
        # nothing allocated (the runtime's own bookkeeping may move a page or two)
        assert before - free_bytes() < 4 << 20
        assert big.free_energy_noalloc() == a_big   # the array keeps the last call's partials: no stale reuse
        b_big = big.free_energy()                # the allocating entry point (its buffers appear now)
        b_small = small.free_energy()
        assert free_bytes() < before - 2 * 2048 * 1028 * 16   # it does allocate its two intermediates
        assert a_big == b_big and a_small == b_small
        assert big.free_energy_noalloc() == b_big
        # the state is untouched, and the next step is the same as without the call
        g = big.get_state()
        big.step(2)
        ref = context(model, 2048)
        ref.step(2)
        for x, y in zip(big.get_state(), ref.get_state()):
            assert np.all(np.isfinite(x)) and np.array_equal(x, y)
        ref.close()
        del g
    finally:
        big.close()
        small.close()


def test_noalloc_guard_and_errors():
    ctx = context("ternary", 32)
    try:
        x = (np.arange(32) * (TWO_PI / 32))[:, None] * np.ones((1, 32))
        ctx.set_real_state(None, 10.0 + 0.1 * np.cos(x), np.full((32, 32), 0.3))
        e, _ = ctx.free_energy_noalloc()
        assert e >= 100 and ctx.status() == C.CHNS_FLAG_ENERGY
    finally:
        ctx.close()
    ns = C.Context(C.params("ns", Nx=32, Ny=32))
    try:
        ns.set_real_state(np.zeros((32, 32)))
        with pytest.raises(C.ChnsError):
            ns.free_energy_noalloc()
    finally:
        ns.close()
