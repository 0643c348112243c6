"""Ensembles (chns_ensemble_*): batched launches per (size, model) group and a
captured HIP graph per step.  Every instance must evolve exactly as a single
context does from the same state (same kernels, same arithmetic), including
across graph replays with eager work (diagnostics, state uploads, spectrum
truncation) in between, and the first instance is checked against the oracle.
"""
import math

import numpy as np
import pytest

import chns_amd as C
import oracle as O

pytestmark = pytest.mark.gpu

TOL_SAME = 1e-13     # batched vs single-instance launches of the same kernels


def member(model, N, seed, sig=0.1):
    """Parameters and a random initial state of one ensemble member."""
    L = 2 * math.pi
    rng = np.random.default_rng(seed)
    s = (0.05 + 0.1 * rng.random(3)) if model == "ternary" else np.full(3, sig)
    kw = dict(Nx=N, Ny=N, lx=L, ly=L, nu=5e-3, dt=5e-4, lam=1e-3, epsilon=1.5 * L / N,
              sigma=float(0.05 + 0.1 * rng.random()), sigma_12=float(s[0]), sigma_13=float(s[1]),
              sigma_23=float(s[2]), f_amp=1.0 if model == "ns" else 0.0, f_k=2 if model == "ns" else 0)
    op = O.Params(model=model, **kw)
    wz = O.random_vorticity(op, seed)
    if model == "ternary":
        c = O.init_noise_real(op, seed, 1.0 / 3.0, 0.05)
    elif model == "binary":
        c = O.init_noise_real(op, seed, 0.5, 0.05)
    else:
        c = ()
    return kw, op, O.state_from_real(op, wz, *c)


MEMBERS = [("ternary", 32, 1), ("binary", 32, 2), ("ternary", 32, 3), ("ternary", 64, 4), ("ns", 32, 5),
           ("binary", 64, 6), ("ternary", 32, 7), ("binary", 16, 8)]


def build(members):
    kws, ops, states = zip(*[member(*m) for m in members])
    ps = [C.params(m[0], **kw) for m, kw in zip(members, kws)]
    return ps, ops, list(states)


def single(p, s, nsteps, truncate_after=None):
    with C.Context(p) as ctx:
        ctx.set_state(*s.fields())
        if truncate_after is None:
            ctx.step(nsteps)
        else:
            ctx.step(truncate_after)
            ctx.truncate_state()
            ctx.step(nsteps - truncate_after)
        return ctx.get_state()


def close(a, b, tol):
    return all(O.rel_l2(x, y) <= tol for x, y in zip(a, b))


def test_groups_and_instances():
    ps, _, states = build(MEMBERS)
    with C.Ensemble(ps) as ens:
        assert len(ens) == len(MEMBERS)
        groups = ens.groups()
        assert sum(g[2] for g in groups) == len(MEMBERS)
        assert len(groups) == len({(m[1], m[0]) for m in MEMBERS})
        for b, s in enumerate(states):
            ens[b].set_state(*s.fields())
        for b, s in enumerate(states):
            assert close(ens[b].get_state(), s.fields(), 0.0)


def test_ensemble_matches_single_contexts_across_replays():
    ps, ops, states = build(MEMBERS)
    with C.Ensemble(ps) as ens:
        for b, s in enumerate(states):
            ens[b].set_state(*s.fields())
        ens.step(9)                           # one 8-step graph replay and one direct step
        ke = ens[0].kinetic_energy()          # eager work between graph replays
        ens[2].truncate_state()               # spectrum() side effect on one instance
        ens.step(8)
        ens.sync()
        got = [ens[b].get_state() for b in range(len(MEMBERS))]
    for b, (p, s) in enumerate(zip(ps, states)):
        ref = single(p, s, 17, truncate_after=9 if b == 2 else None)
        assert close(got[b], ref, TOL_SAME), b
    # member 0 against the oracle after 9 steps (kinetic energy) and 17 steps (state)
    s9 = O.run(ops[0], states[0], 9)
    assert abs(ke[0] - O.kinetic_energy(ops[0], s9)[0]) <= 1e-12 * abs(ke[0])
    assert close(got[0], O.run(ops[0], states[0], 17).fields(), 1e-10)


def test_state_upload_between_steps():
    ps, _, states = build(MEMBERS[:4])
    with C.Ensemble(ps) as ens:
        for b, s in enumerate(states):
            ens[b].set_state(*s.fields())
        ens.step(8)
        ens[1].set_state(*states[1].fields())   # restart instance 1: its row inputs are rebuilt
        ens.step(10)
        got = [ens[b].get_state() for b in range(4)]
    for b in range(4):
        ref = single(ps[b], states[b], 10 if b == 1 else 18)
        assert close(got[b], ref, TOL_SAME), b


def test_tune_restores_state_and_keeps_results():
    ps, _, states = build(MEMBERS)
    with C.Ensemble(ps) as ens:
        for b, s in enumerate(states):
            ens[b].set_state(*s.fields())
        ens.step(8)
        before = [ens[b].get_state() for b in range(len(MEMBERS))]
        splits = ens.tune(reps=2)
        assert len(splits) == len(ens.groups()) and all(s in (1, 2, 4) for s in splits)
        after = [ens[b].get_state() for b in range(len(MEMBERS))]
        ens.step(11)
        got = [ens[b].get_state() for b in range(len(MEMBERS))]
    for a, b in zip(before, after):
        assert close(a, b, 0.0)
    for b, (p, s) in enumerate(zip(ps, states)):
        assert close(got[b], single(p, s, 19), TOL_SAME), b


def test_ensemble_needs_every_state():
    ps, _, states = build(MEMBERS[:2])
    with C.Ensemble(ps) as ens:
        ens[0].set_state(*states[0].fields())
        with pytest.raises(C.ChnsError):
            ens.step(1)


def test_invalid_member_rejected():
    ps, _, _ = build(MEMBERS[:2])
    ps[1].Nx = 48   # not a power of two
    with pytest.raises(C.ChnsError):
        C.Ensemble(ps)


@pytest.mark.parametrize("tuned", [False, True])
def test_c4_workload(tuned):
    """The batched-ensemble workload of the benchmark (bench.ensemble_members():
    32 x 256^2 + 16 x 512^2 + 4 x 1024^2, binary and ternary alternating, each
    with its own seed and sigma): 3 direct steps, then 9 (one 8-step graph
    replay and one direct step; after tune() with its chosen launch splits),
    every member equal to a stand-alone context at 1e-13, and one member of
    each size against the oracle after 3 steps at 1e-10."""
    import bench

    members = bench.ensemble_members()
    assert len(members) == 52
    ps = [C.params(m, **kw) for m, kw, _ in members]
    print(f"C4 ensemble: {len(members)} members, tuned={tuned}", flush=True)
    with C.Ensemble(ps) as ens:
        assert sorted(ens.groups()) == [(8, 1, 16), (8, 2, 16), (9, 1, 8), (9, 2, 8), (10, 1, 2), (10, 2, 2)]
        for b, (m, kw, f) in enumerate(members):
            ens[b].set_real_state(None, *f)
        if tuned:
            splits = ens.tune(reps=1)
            assert all(s in (1, 2, 4) for s in splits)
            sched = ens.schedule()
            assert [s for s, _ in sched] == splits and all(0 <= c < 4 for _, c in sched)
            print("tuned (split, schedule) per group:", list(zip(ens.groups(), sched)), flush=True)
        ens.step(3)
        early = {b: ens[b].get_state() for b in (1, 32, 49)}   # ternary members of 256^2, 512^2, 1024^2
        ens.step(9)
        ens.sync()
        got = [ens[b].get_state() for b in range(len(members))]
        status = [ens[b].status() for b in range(len(members))]
    assert status == [0] * len(members)
    for b, (p, (m, kw, f)) in enumerate(zip(ps, members)):
        with C.Context(p) as ctx:
            ctx.set_real_state(None, *f)
            ctx.step(12)
            ref = ctx.get_state()
        assert close(got[b], ref, TOL_SAME), b
    for b, g in early.items():
        m, kw, f = members[b]
        op = O.Params(model=m, **kw)
        s3 = O.run(op, O.state_from_real(op, None, *f), 3)
        assert close(g, s3.fields(), 1e-10), b


SCHED_MEMBERS = [("ternary", 256, 11), ("binary", 512, 12), ("ternary", 1024, 13), ("ns", 256, 14),
                 ("binary", 64, 15), ("ternary", 512, 16)]


@pytest.mark.parametrize("sched", [0, 1, 2, 3])
def test_every_fft_schedule_matches_single_contexts(sched):
    """Each FFT schedule of the batched passes (bit 0: radix-16 column passes,
    bit 1: half the rows per row workgroup), set on every group with a split
    of 2 where it has two members, evolves each member as a stand-alone
    context does (1e-13: the schedules differ in FFT radix, i.e. round-off)."""
    ps, ops, states = build(SCHED_MEMBERS)
    # smaller dt at the larger grids (the explicit k^2 muF rate grows as N^3 at eps = 1.5 dx)
    for p in ps:
        p.dt = 5e-4 * min(1.0, (64.0 / p.Nx) ** 3)
    refs = [single(p, s, 9) for p, s in zip(ps, states)]
    with C.Ensemble(ps) as ens:
        for b, s in enumerate(states):
            ens[b].set_state(*s.fields())
        for g, (_, _, count) in enumerate(ens.groups()):
            ens.set_schedule(g, min(2, count), sched)
        assert all(sc == sched for _, sc in ens.schedule())
        ens.step(9)
        ens.sync()
        got = [ens[b].get_state() for b in range(len(ps))]
    for b in range(len(ps)):
        assert all(np.all(np.isfinite(x)) for x in got[b])
        assert close(got[b], refs[b], TOL_SAME), (b, sched)


def test_schedule_rejections():
    ps, _, _ = build(MEMBERS[:2])
    with C.Ensemble(ps) as ens:
        with pytest.raises(C.ChnsError):
            ens.set_schedule(0, 1, 4)
        with pytest.raises(C.ChnsError):
            ens.set_schedule(len(ens.groups()), 1, 0)
