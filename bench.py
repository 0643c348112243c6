"""Benchmark: plain ETD2RK step throughput, ternary liquid lens 2048^2, FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Prints one JSON line.  ``value`` is grid-point steps/s (N^2 * steps/s) with
the state resident in HBM; the timed region is exactly K steps bracketed by
device synchronisations of the library's stream.  ``roofline`` reports the
dominant pass: algorithmic bytes per launch (the compulsory traffic of the
five-pass design, coefficient tables excluded) over its mean launch duration,
from HIP events recorded on the library's stream around every launch during K
further steps right after the timed region (events between launches cost
about 7% of the step, so the timed region itself carries none).
``cpu_baseline`` times the C/OpenMP restatement of the step (oracle/chns_step.c)
on the host cores (at most 16, one GPU's share) on a bounded sample of the same
workload, and records the CPU model.  ``roofline.traffic`` is the HBM bytes per launch of that
pass from two child ``rocprofv3 --pmc`` runs of this script (FETCH_SIZE, then
WRITE_SIZE, KiB per dispatch), each scaled by the factor that makes the same
counter read the true size of a plain 256 MiB device-to-device copy
(``--copy-probe``); null when rocprofv3 is unavailable or ``--no-traffic``.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ternary-chns_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# counter scale factors: a 256 MiB hipMemcpy D2D (__amd_rocclr_copyBuffer) reads
# FETCH_SIZE = 131084 KiB and WRITE_SIZE = 262144 KiB on MI355X (ROCm 7.2)
FETCH_SCALE = 2.0
WRITE_SCALE = 1.0


# After 10 ms or more of GPU idle (here: while the host builds the initial
# fields and uploads them) the first ~40 steps at 2048^2 run slow while the
# GPU's clocks ramp up: measured 0.717, 0.678, then 0.667 ms/step for the
# first three blocks of 20 steps after 5 warm-up steps, with or without graph
# replays, in a fresh context or after a 10-200 ms pause.  The warm-up
# therefore runs at least this many steps (~43 ms of load) whatever --warmup
# asks for; the JSON line reports both.
MIN_WARMUP = 64


def effective_warmup(requested: int) -> int:
    return max(requested, MIN_WARMUP)


def kernel_tag(pass_name: str, N: int) -> str:
    """Substring of the demangled kernel name of a pass (ternary model)."""
    L = int(round(math.log2(N)))
    # single-instance launches of the plain step: then the SMOOTH_FILTER switch
    # and the default FFT schedule (rows per workgroup 0, points per thread 8),
    # passes C and D with the conservative vorticity advection (true);
    # ensembles use chns::InstBatch
    row, col = "chns::Inst, false, 0>", "chns::Inst, false, 8>"
    rowc, colc = "chns::Inst, false, 0, true>", "chns::Inst, false, 8, true>"
    # pass B at 2048 is the row-pipelined kernel
    rowb = f"k_rowB_pipe<{L}, 2, chns::Inst>" if L == 11 else f"k_rowB<{L}, 2, {row}"
    return {"colA": f"k_colA<{L}, chns::Inst, false>", "rowB": rowb,
            "colC": f"k_colC<{L}, 2, {colc}", "rowD": f"k_rowD<{L}, 2, {rowc}",
            "colE1": f"k_colE<{L}, 2, 1, {col}", "colE2": f"k_colE<{L}, 2, 2, {col}"}[pass_name]


def measure_traffic(N: int, tag: str, timeout: int = 120):
    """Mean scaled FETCH_SIZE + WRITE_SIZE bytes per dispatch of the kernels whose
    name contains ``tag``, from two child rocprofv3 runs of a 3-step bench."""
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not found"
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="chns_pmc_")
        try:
            cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--",
                   sys.executable, os.path.abspath(__file__), "--steps", "3", "--warmup", "1",
                   "--N", str(N), "--no-cpu-baseline", "--no-traffic", "--adjoint-steps", "0"]
            try:
                subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=timeout)
            except subprocess.TimeoutExpired:
                return None, f"rocprofv3 --pmc {counter} timed out"
            vals = []
            for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
                with open(f) as fh:
                    for r in csv.DictReader(fh):
                        if tag in r["Kernel_Name"] and r["Counter_Name"] == counter:
                            vals.append(float(r["Counter_Value"]))
            if not vals:
                return None, f"no {counter} records for {tag}"
            res[counter] = sum(vals) / len(vals) * 1024.0
        finally:
            shutil.rmtree(d, ignore_errors=True)
    fetch = res["FETCH_SIZE"] * FETCH_SCALE
    write = res["WRITE_SIZE"] * WRITE_SCALE
    return {"fetch": fetch, "write": write, "total": fetch + write}, None


def workload(N=2048):
    L = 2.0 * math.pi
    # dt = 2.5e-4 diverges within 100 steps at 2048^2 for the reference scheme (the
    # explicit k^2 muF term), 2.5e-6 within 600; 1e-6 keeps the lens run finite
    # over 2000 steps.  Timing does not depend on dt.
    return dict(Nx=N, Ny=N, lx=L, ly=L, nu=2e-3, dt=1e-6, lam=5e-4, epsilon=1.5 * L / N,
                sigma_12=0.1, sigma_13=0.08, sigma_23=0.06, Xlambda=2.0)


def pass_bytes(N, Jt, K):
    """Algorithmic HBM bytes per launch of each pass (ternary)."""
    F = N * (N // 2 + 1) * 16
    Ft = N * Jt * 16
    Fk = K * 16
    return {
        "colA": 2 * Fk + 2 * Ft,
        "rowB": 2 * Ft + 2 * F,
        # the conservative vorticity advection: pass C writes T(w) instead of
        # T(i kx w), T(i ky w), and pass D reads it (one Ft less in each)
        "colC": 9 * Ft + 3 * Fk,
        "rowD": 9 * Ft + 4 * F,
        # stage 1 writes the intermediate state a only on its kept modes (the only
        # ones read before stage 2 overwrites it), and stage 2 reads the saved
        # a - z2 N(u) instead of the state: together with the above 6F + 4Ft - 3Fk
        # less per step than the five-pass budget B_alg = 42F + 48Ft + 6Fk, which
        # stays the reported figure
        "colE1": 12 * F + 2 * Ft + 3 * Fk,
        "colE2": 12 * F + 2 * Ft,
    }, 42 * F + 48 * Ft + 6 * Fk


def host_cpu():
    """CPU model (/proc/cpuinfo) and the core counts the OS reports."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_baseline(N, kw, warm=2, steps=150, max_seconds=30.0, threads=None, fields=None, what="lens"):
    """The C/OpenMP restatement of the step (oracle/chns_step.c: the same
    reference functions as the numpy oracle, its own radix-2 FFT, two real
    fields per complex transform) on the host cores: ``warm`` untimed steps,
    then up to ``steps`` timed ones (fewer if ``max_seconds`` runs out), on the
    workload of the GPU leg (``fields``: its real initial c1, c2; the lens by
    default).  ``threads``: the cores this process may use, at most 16 (one
    GPU's share of the host)."""
    import oracle as O
    from oracle import c_oracle as CO

    if threads is None:
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
    used = CO.threads(threads)
    op = O.Params(model="ternary", **kw)
    s0 = O.state_from_real(op, None, *(fields if fields is not None else O.init_lens_real(op)))
    st = CO.Stepper(op)
    try:
        st.set_state(s0)
        st.step(warm)
        done = 0
        t0 = time.perf_counter()
        while done < steps:
            st.step(1)
            done += 1
            if time.perf_counter() - t0 >= max_seconds:
                break
        dt = time.perf_counter() - t0
    finally:
        st.close()
    cpu = host_cpu()
    return {"value": N * N * done / dt, "unit": "grid-pt steps/s", "cores": used, "kind": "port",
            "cpu_model": cpu["model"], "nproc": cpu["nproc"], "affinity": cpu["affinity"],
            "sample": f"{done} ternary ETD2RK steps of the {N}^2 {what} workload after {warm} warm-up steps "
                      f"(C/OpenMP restatement oracle/chns_step.c, {used} threads), {dt:.1f} s"}


def forced_ternary_1024():
    """The forced ternary workload at 1024^2: c1, c2 = 1/3 + U[-0.05, 0.05]
    (SplitMix64 seed 2), all sigma_ij = 0.1, eps = 1.5 dx, lambda = 1e-3,
    nu = 5e-3, Xlambda = 2, Kolmogorov forcing 0.1 sin(4y) on omega; dt = 5e-4
    diverges at this size (the explicit k^2 muF rate grows as N^3 at eps = 1.5 dx),
    so dt = 5e-4 (256/1024)^3; the timing does not depend on dt."""
    from chns_amd import ic

    N, L = 1024, 2.0 * math.pi
    kw = dict(Nx=N, Ny=N, lx=L, ly=L, nu=5e-3, dt=5e-4 * (256.0 / N) ** 3, lam=1e-3, epsilon=1.5 * L / N,
              sigma_12=0.1, sigma_13=0.1, sigma_23=0.1, Xlambda=2.0, f_amp=0.1, f_k=4)
    return kw, ic.noise(N, N, 2, 1.0 / 3.0, 0.05, 2)


def forced_ternary_leg(steps=200, warmup=64, cpu=True):
    """Step throughput of the forced ternary 1024^2 workload on the GPU (after
    its own warm-up) and, with ``cpu``, the C/OpenMP restatement beside it."""
    import chns_amd as C

    kw, (c1, c2) = forced_ternary_1024()
    N = kw["Nx"]
    with C.Context(C.params("ternary", **kw), device=0) as ctx:
        ctx.set_real_state(None, c1, c2)
        ctx.step(warmup)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(steps)
        ctx.sync()
        dt = time.perf_counter() - t0
    out = {"workload": f"ternary {N}^2, c = 1/3 + U[-0.05, 0.05] (seed 2), all sigma_ij = 0.1, Xlambda = 2, "
                       f"forcing 0.1 sin(4y), dt = {kw['dt']:.3g}",
           "steps": steps, "ms_per_step": dt * 1e3 / steps, "value": N * N * steps / dt, "unit": "grid-pt steps/s"}
    if cpu:
        out["cpu_baseline"] = cpu_baseline(N, kw, steps=100, max_seconds=15.0, fields=(c1, c2),
                                           what="forced ternary")
    return out


def ensemble_members():
    """The batched-ensemble workload: 32 x 256^2 + 16 x 512^2 + 4 x 1024^2 instances;
    within each size even b is binary (c = 1/2 + U[-0.05, 0.05]) and odd b ternary
    (c1, c2 = 1/3 + U[-0.05, 0.05]); sigma (each sigma_ij) ~ U[0.05, 0.15] from
    SplitMix64 seeded 1000 + b, the noise seeded 2000 + b; epsilon = 1.5 dx, lambda = 1e-3,
    nu = 5e-3, dt = 5e-4 (256/N)^3 (ternary: times (min |gamma_i| / 0.1)^2 when below 0.1),
    L = 2 pi, omega_0 = 0."""

    from chns_amd import ic

    out = []
    b = 0
    for N, count in ((256, 32), (512, 16), (1024, 4)):
        for k in range(count):
            model = "binary" if k % 2 == 0 else "ternary"
            L = 2.0 * math.pi
            sig = 0.05 + 0.1 * ic._splitmix64(1000 + b, 3)
            # dt = 5e-4 at 256^2; the explicit k^2 muF rate grows like N^3 at epsilon = 1.5 dx
            # and 5e-4 diverges at 512^2 and 1024^2, so dt shrinks as (256/N)^3; a ternary
            # member's rate also grows as its smallest |gamma_i| shrinks (sigma ~ U[0.05, 0.15]
            # gives gamma_i down to ~0), so its dt shrinks as (min |gamma_i| / 0.1)^2 below 0.1
            dt = 5e-4 * (256.0 / N) ** 3
            if model == "ternary":
                s12, s13, s23 = (float(x) for x in sig)
                gmin = min(abs(s12 - s23 + s13), abs(s12 - s13 + s23), abs(s23 - s12 + s13))
                dt *= min(1.0, max(gmin, 1e-3) / 0.1) ** 2
            kw = dict(Nx=N, Ny=N, lx=L, ly=L, nu=5e-3, dt=dt, lam=1e-3, epsilon=1.5 * L / N,
                      sigma=float(sig[0]), sigma_12=float(sig[0]), sigma_13=float(sig[1]), sigma_23=float(sig[2]))
            nf = 1 if model == "binary" else 2
            fields = ic.noise(N, N, 2000 + b, 0.5 if model == "binary" else 1.0 / 3.0, 0.05, nf)
            out.append((model, kw, fields))
            b += 1
    return out


def ensemble_bench(steps: int, warmup: int, legs: str = "both"):
    """Speedup of one graph-launched batched ensemble over the same instances run
    one after another, each alone on the GPU (sum of their timed steps).
    legs="batch" runs the batched leg only (no per-member contexts are created
    and destroyed; used for profiling runs)."""
    import numpy as np

    import chns_amd as C

    members = ensemble_members()
    ps = [C.params(m, **kw) for m, kw, _ in members]
    seq = 0.0
    for p, (m, kw, f) in zip(ps, members) if legs == "both" else ():
        with C.Context(p) as ctx:
            ctx.set_real_state(None, *f)
            ctx.step(effective_warmup(warmup))
            ctx.sync()
            t0 = time.perf_counter()
            ctx.step(steps)
            ctx.sync()
            seq += time.perf_counter() - t0
    res = {}
    with C.Ensemble(ps) as ens:
        for b, (m, kw, f) in enumerate(members):
            ens[b].set_real_state(None, *f)
        for tuned in (False, True):
            splits = ens.tune(reps=3) if tuned else None
            ens.step(effective_warmup(warmup))
            ens.sync()
            t0 = time.perf_counter()
            ens.step(steps)
            ens.sync()
            res[tuned] = (time.perf_counter() - t0, splits, ens.schedule())
        finite = all(bool(np.all(np.isfinite(x))) for b in range(len(ens)) for x in ens[b].get_state())
        groups = ens.groups()
    best = min(res, key=lambda k: res[k][0])
    batch = res[best][0]
    pts = sum(kw["Nx"] * kw["Ny"] for _, kw, _ in members)
    return {
        "metric": "batched ensemble speedup over sequential single-instance runs",
        "value": seq / batch if legs == "both" else None,
        "unit": "x",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": batch * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (SplitMix64 noise, seeds 1000+b / 2000+b)",
        "config": {"workload": "ensemble 32x256^2 + 16x512^2 + 4x1024^2, binary/ternary alternating, "
                               "sigma ~ U[0.05,0.15], dt=5e-4*(256/N)^3 (ternary x (min|gamma|/0.1)^2), HIP graph replays of 8 steps",
                   "instances": len(members), "groups": groups, "parallelism": "single GPU"},
        "graph_replays": os.environ.get("CHNS_NO_GRAPH", "") != "1",
        "sequential_s": seq,
        "batch_s": {"untuned": res[False][0], "tuned": res[True][0]},
        "tuned_splits": res[True][1],
        "tuned_schedules": res[True][2],
        "grid_pt_steps_per_s": pts * steps / batch,
        "finite": finite,
    }


def adjoint_bench(N: int, nsteps: int, ckpt_every: int):
    """Adjoint/forward time ratio of the discrete adjoint on the bench workload:
    forward nsteps with a checkpoint every ckpt_every steps, the c1 shell-spectrum
    loss against a target from a perturbed run, and the reverse sweep (segment
    replay + two vector-Jacobian products per step).  Both legs are timed with
    the state resident in HBM and bracketed by device synchronisations."""
    import numpy as np

    import chns_amd as C
    from chns_amd import ic

    kw = workload(N)
    p = C.params("ternary", **kw)
    c1, c2 = ic.lens(N, N, kw["lx"], kw["ly"], kw["epsilon"])
    with C.Context(p, device=0) as ctx:
        # target: the spectrum of a run from lenses of radius 0.45 instead of 0.4
        t1, t2 = ic.lens(N, N, kw["lx"], kw["ly"], kw["epsilon"], radius=0.45)
        ctx.set_real_state(None, t1, t2)
        ctx.step(nsteps)
        target = ctx.spectrum(truncate_state=False)["sph1"]
        ctx.set_real_state(None, c1, c2)
        ctx.adj_forward(min(nsteps, 2 * ckpt_every), ckpt_every)   # warm-up (allocations, code objects)
        ctx.adj_seed_spectrum(target)
        ctx.adj_backward()
        ctx.set_real_state(None, c1, c2)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.adj_forward(nsteps, ckpt_every)
        ctx.sync()
        t1 = time.perf_counter()
        loss, _ = ctx.adj_seed_spectrum(target)
        t2 = time.perf_counter()
        gw, g1, g2, gs = ctx.adj_backward()
        t3 = time.perf_counter()
        ctx.adj_release()
    fwd, bwd = t1 - t0, t3 - t2
    return {"workload": f"ternary lens {N}^2, {nsteps} steps, checkpoint every {ckpt_every}, "
                        "loss = sum_k (S11(k) - S_target(k))^2",
            "forward_s": fwd, "backward_s": bwd, "adjoint_fwd_ratio": bwd / fwd, "loss": loss,
            "grad_sigma": gs.tolist(),
            "finite": bool(np.all(np.isfinite(g1)) and np.all(np.isfinite(g2)) and np.all(np.isfinite(gs)))}


def copy_probe(nbytes: int, reps: int = 4):
    """Device-to-device copies of a known size (hipMemcpy) for calibrating the
    FETCH_SIZE / WRITE_SIZE counters against a plain copy kernel."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-probe", type=int, default=0, help="only run D2D copies of this many bytes")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU baselines (and the forced ternary 1024^2 leg they are timed beside)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter runs")
    ap.add_argument("--traffic-pass", default="rowD", help="pass whose HBM traffic is counted")
    ap.add_argument("--ensemble", action="store_true", help="measure the batched-ensemble workload instead")
    ap.add_argument("--ensemble-legs", choices=("both", "batch"), default="both",
                    help="ensemble: both legs (the speedup), or the batched leg only")
    ap.add_argument("--adjoint-steps", type=int, default=2000, help="adjoint leg: forward steps (0: skip)")
    ap.add_argument("--adjoint-ckpt", type=int, default=100, help="adjoint leg: checkpoint interval")
    a = ap.parse_args()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.ensemble:
        if a.gpus != 1:
            print("bench.py: the ensemble runs on one GPU", file=sys.stderr)
            sys.exit(2)
        print(json.dumps(ensemble_bench(a.steps if a.steps != 200 else 500, a.warmup, a.ensemble_legs)))
        return
    if a.gpus != 1:
        print(f"bench.py: measures one GPU only (got --gpus {a.gpus}); "
              "a single 2048^2 grid is not sharded across GPUs here", file=sys.stderr)
        sys.exit(2)

    N = a.N
    # counter runs first, in child processes, before this process touches the GPU
    traffic, traffic_err = (None, "skipped") if a.no_traffic else measure_traffic(N, kernel_tag(a.traffic_pass, N))

    import numpy as np

    import chns_amd as C
    from chns_amd import ic

    kw = workload(N)
    p = C.params("ternary", **kw)
    amask = C.dealias_extent(p)
    Jt = int(np.max(np.where(amask >= 0)[0])) + 1
    Kkept = int(np.sum(np.where(amask >= 0, 2 * amask + 1 - (amask >= N // 2), 0)))
    c1, c2 = ic.lens(N, N, kw["lx"], kw["ly"], kw["epsilon"])
    warm = effective_warmup(a.warmup)
    with C.Context(p, device=0) as ctx:
        ctx.set_real_state(None, c1, c2)
        ctx.step(warm)
        ctx.sync()
        # timed region: exactly K steps, no events between the launches
        t0 = time.perf_counter()
        ctx.step(a.steps)
        ctx.sync()
        t1 = time.perf_counter()
        # per-pass kernel durations: HIP events around every launch on the
        # library's stream over K more steps of the same run
        ctx.timing_read()
        ctx.timing(True)
        ctx.step(a.steps)
        tm = ctx.timing_read()
        ctx.timing(False)
        # host boundary: download and upload of the three fields (PCIe + layout transpose)
        t2 = time.perf_counter()
        state = ctx.get_state()
        t3 = time.perf_counter()
        ctx.set_state(*state)
        ctx.sync()
        t4 = time.perf_counter()
    finite = all(bool(np.all(np.isfinite(f))) for f in state)
    ms_per_step = (t1 - t0) * 1e3 / a.steps
    sps = a.steps / (t1 - t0)
    pb, balg = pass_bytes(N, Jt, Kkept)
    dom = a.traffic_pass if tm.get(a.traffic_pass, (0, 0))[1] > 0 else \
        max((k for k in tm if tm[k][1] > 0), key=lambda k: tm[k][0])
    avg_ms = tm[dom][0] / tm[dom][1]
    achieved = pb[dom] / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": "grid-pt ETD2RK steps/s, ternary 2048^2 FP64",
        "value": N * N * sps,
        "unit": "grid-pt steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "warmup_steps_run": warm,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (liquid-lens IC, deterministic)",
        "config": {"workload": f"ternary CHNS liquid lens {N}^2, ETD2RK, dt=1e-6, circular N/4 dealias",
                   "N": N, "model": "ternary", "parallelism": "single GPU"},
        "steps_per_s": sps,
        "hbm_frac_step": balg * sps / (HBM_PEAK_GBS * 1e9),
        "alg_bytes_per_step": balg,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["total"] if traffic and dom == a.traffic_pass else None,
                     "alg_bytes_per_launch": pb[dom], "avg_launch_ms": avg_ms,
                     "traffic_fetch_write": [traffic["fetch"], traffic["write"]] if traffic else None,
                     "traffic_note": traffic_err or f"{a.traffic_pass}: rocprofv3 --pmc, counters scaled "
                                                    f"{FETCH_SCALE:g}/{WRITE_SCALE:g} by a D2D copy"},
        "pass_ms_per_step": {k: tm[k][0] / a.steps for k in tm},
        "kernel_ms_per_step": sum(tm[k][0] for k in tm) / a.steps,
        "host_boundary_ms": {"get_state": (t3 - t2) * 1e3, "set_state": (t4 - t3) * 1e3},
        "finite": finite,
    }
    if a.adjoint_steps > 0:
        out["adjoint"] = adjoint_bench(N, a.adjoint_steps, a.adjoint_ckpt)
    if not a.no_cpu_baseline:
        if N == 2048:
            out["forced_ternary_1024"] = forced_ternary_leg()
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(N, kw)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
