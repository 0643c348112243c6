"""CPU checks of the C ABI library: it loads, exports every declared symbol,
and its host-built tables match the oracle bit for bit (no device calls)."""
import math
import os
import re

import numpy as np
import pytest

import chns_amd as C
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "chns.h")).read()
    return sorted(set(re.findall(r"\b(chns_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = C.load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert set(declared_symbols()) == set(C.EXPORTS)


def test_reference_named_library_exports_its_header():
    """libchns_ref.so exports every function include/chns_ref.h declares: the
    reference's etd2rk, free_energy, kinetic_energy, perimeter, spectrum with
    their names, and the chns_ref_* set-up calls."""
    import ctypes
    txt = open(os.path.join(ROOT, "include", "chns_ref.h")).read()
    names = sorted(set(re.findall(r"^(?:int|void)\s+([a-z0-9_]+)\s*\(", txt, re.M)))
    assert {"etd2rk", "free_energy", "kinetic_energy", "perimeter", "spectrum"} <= set(names)
    C.load_library()
    lib = ctypes.CDLL(os.path.join(os.path.dirname(C.LIB_PATH), "libchns_ref.so"))
    for name in names:
        assert hasattr(lib, name), name


CASES = [
    ("ns", 64, 2 * math.pi), ("binary", 128, 2 * math.pi), ("ternary", 128, 4 * math.pi),
    ("ternary", 2048, 2 * math.pi), ("binary", 512, 4 * math.pi),
]


@pytest.mark.parametrize("model,N,L", CASES)
def test_etd_tables_bit_exact(model, N, L):
    kw = dict(Nx=N, Ny=N, lx=L, ly=L, nu=2e-3, dt=2.5e-4, lam=5e-4, epsilon=1.5 * L / N, sigma=0.1,
              sigma_12=0.1, sigma_13=0.08, sigma_23=0.06)
    t = C.etd_tables(C.params(model, **kw))
    o = O.etd_tables(O.Params(model=model, **kw))
    for k in o:
        assert np.array_equal(t[k].view(np.uint64), o[k].view(np.uint64)), k


@pytest.mark.parametrize("model", ["ns", "binary", "ternary"])
def test_etd_tables_reference_defaults(model):
    t = C.etd_tables(C.params(model))
    o = O.etd_tables(O.Params(model=model))
    for k in o:
        assert np.array_equal(t[k].view(np.uint64), o[k].view(np.uint64)), k


@pytest.mark.parametrize("rule", [0, 1, 2])
@pytest.mark.parametrize("N,L", [(64, 2 * math.pi), (128, 4 * math.pi), (2048, 2 * math.pi), (256, 3.0)])
def test_dealias_extent_matches_mask(N, L, rule):
    a = C.dealias_extent(C.params("ternary", Nx=N, Ny=N, lx=L, ly=L, dealias=rule))
    m = O.dealias_mask(O.Params(Nx=N, Ny=N, lx=L, ly=L, dealias=rule))
    for j in range(N // 2 + 1):
        kept = np.where(m[:, j])[0]
        expect = -1 if kept.size == 0 else int(np.minimum(kept, N - kept).max())
        assert a[j] == expect


def test_params_struct_mirror_has_the_c_layout():
    """CParams mirrors chns_params field for field: the defaults written by the
    library land in the fields of the same names (dealias, the last field,
    included)."""
    p = C.params("binary")
    assert (p.Nx, p.model, p.sigma, p.lambda_, p.rho1, p.g, p.vpm, p.sponge, p.dealias) == \
        (1024, 1, 2.0, 1e-6, 1.044, 3.0, 0, 0, 0)
    assert C.params("ternary", dealias=2).dealias == 2


def test_invalid_params_rejected():
    lib = C.load_library()
    for bad in (dict(dealias=3), dict(dealias=1, smooth_filter=1)):
        p = C.params("ternary", Nx=64, Ny=64, **bad)
        assert lib.chns_dealias_extent(C.ctypes.byref(p), np.zeros(33, np.int32).ctypes.data_as(
            C.ctypes.POINTER(C.ctypes.c_int))) == C.CHNS_E_ARG
    p = C.params("ternary", Nx=100, Ny=100)
    out = np.zeros(100 * 51)
    assert lib.chns_etd_tables(C.ctypes.byref(p), *[C._dp(out)] * 6) == C.CHNS_E_ARG
    p = C.params("ternary", Nx=64, Ny=32)
    assert lib.chns_dealias_extent(C.ctypes.byref(p), np.zeros(17, np.int32).ctypes.data_as(
        C.ctypes.POINTER(C.ctypes.c_int))) == C.CHNS_E_ARG


def test_bench_kernel_tags_name_built_kernels():
    """Each pass tag bench.py looks for in rocprofv3 output names exactly one
    kernel of the library (the plain single-instance step at 2048^2)."""
    import shutil
    import subprocess
    import bench
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not found")
    out = subprocess.run([nm, "-C", C.LIB_PATH], capture_output=True, text=True, check=True).stdout
    kernels = {ln.split("__device_stub__", 1)[1] for ln in out.splitlines() if "__device_stub__" in ln}
    for p in ("rowB", "colC", "rowD", "colE1", "colE2"):
        tag = bench.kernel_tag(p, 2048)
        hits = [k for k in kernels if tag in k]
        assert len(hits) == 1, (p, tag, hits)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with a message, before any device call."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
