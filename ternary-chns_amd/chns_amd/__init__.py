"""Python host side of the MI355X CHNS-2D hot path (ctypes over libchns_amd.so).

The library is the product: every numeric operation below runs in the HIP
kernels behind the C ABI declared in ``include/chns.h``.  This module only
moves arrays across the boundary and mirrors the reference's entry points
(``etd2rk``, ``truncation`` of the state as ``spectrum`` does, the
``INIT_READ_DATA`` initialisation) for Python callers and tests.  If the
shared library is missing the import fails loudly; there is no fallback.

Array conventions are the reference's: complex half-spectra ``[Nx][Ny/2+1]``
(``IDXC``, global.h:13) as numpy complex128, real fields ``[Nx][Ny]`` float64.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libchns_amd.so")

MODEL_NS, MODEL_BINARY, MODEL_TERNARY = 0, 1, 2
_MODELS = {"ns": MODEL_NS, "binary": MODEL_BINARY, "ternary": MODEL_TERNARY}

CHNS_OK, CHNS_E_ARG, CHNS_E_HIP, CHNS_E_STATE, CHNS_E_DIVERGED = 0, -1, -2, -3, -4
CHNS_FLAG_NONFINITE, CHNS_FLAG_ENERGY = 1, 2

EXPORTS = (
    "chns_params_default", "chns_create", "chns_destroy", "chns_last_error", "chns_set_state",
    "chns_get_state", "chns_set_real_state", "chns_step", "chns_sync", "chns_truncate_state",
    "chns_etd_tables", "chns_dealias_extent", "chns_filter_table", "chns_timing_enable", "chns_timing_read",
    "chns_free_energy", "chns_kinetic_energy", "chns_perimeter", "chns_spectrum", "chns_real_field",
    "chns_diag_sizes", "chns_ensemble_create", "chns_ensemble_destroy", "chns_ensemble_last_error",
    "chns_ensemble_size", "chns_ensemble_instance", "chns_ensemble_step", "chns_ensemble_sync",
    "chns_ensemble_tune", "chns_ensemble_groups", "chns_adj_forward", "chns_adj_seed_spectrum", "chns_adj_seed",
    "chns_adj_backward", "chns_adj_get", "chns_adj_release", "chns_adj_set_store", "chns_graphs_enable",
    "chns_status", "chns_free_energy_noalloc", "chns_adj_seed_free_energy", "chns_ensemble_schedule",
    "chns_ensemble_set_schedule", "chns_kinetic_energy_counts",
)
# coefficient partial sums behind the sigma gradient (chns_adj_get)
ADJ_PARTIALS = ("fac", "G1", "G2", "G3", "g1i", "g2i", "g3i", "mc1", "mc2",
                "A11", "A12", "A21", "A22", "lg1", "lg2")
PASS_NAMES = ("colA", "rowB", "colC", "rowD", "colE1", "colE2")


class CParams(ctypes.Structure):
    """Mirror of ``chns_params`` (include/chns.h)."""

    _fields_ = [
        ("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("model", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("nu", ctypes.c_double), ("dt", ctypes.c_double),
        ("sigma", ctypes.c_double), ("epsilon", ctypes.c_double), ("lambda_", ctypes.c_double),
        ("sigma_12", ctypes.c_double), ("sigma_13", ctypes.c_double), ("sigma_23", ctypes.c_double),
        ("Xlambda", ctypes.c_double),
        ("smooth_filter", ctypes.c_int),
        ("f_amp", ctypes.c_double), ("f_k", ctypes.c_int),
        ("gravity", ctypes.c_int),
        ("rho1", ctypes.c_double), ("rho2", ctypes.c_double), ("rho3", ctypes.c_double), ("g", ctypes.c_double),
        ("vpm", ctypes.c_int), ("sponge", ctypes.c_int),
        ("dealias", ctypes.c_int),
    ]


class ChnsError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libchns_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ChnsError(f"{path} is missing: run `python ternary-chns_amd/build.py` (no CPU fallback exists)")
    lib = ctypes.CDLL(path)
    P, D, I = ctypes.POINTER(CParams), ctypes.POINTER(ctypes.c_double), ctypes.c_int
    V = ctypes.c_void_p
    lib.chns_params_default.argtypes = [P, I]
    lib.chns_params_default.restype = None
    lib.chns_create.argtypes = [ctypes.POINTER(V), P, I]
    lib.chns_destroy.argtypes = [V]
    lib.chns_destroy.restype = None
    lib.chns_last_error.argtypes = [V]
    lib.chns_last_error.restype = ctypes.c_char_p
    lib.chns_set_state.argtypes = [V, D, D, D]
    lib.chns_get_state.argtypes = [V, D, D, D]
    lib.chns_set_real_state.argtypes = [V, D, D, D]
    lib.chns_step.argtypes = [V, I]
    lib.chns_sync.argtypes = [V]
    lib.chns_truncate_state.argtypes = [V]
    lib.chns_etd_tables.argtypes = [P, D, D, D, D, D, D]
    lib.chns_dealias_extent.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
    lib.chns_filter_table.argtypes = [P, D]
    lib.chns_timing_enable.argtypes = [V, I]
    lib.chns_timing_read.argtypes = [V, D, ctypes.POINTER(ctypes.c_int)]
    lib.chns_ensemble_create.argtypes = [ctypes.POINTER(V), P, I, I]
    lib.chns_ensemble_destroy.argtypes = [V]
    lib.chns_ensemble_destroy.restype = None
    lib.chns_ensemble_last_error.argtypes = [V]
    lib.chns_ensemble_last_error.restype = ctypes.c_char_p
    lib.chns_ensemble_size.argtypes = [V]
    lib.chns_ensemble_instance.argtypes = [V, I]
    lib.chns_ensemble_instance.restype = V
    lib.chns_ensemble_step.argtypes = [V, I]
    lib.chns_ensemble_sync.argtypes = [V]
    lib.chns_ensemble_tune.argtypes = [V, I, ctypes.POINTER(I)]
    lib.chns_ensemble_groups.argtypes = [V, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I)]
    lib.chns_free_energy.argtypes = [V, D, D]
    lib.chns_kinetic_energy.argtypes = [V, D, D, D]
    lib.chns_kinetic_energy_counts.argtypes = [V, D, D, D, D, D]
    lib.chns_perimeter.argtypes = [V, D]
    lib.chns_spectrum.argtypes = [V, I, D, D, D, D, D, D, D]
    lib.chns_real_field.argtypes = [V, I, I, D]
    lib.chns_diag_sizes.argtypes = [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.chns_adj_forward.argtypes = [V, I, I]
    lib.chns_adj_seed_spectrum.argtypes = [V, D, D, D]
    lib.chns_adj_seed.argtypes = [V, D, D, D]
    lib.chns_adj_backward.argtypes = [V, D, D, D, D]
    lib.chns_adj_get.argtypes = [V, D, D, D, D]
    lib.chns_adj_release.argtypes = [V]
    lib.chns_adj_set_store.argtypes = [V, I]
    lib.chns_graphs_enable.argtypes = [V, I]
    lib.chns_status.argtypes = [V, ctypes.POINTER(I)]
    lib.chns_free_energy_noalloc.argtypes = [V, D, D]
    lib.chns_adj_seed_free_energy.argtypes = [V, D, D, D]
    lib.chns_ensemble_schedule.argtypes = [V, ctypes.POINTER(I), ctypes.POINTER(I)]
    lib.chns_ensemble_set_schedule.argtypes = [V, I, I, I]
    _lib = lib
    return lib


def params(model="ternary", **kw) -> CParams:
    """chns_params with the reference defaults for ``model``, then overrides.

    Keyword ``lam`` (or ``lambda_``) is the mobility ``lambda``; ``epsilon``
    None or <= 0 selects 3*dx.
    """
    lib = load_library()
    m = _MODELS[model] if isinstance(model, str) else int(model)
    p = CParams()
    lib.chns_params_default(ctypes.byref(p), m)
    for k, v in kw.items():
        if k == "lam":
            k = "lambda_"
        if k == "epsilon" and v is None:
            v = -1.0
        if k == "model":
            v = _MODELS[v] if isinstance(v, str) else int(v)
        setattr(p, k, v)
    return p


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _nphase(model):
    return {MODEL_NS: 0, MODEL_BINARY: 1, MODEL_TERNARY: 2}[model]


class Context:
    """A device-resident CHNS state (one ``chns_ctx``)."""

    def __init__(self, p: CParams, device: int = 0, _handle=None):
        self.lib = load_library()
        self.p = p
        self.Nx, self.Ny = p.Nx, p.Ny
        self.Ny2 = p.Ny // 2 + 1
        self.nph = _nphase(p.model)
        self._owned = _handle is None
        if _handle is not None:          # an ensemble's instance (not owned)
            self.h = _handle
            return
        h = ctypes.c_void_p()
        rc = self.lib.chns_create(ctypes.byref(h), ctypes.byref(p), device)
        if rc != CHNS_OK:
            raise ChnsError(f"chns_create failed ({rc})")
        self.h = h

    def _check(self, rc, what):
        if rc != CHNS_OK:
            msg = self.lib.chns_last_error(self.h).decode()
            raise ChnsError(f"{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            if self._owned:
                self.lib.chns_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state ------------------------------------------------------------
    def _cplx(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        if a.shape != (self.Nx, self.Ny2):
            raise ValueError(f"expected shape {(self.Nx, self.Ny2)}, got {a.shape}")
        return a.view(np.float64)

    def set_state(self, cwz, cphi1=None, cphi2=None):
        arrs = [self._cplx(cwz)]
        arrs += [self._cplx(cphi1) if self.nph >= 1 else None, self._cplx(cphi2) if self.nph >= 2 else None]
        self._check(self.lib.chns_set_state(self.h, *[_dp(a) for a in arrs]), "chns_set_state")

    def get_state(self):
        outs = [np.empty((self.Nx, self.Ny2), np.complex128) for _ in range(1 + self.nph)]
        ptrs = [_dp(o.view(np.float64)) for o in outs] + [None] * (2 - self.nph)
        self._check(self.lib.chns_get_state(self.h, *ptrs), "chns_get_state")
        return outs

    def set_real_state(self, wz, phi1=None, phi2=None):
        def real(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.shape != (self.Nx, self.Ny):
                raise ValueError("real fields are [Nx][Ny]")
            return a
        arrs = [real(wz), real(phi1) if self.nph >= 1 else None, real(phi2) if self.nph >= 2 else None]
        self._check(self.lib.chns_set_real_state(self.h, *[_dp(a) for a in arrs]), "chns_set_real_state")

    # -- stepping ---------------------------------------------------------
    def step(self, nsteps: int = 1):
        self._check(self.lib.chns_step(self.h, int(nsteps)), "chns_step")

    def sync(self):
        self._check(self.lib.chns_sync(self.h), "chns_sync")

    def truncate_state(self):
        self._check(self.lib.chns_truncate_state(self.h), "chns_truncate_state")

    def graphs(self, enable: bool = True):
        """HIP graph replays of step() on (default) or off (every kernel launched directly)."""
        self._check(self.lib.chns_graphs_enable(self.h, int(bool(enable))), "chns_graphs_enable")

    def status(self):
        """Divergence flags (CHNS_FLAG_NONFINITE | CHNS_FLAG_ENERGY), 0 when healthy; waits
        for the queued work.  Does not raise on divergence: the caller decides, as the
        reference's MPI_Abort at free_energy.c:151-154 did."""
        f = ctypes.c_int()
        rc = self.lib.chns_status(self.h, ctypes.byref(f))
        if rc not in (CHNS_OK, CHNS_E_DIVERGED):
            self._check(rc, "chns_status")
        return f.value

    # -- diagnostics (free_energy.c, kinetic_energy.c, perimeter.c, spectrum.c)
    def free_energy(self):
        """(ephi, egphi) of the untruncated phase fields (free_energy.c:13-157)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.chns_free_energy(self.h, ctypes.byref(a), ctypes.byref(b)), "chns_free_energy")
        return a.value, b.value

    def free_energy_noalloc(self):
        """free_energy() with no device allocation (static scratch, step buffers reused)."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.chns_free_energy_noalloc(self.h, ctypes.byref(a), ctypes.byref(b)),
                    "chns_free_energy_noalloc")
        return a.value, b.value

    def kinetic_energy(self):
        """(eu,), (eu, eph) or (eu, eph1, eph2) as kinetic_energy.c prints them."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.chns_kinetic_energy(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
                    "chns_kinetic_energy")
        return (a.value, b.value, c.value)[: 1 + self.nph]

    def kinetic_energy_counts(self):
        """(kinetic_energy(), [n1[, n2]]): the energies and the per-phase point
        counts (c_i >= 0.5) they are averaged over (kinetic_energy.c:62-79)."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n1, n2 = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.chns_kinetic_energy_counts(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                                        ctypes.byref(n1), ctypes.byref(n2)),
                    "chns_kinetic_energy_counts")
        return (a.value, b.value, c.value)[: 1 + self.nph], [int(n1.value), int(n2.value)][: self.nph]

    def perimeter(self):
        """[(interface, inside) per phase field] (perimeter.c:127-238)."""
        q = np.zeros(4)
        self._check(self.lib.chns_perimeter(self.h, _dp(q)), "chns_perimeter")
        return [(int(q[2 * f]), int(q[2 * f + 1])) for f in range(self.nph)]

    def spectrum(self, truncate_state: bool = True):
        """Shell spectra of spectrum.c; truncates the phase fields like spectrum() does."""
        ns, nxy = ctypes.c_int(), ctypes.c_int()
        self.lib.chns_diag_sizes(ctypes.byref(self.p), ctypes.byref(ns), ctypes.byref(nxy))
        out = {k: np.zeros(ns.value) for k in ("ek", "ensk", "s1", "s2", "s12")}
        out["ekx"] = np.zeros(nxy.value)
        out["eky"] = np.zeros(nxy.value)
        self._check(self.lib.chns_spectrum(self.h, int(bool(truncate_state)), *[_dp(out[k]) for k in
                    ("ek", "ensk", "s1", "s2", "s12", "ekx", "eky")]), "chns_spectrum")
        if self.nph == 1:
            out["sph"] = out.pop("s1")
            del out["s2"], out["s12"]
        elif self.nph == 2:
            out["sph1"], out["sph2"], out["sp12"] = out.pop("s1"), out.pop("s2"), out.pop("s12")
        else:
            del out["s1"], out["s2"], out["s12"]
        return out

    def real_field(self, which: int, truncate: bool = True):
        """Packed real [Nx][Ny] field: 0 vorticity, 1/2 phase fields (main.c:332-358)."""
        out = np.empty((self.Nx, self.Ny))
        self._check(self.lib.chns_real_field(self.h, int(which), int(bool(truncate)), _dp(out)), "chns_real_field")
        return out

    # -- discrete adjoint (ternary) ---------------------------------------
    def adj_forward(self, nsteps: int, ckpt_every: int = 100):
        """Advance nsteps steps storing the state every ckpt_every steps (for the reverse sweep)."""
        self._check(self.lib.chns_adj_forward(self.h, int(nsteps), int(ckpt_every)), "chns_adj_forward")

    def adj_seed_spectrum(self, target):
        """Loss sum_s (sph1(s) - target(s))^2 at the current state and its adjoint seed.
        Returns (loss, sph1)."""
        ns, nxy = ctypes.c_int(), ctypes.c_int()
        self.lib.chns_diag_sizes(ctypes.byref(self.p), ctypes.byref(ns), ctypes.byref(nxy))
        t = np.ascontiguousarray(target, dtype=np.float64)
        if t.shape != (ns.value,):
            raise ValueError(f"target needs {ns.value} shells")
        loss = ctypes.c_double()
        S = np.zeros(ns.value)
        self._check(self.lib.chns_adj_seed_spectrum(self.h, _dp(t), ctypes.byref(loss), _dp(S)),
                    "chns_adj_seed_spectrum")
        return loss.value, S

    def adj_seed_free_energy(self):
        """Loss = ephi + egphi of the truncated final phase fields and its seed.
        Returns (loss, ephi, egphi)."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.chns_adj_seed_free_energy(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
                    "chns_adj_seed_free_energy")
        return a.value, b.value, c.value

    def adj_seed(self, wbar=None, c1bar=None, c2bar=None):
        """Seed with spectral adjoints of the final state (IDXC complex, None = zero)."""
        arrs = [None if a is None else self._cplx(a) for a in (wbar, c1bar, c2bar)]
        self._check(self.lib.chns_adj_seed(self.h, *[_dp(a) for a in arrs]), "chns_adj_seed")

    def adj_backward(self):
        """Reverse sweep: (g_wz, g_phi1, g_phi2, g_sigma[s12, s13, s23]) with the real-space
        gradients [Nx][Ny] with respect to the initial real fields."""
        g = [np.empty((self.Nx, self.Ny)) for _ in range(3)]
        gs = np.zeros(3)
        self._check(self.lib.chns_adj_backward(self.h, *[_dp(a) for a in g], _dp(gs)), "chns_adj_backward")
        return g[0], g[1], g[2], gs

    def adj_get(self):
        """Spectral adjoints of the initial state (IDXC) and the coefficient partials."""
        outs = [np.empty((self.Nx, self.Ny2), np.complex128) for _ in range(3)]
        part = np.zeros(len(ADJ_PARTIALS))
        self._check(self.lib.chns_adj_get(self.h, *[_dp(o.view(np.float64)) for o in outs], _dp(part)),
                    "chns_adj_get")
        return outs, dict(zip(ADJ_PARTIALS, part.tolist()))

    def adj_set_store(self, store: bool):
        """Keep the forward intermediates of a replayed segment (default) or recompute them."""
        self._check(self.lib.chns_adj_set_store(self.h, int(bool(store))), "chns_adj_set_store")

    def adj_release(self):
        self._check(self.lib.chns_adj_release(self.h), "chns_adj_release")

    # -- measurement ------------------------------------------------------
    def timing(self, enable: bool = True):
        """Record HIP events around every pass launch on the context's stream."""
        self._check(self.lib.chns_timing_enable(self.h, int(bool(enable))), "chns_timing_enable")

    def timing_read(self):
        """{pass: (total_ms, launches)} since the last read (synchronises)."""
        ms = np.zeros(len(PASS_NAMES))
        n = np.zeros(len(PASS_NAMES), dtype=np.int32)
        self._check(self.lib.chns_timing_read(self.h, _dp(ms), n.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
                    "chns_timing_read")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(PASS_NAMES)}


class Ensemble:
    """Many independent instances on one GPU (``chns_ensemble``): batched
    launches per (size, model) group, one captured HIP graph per step."""

    def __init__(self, plist, device: int = 0):
        self.lib = load_library()
        self.plist = list(plist)
        arr = (CParams * len(self.plist))(*self.plist)
        h = ctypes.c_void_p()
        rc = self.lib.chns_ensemble_create(ctypes.byref(h), arr, len(self.plist), device)
        if rc != CHNS_OK:
            raise ChnsError(f"chns_ensemble_create failed ({rc})")
        self.h = h
        self.instances = [Context(p, _handle=ctypes.c_void_p(self.lib.chns_ensemble_instance(h, b)))
                          for b, p in enumerate(self.plist)]

    def _check(self, rc, what):
        if rc != CHNS_OK:
            raise ChnsError(f"{what} failed ({rc}): {self.lib.chns_ensemble_last_error(self.h).decode()}")

    def __len__(self):
        return len(self.instances)

    def __getitem__(self, b) -> Context:
        return self.instances[b]

    def step(self, nsteps: int = 1):
        self._check(self.lib.chns_ensemble_step(self.h, int(nsteps)), "chns_ensemble_step")

    def sync(self):
        self._check(self.lib.chns_ensemble_sync(self.h), "chns_ensemble_sync")

    def groups(self):
        """[(log2 N, model, count)] in group order."""
        n = self.lib.chns_ensemble_groups(self.h, None, None, None)
        a, b, c = (ctypes.c_int * n)(), (ctypes.c_int * n)(), (ctypes.c_int * n)()
        self.lib.chns_ensemble_groups(self.h, a, b, c)
        return [(a[k], b[k], c[k]) for k in range(n)]

    def tune(self, reps: int = 5):
        """Autotune the launch split and FFT schedule per group; returns the chosen splits."""
        n = len(self.groups())
        out = (ctypes.c_int * n)()
        self._check(self.lib.chns_ensemble_tune(self.h, int(reps), out), "chns_ensemble_tune")
        return list(out)

    def schedule(self):
        """[(split, sched)] per group (sched bit 0: radix-16 column passes, bit 1: half-row workgroups)."""
        n = len(self.groups())
        a, b = (ctypes.c_int * n)(), (ctypes.c_int * n)()
        self.lib.chns_ensemble_schedule(self.h, a, b)
        return list(zip(list(a), list(b)))

    def set_schedule(self, group: int, split: int, sched: int):
        self._check(self.lib.chns_ensemble_set_schedule(self.h, int(group), int(split), int(sched)),
                    "chns_ensemble_set_schedule")

    def close(self):
        if getattr(self, "h", None):
            for c in self.instances:
                c.h = None
            self.lib.chns_ensemble_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ---------------------------------------------------------------- reference-named mirrors
def etd2rk(ctx: Context, nsteps: int = 1):
    """etd2rk(cwz, cphi1, cphi2) (etd2rk.c:14-224) on the device-resident state."""
    ctx.step(nsteps)


def spectrum_truncate(ctx: Context):
    """State side effect of spectrum() (spectrum.c:22-27)."""
    ctx.truncate_state()


def etd_tables(p: CParams):
    """Host-built ETD2RK coefficient tables (IDXC order), no device needed."""
    lib = load_library()
    shape = (p.Nx, p.Ny // 2 + 1)
    names = ("eta", "z1w", "z2w", "beta", "z1c", "z2c")
    out = {k: np.zeros(shape) for k in names}
    rc = lib.chns_etd_tables(ctypes.byref(p), *[_dp(out[k]) for k in names])
    if rc != CHNS_OK:
        raise ChnsError(f"chns_etd_tables failed ({rc})")
    return out


def dealias_extent(p: CParams):
    """Largest kept |kx index| per ky column (-1: column fully truncated)."""
    lib = load_library()
    a = np.zeros(p.Ny // 2 + 1, dtype=np.int32)
    rc = lib.chns_dealias_extent(ctypes.byref(p), a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    if rc != CHNS_OK:
        raise ChnsError(f"chns_dealias_extent failed ({rc})")
    return a


def filter_table(p: CParams):
    """SMOOTH_FILTER factor exp(-36 pp^18) per mode (IDXC order), no device needed."""
    lib = load_library()
    out = np.zeros((p.Nx, p.Ny // 2 + 1))
    rc = lib.chns_filter_table(ctypes.byref(p), _dp(out))
    if rc != CHNS_OK:
        raise ChnsError(f"chns_filter_table failed ({rc})")
    return out


def two_pi():
    return 2.0 * math.pi
