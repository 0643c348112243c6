"""ctypes front of the C restatement of the step (oracle/chns_step.c).

TEST INFRASTRUCTURE ONLY, like the rest of ``oracle/``: the checker for the
HIP library at sizes where the numpy oracle is slow, and the CPU baseline of
bench.py.  Built by ``make -C oracle`` (``__graft_entry__.build`` runs it).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libchns_oracle.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):   # not built yet: build it (gcc, a few seconds)
            import subprocess
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        lib = ctypes.CDLL(LIB_PATH)
        D = ctypes.POINTER(ctypes.c_double)
        lib.chnso_create.argtypes = [D, ctypes.POINTER(ctypes.c_int)]
        lib.chnso_create.restype = ctypes.c_void_p
        lib.chnso_destroy.argtypes = [ctypes.c_void_p]
        lib.chnso_set_state.argtypes = [ctypes.c_void_p, D, D, D]
        lib.chnso_get_state.argtypes = [ctypes.c_void_p, D, D, D]
        lib.chnso_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.chnso_truncate_state.argtypes = [ctypes.c_void_p]
        lib.chnso_threads.argtypes = [ctypes.c_int]
        lib.chnso_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def threads(n: int = 0) -> int:
    """Set the OpenMP thread count (n > 0) and return the count in use."""
    return load().chnso_threads(int(n))


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Stepper:
    """The ETD2RK step of ``oracle.Params`` p (plain step: no optional physics,
    the reference's N/4 truncation)."""

    def __init__(self, p):
        if p.gravity or p.vpm or p.sponge or p.smooth_filter or getattr(p, "dealias", 0) or p.Nx != p.Ny:
            raise ValueError("the C restatement covers the plain step on square grids")
        lib = load()
        dp = np.array([p.lx, p.nu, p.dt, p.sigma, p.epsilon, p.lam, p.sigma_12, p.sigma_13, p.sigma_23,
                       p.Xlambda, p.f_amp], dtype=np.float64)
        ip = np.array([p.Nx, p.model, p.f_k], dtype=np.int32)
        self.h = lib.chnso_create(_dp(dp), ip.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        if not self.h:
            raise ValueError("unsupported grid")
        self.p = p
        self.nph = {0: 0, 1: 1, 2: 2}[p.model]

    def close(self):
        if getattr(self, "h", None):
            load().chnso_destroy(self.h)
            self.h = None

    __del__ = close

    def set_state(self, s):
        f = [np.ascontiguousarray(x, dtype=np.complex128) for x in s.fields()]
        self._keep = f
        load().chnso_set_state(self.h, *[None if x is None else _dp(x.view(np.float64)) for x in f + [None] * (3 - len(f))])

    def get_state(self):
        import oracle as O
        out = [np.empty((self.p.Nx, self.p.Ny2), np.complex128) for _ in range(1 + self.nph)]
        load().chnso_get_state(self.h, *[None if x is None else _dp(x.view(np.float64)) for x in out + [None] * (2 - self.nph)])
        return O.State(*out)

    def step(self, n: int = 1):
        load().chnso_step(self.h, int(n))

    def truncate_state(self):
        load().chnso_truncate_state(self.h)


def run(p, s, nsteps: int, spectrum_every: int = 0):
    """oracle.run() on the C restatement."""
    st = Stepper(p)
    try:
        st.set_state(s)
        if spectrum_every:
            done = 0
            while done < nsteps:
                k = min(spectrum_every - done % spectrum_every, nsteps - done)
                st.step(k)
                done += k
                if done % spectrum_every == 0:
                    st.truncate_state()
        else:
            st.step(nsteps)
        return st.get_state()
    finally:
        st.close()
