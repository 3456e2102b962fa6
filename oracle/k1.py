"""ORACLE (test infrastructure only): the reference's serial FP64 plasma recurrence.

Restates ``src/3+1D-UPPE-ERK4(3)IP-envelope/media_physics.py:11-66`` (kernel text) with the
float32 argument casts of its launch (``media_physics.py:182-189``).  Two implementations:

* :func:`k1_numpy` - NumPy, vectorised over lines, loop over t (the definition);
* :func:`k1` - the C restatement ``plasma_k1.c`` via ctypes (same arithmetic, fast).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` use this.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libplasma_k1.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        lib = ctypes.CDLL(_SO)
        lib.k1_plasma_lines.restype = None
        lib.k1_plasma_lines.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_long, ctypes.c_long]
        _lib = lib
    return _lib


def k1(w_ion, w_ava, rate_rec, dt, rho_nt):
    """rho, gen = K1(W_ion, W_ava) over the last axis (t); inputs cast to float32."""
    w_ion = np.ascontiguousarray(w_ion, dtype=np.float32)
    w_ava = np.ascontiguousarray(np.broadcast_to(w_ava, w_ion.shape), dtype=np.float32)
    nt = w_ion.shape[-1]
    nlines = w_ion.size // nt if nt else 0
    rho = np.zeros(w_ion.shape, np.float64)
    gen = np.zeros(w_ion.shape, np.float64)
    if w_ion.size:
        _load().k1_plasma_lines(
            w_ion.ctypes.data, w_ava.ctypes.data, rho.ctypes.data, gen.ctypes.data,
            np.float32(rate_rec), np.float32(dt), np.float32(rho_nt), nt, nlines)
    return rho, gen


def k1_numpy(w_ion, w_ava, rate_rec, dt, rho_nt):
    """Same recurrence in NumPy (vectorised over lines); the definition the C file follows."""
    wi = np.asarray(w_ion, np.float32).astype(np.float64)
    wa = np.broadcast_to(np.asarray(w_ava, np.float32), wi.shape).astype(np.float64)
    rr = float(np.float32(rate_rec))
    dtd = float(np.float32(dt))
    rnt = float(np.float32(rho_nt))
    rho_out = np.zeros(wi.shape)
    gen_out = np.zeros(wi.shape)
    rho = np.zeros(wi.shape[:-1])
    for t in range(wi.shape[-1]):
        rho_out[..., t] = rho
        a = wi[..., t]
        b = wa[..., t]
        depl = np.maximum((rnt - rho) / rnt, 0.0)
        gen_out[..., t] = (a * rnt + b * rho) * depl
        src = a * rnt * depl
        g = b * depl - rr
        gdt = np.clip(g * dtd, -10.0, 10.0)
        e = np.exp(gdt)
        small = np.abs(gdt) < 1e-5
        with np.errstate(divide="ignore", invalid="ignore"):
            f = np.where(small, dtd * (1.0 + 0.5 * gdt), (e - 1.0) / np.where(small, 1.0, g))
        rho = np.clip(rho * e + src * f, 0.0, rnt)
    return rho_out, gen_out
