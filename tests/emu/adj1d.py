"""Runs the emulated 1D forward + adjoint (tests/emu/adj1d_driver.cpp) on a pulse and a solver
of the drop-in class, and returns the recorded step count, A(L) and the gradients.
Test infrastructure: the host build of uppe1d.hip under AddressSanitizer."""
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BUILD = os.path.join(ROOT, "tests", "emu", "_build")
EXE = os.path.join(BUILD, "adj1d")


def build(force=False):
    from emu_build import build as b   # noqa: E402  (tests/emu on sys.path)
    src = os.path.join(ROOT, "project_amd", "csrc", "uppe1d.hip")
    drv = os.path.join(HERE, "adj1d_driver.cpp")
    deps = [src, drv, os.path.join(HERE, "hip", "hip_runtime.h"),
            os.path.join(ROOT, "project_amd", "csrc", "fft_common.hpp"),
            os.path.join(ROOT, "include", "uppe.h")]
    if not force and os.path.exists(EXE) and \
            os.path.getmtime(EXE) > max(os.path.getmtime(p) for p in deps):
        return EXE
    return b(src, drv, EXE)


def run(solver, t, a0, L, tol, S_target, ckpt_every=16, hs_cap=20000, order_seed=0,
        max_step=1e-2, min_step=1e-11, timeout=900):
    t = np.asarray(t, np.float64)
    n = len(t)
    dt = t[1] - t[0]
    omega = 2 * np.pi * np.fft.fftfreq(n, dt)
    D, U, H = solver._tables(omega)
    assert H is None
    d = solver._desc(n, dt, L, tol, max_step, min_step, 0)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("4i", n, ckpt_every, hs_cap, 0))
            f.write(bytes(d))
            f.write(np.ascontiguousarray(D.real, np.float64).tobytes())
            f.write(np.ascontiguousarray(D.imag, np.float64).tobytes())
            f.write(np.ascontiguousarray(U, np.float64).tobytes())
            f.write(np.ascontiguousarray(a0, np.complex64).tobytes())
            f.write(np.ascontiguousarray(S_target, np.float64).tobytes())
        env = dict(os.environ, EMU_ORDER_SEED=str(order_seed),
                   ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1")
        r = subprocess.run([EXE, fin, fout], capture_output=True, text=True, env=env,
                           timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError("emulated run failed (rc %d):\n%s" % (r.returncode,
                                                                   r.stderr[-8000:]))
        b = open(fout, "rb").read()
    M, attempts, rcf, rca = struct.unpack("4i", b[:16])
    o = 16
    out = np.frombuffer(b, np.complex64, n, o); o += 8 * n
    lam = np.frombuffer(b, np.complex64, n, o); o += 8 * n
    gD = np.frombuffer(b, np.float64, 2 * n, o); o += 16 * n
    gU = np.frombuffer(b, np.float64, n, o); o += 8 * n
    gG = np.frombuffer(b, np.float64, 1, o)[0]; o += 8
    hs = np.frombuffer(b, np.float64, M, o)
    return dict(M=M, attempts=attempts, rc_fwd=rcf, rc_adj=rca, out=out.copy(),
                gA0=lam.copy(), gD=gD[0::2] + 1j * gD[1::2], gU=gU.copy(), gG=gG, hs=hs.copy(),
                stderr=r.stderr)
