"""Adjoint (reverse-mode) ERK4(3)-IP for the 3+1D envelope solver behind
``torch.autograd.Function``.

No counterpart in the reference.  The forward propagation is first run adaptively to fix the
accepted step sequence (the gradient is that of the discrete map with those steps).  The
frozen forward replays the steps on the GPU and keeps a checkpoint of the state every
``ckpt_every`` steps (default ceil(sqrt(M)) for M steps).  The backward pass recomputes each
segment from its checkpoint, recording the state and the stage fields X2, X3, X4 of every step
of the segment, then runs the adjoint step kernels in reverse (``uppe3d_adjoint_step``).

The differentiated map is the state map B0 -> B(L) in the solver's internal layout
[omega][x][y] (B = FFT_t(A) / A_scale, omega in FFT order).  The input parametrisation
(spatial phase, spectral phase) and the objective are ordinary torch code around it.  A_scale
is a pure rescaling of the state (solver_3d.py:166-170) and is held fixed.  The plasma term is
not differentiated: the backward pass raises if any line of the propagation had a plasma term
above FP32 resolution.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib


class FrozenPropagator3D:
    """GPU propagation of one 3+1D field over a frozen accepted-step sequence."""

    def __init__(self, solver, A0, L, tol=1e-4, max_step=1e-2, min_step=1e-9, ckpt_every=None,
                 max_attempts=100000):
        self.solver = solver
        self.L, self.tol, self.max_step, self.min_step = L, tol, max_step, min_step
        self.shape = (solver.Nt, solver.Nx, solver.Ny)
        self.dev = torch.device("cuda")
        a = solver.load(A0)                       # plan with A_scale = max|A0|
        self.A_scale = solver.A_scale
        self.B0 = self.get_state()                # state of A0
        self.hs = self._record(max_attempts)
        self.M = len(self.hs)
        self.C = int(ckpt_every) if ckpt_every else max(1, int(math.ceil(math.sqrt(self.M))))
        del a

    # -- plan helpers ----------------------------------------------------------------------
    def _plan(self):
        return self.solver._plan

    def get_state(self):
        out = torch.empty(self.shape, dtype=torch.complex64, device=self.dev)
        _lib.check(_lib.load().uppe3d_get_state(self._plan(), _lib.ptr(out),
                                                _lib.stream_handle()), "uppe3d_get_state")
        return out

    def set_state(self, B):
        B = B.to(self.dev, torch.complex64).contiguous()
        _lib.check(_lib.load().uppe3d_set_state(self._plan(), _lib.ptr(B),
                                                _lib.stream_handle()), "uppe3d_set_state")

    def _replay(self, h):
        _lib.check(_lib.load().uppe3d_replay_attempt(self._plan(), float(h), 1,
                                                     _lib.stream_handle()),
                   "uppe3d_replay_attempt")

    def _record(self, max_attempts):
        """Adaptive run from the loaded field; returns the accepted step sizes."""
        lib = _lib.load()
        s = _lib.stream_handle()
        _lib.check(lib.uppe3d_begin(self._plan(), self.L, self.tol, self.max_step,
                                    self.min_step, s), "uppe3d_begin")
        hs = []
        for _ in range(max_attempts):
            st = self.solver.status()
            if st.done:
                break
            h = st.h
            _lib.check(lib.uppe3d_attempt(self._plan(), s), "uppe3d_attempt")
            st2 = self.solver.status()
            if st2.nonfinite:
                raise FloatingPointError("non-finite error norm at z=%g" % st.z)
            if st2.accepted > st.accepted:
                hs.append(h)
        else:
            raise RuntimeError("step recording exceeded max_attempts")
        self.record_status = self.solver.status()
        return hs

    # -- frozen forward / backward ---------------------------------------------------------
    def forward(self, B0):
        """B(L) of the frozen map from state B0 (layout [omega][x][y]); keeps checkpoints."""
        lib = _lib.load()
        s = _lib.stream_handle()
        _lib.check(lib.uppe3d_begin(self._plan(), 1e300, 1.0, 1.0, 0.0, s), "uppe3d_begin")
        self.set_state(B0)
        nck = (self.M + self.C - 1) // self.C
        self.ckpt = torch.empty((max(nck, 1),) + self.shape, dtype=torch.complex64,
                                device=self.dev)
        for k, h in enumerate(self.hs):
            if k % self.C == 0:
                self.ckpt[k // self.C].copy_(self.get_state())
            self._replay(h)
        return self.get_state()

    def backward(self, Bbar):
        """dL/dB0 from dL/dB(L) (same layout), through the frozen steps in reverse."""
        lib = _lib.load()
        s = _lib.stream_handle()
        nb = ctypes.c_size_t()
        _lib.check(lib.uppe3d_adjoint_workspace_size(ctypes.byref(self.solver._desc),
                                                     ctypes.byref(nb)), "adjoint ws")
        aws = torch.empty(nb.value, dtype=torch.uint8, device=self.dev)
        Bb = Bbar.to(self.dev, torch.complex64).contiguous().clone()
        C = self.C
        seg = torch.empty((C,) + self.shape, dtype=torch.complex64, device=self.dev)
        xs = torch.empty((3, C) + self.shape, dtype=torch.complex64, device=self.dev)
        plan = self._plan()
        try:
            for j in reversed(range((self.M + C - 1) // C)):
                k0, k1 = j * C, min(self.M, (j + 1) * C)
                self.set_state(self.ckpt[j])
                for k in range(k0, k1):
                    i = k - k0
                    seg[i].copy_(self.get_state())
                    _lib.check(lib.uppe3d_set_stage_record(
                        plan, _lib.ptr(xs[0, i]), _lib.ptr(xs[1, i]), _lib.ptr(xs[2, i])),
                        "stage record")
                    self._replay(self.hs[k])
                _lib.check(lib.uppe3d_set_stage_record(plan, None, None, None), "stage record")
                for k in reversed(range(k0, k1)):
                    i = k - k0
                    _lib.check(lib.uppe3d_adjoint_step(
                        plan, float(self.hs[k]), _lib.ptr(seg[i]), _lib.ptr(xs[0, i]),
                        _lib.ptr(xs[1, i]), _lib.ptr(xs[2, i]), _lib.ptr(Bb), _lib.ptr(aws), s),
                        "uppe3d_adjoint_step")
        finally:
            lib.uppe3d_set_stage_record(plan, None, None, None)
        flag = ctypes.c_int(0)
        _lib.check(lib.uppe3d_adjoint_check(plan, s, ctypes.byref(flag)), "adjoint check")
        if flag.value:
            raise NotImplementedError("the plasma term is significant on some lines; the 3D "
                                      "adjoint does not differentiate it")
        return Bb


class UPPEPropagate3D(torch.autograd.Function):
    """B(L) = frozen propagation of the state B0 (layout [omega][x][y], complex64)."""

    @staticmethod
    def forward(ctx, B0, prop):
        ctx.prop = prop
        return prop.forward(B0.detach())

    @staticmethod
    def backward(ctx, gout):
        return ctx.prop.backward(gout), None


# -- config 4 objective: on-axis spectral energy in a wavelength band ------------------------
def to_state(A, A_scale):
    """(Nx, Ny, Nt) time-domain field -> state [omega][x][y] = FFT_t(A) / A_scale (torch)."""
    return (torch.fft.fft(A, dim=2) / A_scale).permute(2, 0, 1).contiguous()


def phase_field(A_base, phi, psi):
    """A_base(x, y, t) with spatial phase phi(x, y) and spectral phase psi(omega, FFT order)."""
    Aw = torch.fft.fft(A_base * torch.exp(1j * phi)[:, :, None], dim=2)
    return torch.fft.ifft(Aw * torch.exp(1j * psi)[None, None, :], dim=2)


def band_mask(omega, omega0, lam_lo, lam_hi):
    """Boolean mask (FFT order) of the frequencies whose wavelength lies in [lam_lo, lam_hi]."""
    c = 299792458.0
    wabs = np.asarray(omega, np.float64) + omega0
    with np.errstate(divide="ignore"):
        lam = np.where(wabs > 0, 2 * np.pi * c / np.maximum(wabs, 1e-30), np.inf)
    return (lam >= lam_lo) & (lam <= lam_hi)


def onaxis_band_energy(B, A_scale, mask, ix, iy):
    """sum over the band (boolean mask over omega, FFT order) of |A_scale B(omega, ix, iy)|^2."""
    idx = torch.as_tensor(np.nonzero(np.asarray(mask))[0], device=B.device)
    return (A_scale ** 2) * torch.sum(torch.abs(B[idx, ix, iy]) ** 2)
