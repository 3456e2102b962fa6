"""Adjoint (reverse-mode) ERK4(3)-IP for the 1D solver behind ``torch.autograd.Function``.

No counterpart in the reference (its only "optimisation" is brute-force parameter scans,
``src/1D-ERK4(3)IP UPPE/test_spectrum_pulse-coupling_efficiency-out.py:284-311``); it serves
the reference's stated goal of computing optimal parameters.

The forward pass runs the adaptive propagation once on the GPU, recording the accepted step
sequence and a checkpoint of the state every ``ckpt_every`` accepted steps; that sequence is
then frozen (the gradient is the exact derivative of the discrete map with those steps).
The backward pass (one HIP kernel) recomputes each segment from its checkpoint and walks the
ERK4(3)-IP stages in reverse, returning the gradients of the loss with respect to
* the input spectrum A0(omega),
* the dispersion operator D(omega) and the UPPE factor U(omega) tables,
* the nonlinear coefficient gamma.
Physical parameters (spectral phase, pulse energy, gas pressure) enter through these; the
pressure derivatives of the host tables are formed in FP64 by central differences of the
solver's own hooks (``pressure_jacobian``).

Complex gradients follow torch's convention dL/dRe z + i dL/dIm z.
"""
import ctypes

import numpy as np
import torch

from . import _lib


class FrozenPropagator1D:
    """GPU propagation of one 1D instance with step recording / frozen-step replay."""

    def __init__(self, solver, t, L, tol=1e-6, max_step=1e-2, min_step=1e-11, ckpt_every=16,
                 hs_cap=20000):
        self.solver = solver
        self.t = np.asarray(t, np.float64)
        self.n = len(self.t)
        self.dt = self.t[1] - self.t[0]
        self.omega = 2 * np.pi * np.fft.fftfreq(self.n, self.dt)
        self.L, self.tol, self.max_step, self.min_step = L, tol, max_step, min_step
        self.C = ckpt_every
        self.hs_cap = hs_cap
        self.dev = torch.device("cuda")
        self.hs = None          # frozen step sequence (device float64), set by the first run
        self.M = 0

    def tables(self):
        D, U, H = self.solver._tables(self.omega)
        if H is not None:
            raise NotImplementedError("the adjoint does not cover the Raman response")
        return D, U, float(self.solver.gamma)

    def run(self, A0w, D, U, gamma, record=True):
        """A0w: complex64 torch (n,) omega domain; D complex128, U float64 (numpy); gamma."""
        lib = _lib.load()
        n = self.n
        a0 = torch.fft.ifft(A0w.to(self.dev, torch.complex64)).contiguous()
        d = self.solver._desc(n, self.dt, self.L, self.tol, self.max_step, self.min_step, 0)
        d.gamma = float(gamma)
        nb = ctypes.c_size_t()
        _lib.check(lib.uppe1d_workspace_size(ctypes.byref(d), ctypes.byref(nb)), "ws size")
        self.ws = torch.empty(nb.value, dtype=torch.uint8, device=self.dev)
        out = torch.empty(n, dtype=torch.complex64, device=self.dev)
        rec = _lib.Uppe1DRecord()
        replay = self.hs is not None
        if record or replay:
            cap = self.M if replay else self.hs_cap
            self.ckpt = torch.empty(((cap + self.C - 1) // self.C, n), dtype=torch.complex64,
                                    device=self.dev)
            hs_new = torch.empty(max(cap, 1), dtype=torch.float64, device=self.dev)
            rec.hs = ctypes.c_void_p(hs_new.data_ptr())
            rec.hs_cap = cap
            rec.ckpt = ctypes.c_void_p(self.ckpt.data_ptr())
            rec.ckpt_every = self.C
            rec.ckpt_cap = self.ckpt.shape[0]
            if replay:
                rec.replay_hs = ctypes.c_void_p(self.hs.data_ptr())
                rec.replay_M = self.M
        Dre = np.ascontiguousarray(np.real(D), np.float64)
        Dim = np.ascontiguousarray(np.imag(D), np.float64)
        Uh = np.ascontiguousarray(U, np.float64)
        st = _lib.Uppe1DStatus()
        rc = lib.uppe1d_propagate(ctypes.byref(d), _lib.host_ptr(Dre), _lib.host_ptr(Dim),
                                  _lib.host_ptr(Uh), None, None, _lib.ptr(a0), _lib.ptr(self.ws),
                                  None, None, _lib.ptr(out), ctypes.byref(rec),
                                  _lib.stream_handle(), ctypes.byref(st))
        if rc == -3:
            raise FloatingPointError(lib.uppe_last_error().decode())
        _lib.check(rc, "uppe1d_propagate")
        if st.overflow & 2:
            raise RuntimeError("step record capacity exceeded; raise hs_cap")
        if not replay:
            self.M = st.accepted
            self.hs = hs_new[:self.M].clone()
        self.last_status = st
        return out

    def backward(self, lam):
        """lam = dL/dA(L) (omega). Returns dL/dA0(omega), dL/dD, dL/dU, dL/dgamma."""
        lib = _lib.load()
        n = self.n
        lam = lam.to(self.dev, torch.complex64).contiguous().clone()
        nb = ctypes.c_size_t()
        _lib.check(lib.uppe1d_adjoint_workspace_size(n, self.C, ctypes.byref(nb)), "adj ws")
        aws = torch.empty(nb.value, dtype=torch.uint8, device=self.dev)
        gD = torch.empty(2 * n, dtype=torch.float64, device=self.dev)
        gU = torch.empty(n, dtype=torch.float64, device=self.dev)
        gG = torch.empty(1, dtype=torch.float64, device=self.dev)
        _lib.check(lib.uppe1d_adjoint(_lib.ptr(self.ws), 1, 0, self.M, _lib.ptr(self.hs),
                                      _lib.ptr(self.ckpt), self.C, _lib.ptr(lam), _lib.ptr(aws),
                                      _lib.ptr(gD), _lib.ptr(gU), _lib.ptr(gG),
                                      _lib.stream_handle()), "uppe1d_adjoint")
        gDc = torch.complex(gD[0::2], gD[1::2])
        return lam, gDc, gU, gG[0]


class UPPEPropagate1D(torch.autograd.Function):
    """A(L)(omega) = propagate(A0(omega); D, U, gamma) with the step sequence frozen by the
    propagator's first run."""

    @staticmethod
    def forward(ctx, A0w, D, U, gamma, prop):
        out = prop.run(A0w, D.detach().cpu().numpy(), U.detach().cpu().numpy(),
                       float(gamma.detach().cpu()), record=True)
        ctx.prop = prop
        ctx.dtypes = (A0w.dtype, D.dtype, U.dtype, gamma.dtype)
        return out

    @staticmethod
    def backward(ctx, gout):
        gA0, gD, gU, gG = ctx.prop.backward(gout)
        tA, tD, tU, tG = ctx.dtypes
        return gA0.to(tA), gD.to(tD), gU.to(tU), gG.to(tG).reshape(()), None


# -- config 3: spectral-shaping objective and physical parameters ------------------------
def super_gaussian_target(omega, A0w, seed=3, widen=2.0):
    """Seeded smooth super-Gaussian spectrum of `widen` x the input RMS bandwidth, scaled to
    the input spectral energy (so the loss is insensitive to units)."""
    rng = np.random.default_rng(seed)
    S0 = np.abs(A0w) ** 2
    wc = np.sum(omega * S0) / np.sum(S0)
    bw = np.sqrt(np.sum((omega - wc) ** 2 * S0) / np.sum(S0))
    order = rng.uniform(2.0, 4.0)
    shift = rng.uniform(-0.2, 0.2) * bw
    S = np.exp(-np.abs((omega - wc - shift) / (widen * bw * np.sqrt(2.0))) ** (2 * order))
    return S * (np.sum(S0) / np.sum(S))


def spectral_loss(Aw, S_target):
    return torch.sum((torch.abs(Aw) ** 2 - S_target) ** 2)


def pressure_jacobian(make_solver, omega, P, rel=1e-6):
    """Central FP64 differences of the host tables (D, U, gamma) in the gas pressure P, built
    from the solver's own refractive-index and constructor hooks."""
    def tabs(p):
        s = make_solver(p)
        D, U, _ = s._tables(omega)
        return D, U, s.gamma
    hP = P * rel
    Dp, Up, gp = tabs(P + hP)
    Dm, Um, gm = tabs(P - hP)
    return (Dp - Dm) / (2 * hP), (Up - Um) / (2 * hP), (gp - gm) / (2 * hP)
