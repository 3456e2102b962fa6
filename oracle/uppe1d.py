"""ORACLE (test infrastructure only): FP64 NumPy restatement of the reference's 1D
ERK4(3)-IP UPPE solver in gas.

Never imported by the product package.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` use it, as the checker.

Follows, function by function (reference paths relative to the reference root):

* ``src/1D-ERK4(3)IP UPPE gas/erk43ip_method.py:21-70``   constructor semantics (gamma rule, n0)
* ``erk43ip_method.py:72-102``                            Sellmeier table and fallback
* ``src/1D-ERK4(3)IP UPPE gas/gas.py:14-40``              argon Peck & Fisher patch, P/T scaling
* ``erk43ip_method.py:104-148``                           dispersion operator (central-difference k1)
* ``erk43ip_method.py:150-179``                           Raman response (dt floor 1e-15 quirk)
* ``erk43ip_method.py:181-195``                           linear operator
* ``erk43ip_method.py:197-255``                           nonlinear operator (UPPE factor)
* ``erk43ip_method.py:256-286``                           ERK4(3) step with FSAL rule
* ``erk43ip_method.py:288-369``                           adaptive propagate, history rule
* ``erk43ip_method.py:375-446``                           pulse generators

Extension with no counterpart in the 1D reference (the configs ask for it): a plasma term
restated from the 3D medium, ``src/3+1D-UPPE-ERK4(3)IP-envelope/media_physics.py:156-205``,
with W = sigma_K * I**K (float32 cast), the K1 recurrence (``oracle/k1.py``) and
loss + defocus terms.  Intensity is |A|**2 / area.

Pinned by: golden vectors written from the real reference (``tests/golden/``), closed forms
(linear propagation, SPM, Peck & Fisher values), invariants (energy, photon number) and the
order of convergence (``tests/test_oracle_1d.py``).
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
from scipy.fft import fft, ifft, fftfreq, fftshift, ifftshift

from . import k1 as _k1

C_LIGHT = 299792458.0
E_CHARGE = 1.60217663e-19
EPS0 = 8.85418781e-12
M_E = 9.1093837e-31

# erk43ip_method.py:72-91 (unknown materials fall back to the fused-silica entry)
SELLMEIER = {
    "fused_silica": ((0.6961663, 0.4079426, 0.8974794),
                     (0.0684043 ** 2, 0.1162414 ** 2, 9.896161 ** 2), (0.21e-6, 6.7e-6)),
    "sapphire": ((1.4313493, 0.65054713, 5.3414021),
                 (0.0726631 ** 2, 0.1193242 ** 2, 18.028251 ** 2), (0.2e-6, 5.5e-6)),
    "yag": ((2.282, 3.27644, 4.44174), (0.01185 ** 2, 0.0282 ** 2, 282.734 ** 2),
            (0.4e-6, 5e-6)),
}


def sellmeier_index(lam, material):
    b, c, _ = SELLMEIER.get(material, SELLMEIER["fused_silica"])
    l2 = (np.asarray(lam, np.float64) * 1e6) ** 2
    n2 = np.ones_like(l2)
    for bi, ci in zip(b, c):
        n2 = n2 + bi * l2 / (l2 - ci)
    return np.sqrt(n2)


def argon_peck_fisher_index(lam, P, T):
    """gas.py:20-40: (n-1)*1e8 at 15 C / 1 atm, scaled by (P/101325)*(288.15/T)."""
    s2 = 1.0 / (np.asarray(lam, np.float64) * 1e6) ** 2
    nm1 = (1205582.27 / (130.0 - s2) + 317222.05 / (38.9 - s2)) * 1e-8
    return 1.0 + nm1 * (P / 101325.0) * (288.15 / T)


@dataclass
class Plasma1D:
    """Parameters of the 1D plasma extension (defaults: the reference's Argon class at 1 atm,
    media_physics.py:396-432, with K = 8)."""
    K: int = 8
    sigma_K: float = 1.0e-130
    Ui: float = 15.76 * E_CHARGE
    rho_nt: float = 2.45e25
    tau_rec: float = 1.0e-9
    tau_c: float = 190e-15
    sigma_drude: Optional[float] = None      # None: Drude formula with n0 = 1 (Argon class)
    avalanche: bool = False                  # Argon: avalanche rate is zero (:431-432)
    area: Optional[float] = None             # None: pi * beam_radius**2 of the solver

    def drude(self, omega0):
        if self.sigma_drude is not None:
            return self.sigma_drude
        return (E_CHARGE ** 2 * self.tau_c) / (
            C_LIGHT * EPS0 * 1.0 * M_E * (1.0 + omega0 ** 2 * self.tau_c ** 2))


@dataclass
class Params1D:
    material: str = "fused_silica"
    gamma: Optional[float] = None
    n2: Optional[float] = None
    beam_radius: Optional[float] = None
    alpha: float = 0.0
    center_wavelength: float = 1064e-9
    use_raman: bool = True
    f_R: float = 0.18
    tau1: float = 12.2e-15
    tau2: float = 32e-15
    use_self_steepening: bool = True
    gas_P: float = 101325.0
    gas_T: float = 293.15
    plasma: Optional[Plasma1D] = None


class Oracle1D:
    """State of one reference solver instance (constructor rule of erk43ip_method.py:21-70)."""

    def __init__(self, p: Params1D, precision="double"):
        """precision "single": the same algorithm with the field, D, the UPPE factor and the
        nonlinear terms in complex64 / float32 (the plasma recurrence stays FP64) - a measure of
        what FP32 arithmetic costs this algorithm (the reference's 1D solver is FP64 only)."""
        self.p = p
        self.precision = precision
        self.lambda0 = p.center_wavelength
        self.omega0 = 2 * np.pi * C_LIGHT / self.lambda0
        if p.gamma is not None:
            self.gamma, self.n2, self.beam_radius = p.gamma, 0.0, 0.0
            self.beam_area = None
        elif p.n2 is not None and p.beam_radius is not None:
            self.n2, self.beam_radius = p.n2, p.beam_radius
            self.beam_area = np.pi * p.beam_radius ** 2
            self.gamma = (2 * np.pi / self.lambda0) * (p.n2 / self.beam_area)
        else:
            self.gamma, self.n2, self.beam_radius, self.beam_area = 0.0, 0.0, 0.0, None
        self.n0 = float(self.index(np.array([self.lambda0]))[0])
        self._omega = None

    # -- material -----------------------------------------------------------------------
    def index(self, lam):
        if self.p.material == "argon_gas":
            return argon_peck_fisher_index(lam, self.p.gas_P, self.p.gas_T)
        return sellmeier_index(lam, self.p.material)

    def valid_range(self):
        return SELLMEIER.get(self.p.material, SELLMEIER["fused_silica"])[2]

    # -- operators ----------------------------------------------------------------------
    def setup(self, omega):
        """Dispersion operator, n(omega) cache and Raman response for one omega grid."""
        w0 = self.omega0
        wabs = w0 + omega
        wsafe = np.where(np.abs(wabs) < 1e-10, 1e-10, np.abs(wabs))
        lo, hi = self.valid_range()
        lam = np.clip(2 * np.pi * C_LIGHT / wsafe, lo, hi)
        n_w = self.index(lam)
        k = n_w * wabs / C_LIGHT
        dw = w0 * 1e-5
        n_c = self.index(np.array([self.lambda0]))[0]
        n_p = self.index(np.array([2 * np.pi * C_LIGHT / (w0 + dw)]))[0]
        n_m = self.index(np.array([2 * np.pi * C_LIGHT / (w0 - dw)]))[0]
        k0 = n_c * w0 / C_LIGHT
        k1 = (n_p * (w0 + dw) / C_LIGHT - n_m * (w0 - dw) / C_LIGHT) / (2 * dw)
        self.D = 1j * (k - k0 - k1 * omega) - self.p.alpha / 2
        self.n_w = n_w
        self.k0, self.k1 = k0, k1
        self.omega = omega
        self.H_raman = self._raman(omega) if self.p.use_raman else None
        self.uppe = (wabs / w0) * (self.n0 / np.where(n_w < 1.0, 1.0, n_w))
        if self.precision == "single":
            self.D = self.D.astype(np.complex64)
            self.uppe = self.uppe.astype(np.float32)
            if self.H_raman is not None:
                self.H_raman = self.H_raman.astype(np.complex64)

    def _raman(self, omega):
        n = len(omega)
        dt = 2 * np.pi / (n * abs(omega[1] - omega[0])) if n > 1 else 1e-15
        dt = max(dt, 1e-15)
        t = np.arange(n) * dt
        t1, t2 = max(self.p.tau1, 1e-20), max(self.p.tau2, 1e-20)
        h = (t1 ** 2 + t2 ** 2) / (t1 * t2 ** 2) * np.exp(-t / t2) * np.sin(t / t1)
        h[0] = 0.0
        H = fft(h) * dt
        return H / H[0] if abs(H[0]) > 1e-15 else np.ones(n, complex)

    def lin(self, A, h):
        return ifft(fft(A) * np.exp(h * self.D))

    def plasma_term(self, A, dt):
        pl = self.p.plasma
        area = pl.area if pl.area is not None else self.beam_area
        I = np.abs(A).astype(np.float64) ** 2 / area
        w_ion = (pl.sigma_K * I ** pl.K).astype(np.float32)
        sig = pl.drude(self.omega0)
        w_ava = ((sig / pl.Ui) * I if pl.avalanche else np.zeros_like(I)).astype(np.float32)
        rho, gen = _k1.k1(w_ion, w_ava, 1.0 / pl.tau_rec, dt, pl.rho_nt)
        loss = -(gen * pl.Ui) / (2.0 * I + 1e-30) * A
        defocus = -(sig / 2.0) * (1.0 + 1j * self.omega0 * pl.tau_c) * rho * A
        return loss + defocus

    def nonlin(self, A):
        I = np.abs(A) ** 2
        if self.H_raman is not None:
            I = (1 - self.p.f_R) * I + self.p.f_R * ifft(fft(I) * self.H_raman)
        P = A * I
        if self.p.use_self_steepening:
            out = ifft(1j * self.gamma * self.uppe * fft(P))
        else:
            out = 1j * self.gamma * P
        if self.p.plasma is not None:
            out = out + self.plasma_term(A, self.dt)
        return out.astype(A.dtype, copy=False)

    # -- integrator ---------------------------------------------------------------------
    def step(self, A, h, a5_prev=None, h_prev=None):
        """One ERK4(3)-IP attempt (erk43ip_method.py:256-286). Returns A4, err, alpha5'."""
        v = self.lin(A, h / 2)
        if a5_prev is not None and h_prev is not None and abs(h - h_prev) < 1e-10:
            a1 = self.lin(a5_prev, h / 2)
        else:
            a1 = self.lin(self.nonlin(A), h / 2)
        a2 = self.nonlin(v + h / 2 * a1)
        a3 = self.nonlin(v + h / 2 * a2)
        a4 = self.nonlin(self.lin(v + h * a3, h / 2))
        beta = self.lin(v + h / 6 * (a1 + 2 * a2 + 2 * a3), h / 2)
        A4 = beta + h / 6 * a4
        a5 = self.nonlin(A4)
        A3 = beta + (h / 30) * (2 * a4 + 3 * a5)
        nrm = np.sqrt(np.sum(np.abs(A4) ** 2))
        if nrm < 1e-20:
            nrm = 1.0
        err = np.sqrt(np.sum(np.abs(A4 - A3) ** 2)) / nrm
        return A4, err, a5

    def propagate(self, A0, t, L, tol=1e-4, max_step=1e-2, min_step=1e-11, replay=None,
                  max_attempts=None):
        """Adaptive loop of erk43ip_method.py:288-369.

        Returns (z_hist, A_hist, omega, log).  ``log`` lists every attempt as
        (h, err, accepted) - a forced min-step attempt follows its rejected one with
        accepted = 2.  With ``replay`` (a log from another run) the attempts and their
        accept decisions are taken from it instead of the controller, so two runs can be
        compared with identical step sequences.
        """
        n = len(t)
        self.dt = t[1] - t[0]
        omega = 2 * np.pi * fftfreq(n, self.dt)
        self.setup(omega)
        z, h = 0.0, max_step / 100
        A = np.array(A0, dtype=np.complex64 if self.precision == "single" else complex)
        zh, Ah = [0.0], [A.copy()]
        e_prev = a5_prev = h_prev = None
        log = []

        def record():
            if len(zh) < 1000 or (z - zh[-1]) > L / 500:
                zh.append(z)
                Ah.append(A.copy())

        if replay is not None:
            for (hr, _, acc) in replay:
                A_new, err, a5 = self.step(A, hr, a5_prev, h_prev)
                log.append((hr, err, acc))
                if acc:
                    z += hr
                    A, a5_prev, h_prev = A_new, a5, hr
                    record()
            return np.array(zh), np.array(Ah), omega, log

        while z < L:
            if max_attempts is not None and len(log) >= max_attempts:
                break
            if z + h > L:
                h = L - z
            A_new, err, a5 = self.step(A, h, a5_prev, h_prev)
            if err < tol:
                log.append((h, err, 1))
                z += h
                A, a5_prev, h_prev = A_new, a5, h
                record()
                fac = 0.95 * (tol / (err + 1e-30)) ** 0.25
                if e_prev is not None and e_prev > 1e-30:
                    fac *= (e_prev / (err + 1e-30)) ** 0.1
                e_prev = err
                h = min(h * min(fac, 2.0), max_step)
            else:
                log.append((h, err, 0))
                fac = 0.95 * (tol / (err + 1e-30)) ** 0.25
                h = max(h * max(fac, 0.1), min_step)
                if h <= min_step:
                    A_new, err, a5 = self.step(A, h, a5_prev, h_prev)
                    log.append((h, err, 2))
                    z += h
                    A, a5_prev, h_prev = A_new, a5, h
                    e_prev = None
                    record()
        return np.array(zh), np.array(Ah), omega, log


# -- pulse generators (erk43ip_method.py:375-446) -------------------------------------------
def gaussian_pulse(t, energy, fwhm, chirp=0.0):
    p0 = energy / (fwhm * np.sqrt(np.pi / (4 * np.log(2))))
    t0 = fwhm / (2 * np.sqrt(np.log(2)))
    return np.sqrt(p0) * np.exp(-(1 + 1j * chirp) * (t / t0) ** 2 / 2)


def dispersed_pulse(t, energy, fwhm, GDD=0.0, TOD=0.0):
    p0 = energy / (fwhm * np.sqrt(np.pi / (4 * np.log(2))))
    t0 = fwhm / (2 * np.sqrt(np.log(2)))
    a = np.sqrt(p0) * np.exp(-(t / t0) ** 2 / 2)
    if GDD == 0 and TOD == 0:
        return a
    n = len(t)
    w = 2 * np.pi * fftfreq(n, t[1] - t[0])
    spec = fft(ifftshift(a)) * np.exp(1j * (0.5 * GDD * w ** 2 + TOD * w ** 3 / 6.0))
    return fftshift(ifft(spec))
