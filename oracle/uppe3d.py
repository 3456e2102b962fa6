"""ORACLE (test infrastructure only): FP64 NumPy restatement of the reference's 3+1D envelope
ERK4(3)-IP UPPE solver and its gas media, with ``precision='double'`` semantics.

Never imported by the product package.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` use it, as the checker.  The reference itself imports cupy at
module top and cannot run here; this file is the "reference FP64 CPU solver" for 3D.

Follows (reference paths relative to the reference root, directory
``src/3+1D-UPPE-ERK4(3)IP-envelope/``):

* ``media_physics.py:72-104``   OpticalMedium constants
* ``media_physics.py:113-119``  shock operator
* ``media_physics.py:121-146``  Keldysh MPI + tunnel LUT, log10 interpolation
* ``media_physics.py:152-154``  base avalanche rate (Argon and the other gases: zero; fused
  silica keeps it)
* ``media_physics.py:208-250``  FusedSilica (Sellmeier, n0 = 1.45; Raman not restated)
* ``media_physics.py:156-205``  nonlinear response; W cast to float32, K1 (``oracle/k1.py``)
* ``media_physics.py:282-510``  Air / He / Ne / Ar / Kr / Xe parameters and Cauchy index
* ``solver_3d.py:8-54``         grids, filter, valid-frequency mask
* ``solver_3d.py:56-87``        dispersion operator (forward-difference k1, difference of squares)
* ``solver_3d.py:89-98``        absorber
* ``solver_3d.py:100-160``      ERK4(3)-IP step in (kx, ky, omega), L-infinity error
* ``solver_3d.py:162-233``      propagate: A_scale, controller with fail_count, output spectrum

One deliberate difference: the transverse wavenumber is K_perp^2[x, y] = kx[x]^2 + ky[y]^2.
The reference builds it with ``meshgrid(ky, kx, indexing='ij')`` (``solver_3d.py:30``), i.e.
transposed; the two agree whenever Nx == Ny and dx == dy (the only square case it supports).
"""
from dataclasses import dataclass

import numpy as np
import scipy.fft as sfft

from . import k1 as _k1

C_LIGHT = 299792458.0
E_CHARGE = 1.60217663e-19
EPS0 = 8.85418781e-12
M_E = 9.1093837e-31
HBAR = 1.054571817e-34

# media_physics.py:282-510: (cauchy A, B [um^2], C [um^4], n2/atm, n4/atm^2, Ui [eV], K,
# sigma_K/atm, tau_c*atm); rho_nt = 2.45e25 * P, tau_rec = 1 ns for all noble gases.
GASES = {
    "Helium": (3.48e-5, 5.4e-8, 1.2e-10, 4.0e-25, -1.0e-38, 24.59, 17, 0.0, 500e-15),
    "Neon": (6.66e-5, 2.4e-8, 1.7e-10, 2.0e-24, -5.0e-38, 21.56, 15, 0.0, 350e-15),
    "Argon": (27.9e-5, 12.8e-8, 5.2e-10, 1.0e-23, -1.0e-36, 15.76, 11, 1.0e-188, 190e-15),
    "Krypton": (41.9e-5, 21.3e-8, 8.4e-10, 2.8e-23, -2.5e-36, 14.00, 10, 1.0e-168, 150e-15),
    "Xenon": (68.7e-5, 40.6e-8, 16.3e-10, 6.5e-23, -8.0e-36, 12.13, 8, 1.0e-130, 120e-15),
}


@dataclass
class GasMedium:
    """A noble-gas medium of the reference (double precision), parameters by name."""
    name: str = "Argon"
    center_wavelength: float = 800e-9
    pressure_atm: float = 1.0
    use_shock: bool = False
    n4_enabled: bool = True          # False drops the HOKE term (test variants only)
    lut_points: int = 2000
    tunnel_guard: float = 1e-20      # the reference's "+ 1e-20" in the tunnel exponent

    def __post_init__(self):
        A, B, Cc, n2, n4, ui, K, sk, tc = GASES[self.name]
        P = self.pressure_atm
        self.lambda0 = self.center_wavelength
        self.c = C_LIGHT
        self.omega0 = 2 * np.pi * C_LIGHT / self.lambda0
        self.cauchy = (A, B, Cc)
        self.n0 = 1.0
        self.n2 = n2 * P
        self.n4 = n4 * P * P if self.n4_enabled else 0.0
        self.Ui = ui * E_CHARGE
        self.rho_nt = 2.45e25 * P
        self.K = K
        self.tau_rec = 1.0e-9
        self.tau_c = tc / P
        self.sigma_drude = (E_CHARGE ** 2 * self.tau_c) / (
            C_LIGHT * EPS0 * 1.0 * M_E * (1.0 + self.omega0 ** 2 * self.tau_c ** 2))
        self.sigma_K = sk * P
        lg = np.linspace(np.log10(1e12), np.log10(1e20), self.lut_points)
        self.lut_log10_I = lg
        self.lut_W = self.keldysh_rate(10 ** lg)

    def refractive_index(self, omega):
        w = np.asarray(omega, np.float64)
        w = np.where(np.abs(w) < 1e-12, 1e-12, w)
        l2 = (2 * np.pi * C_LIGHT / np.abs(w) * 1e6) ** 2
        A, B, Cc = self.cauchy
        return 1.0 + (A + B / l2 + Cc / l2 ** 2) * self.pressure_atm

    def keldysh_rate(self, I):
        w_mpi = self.sigma_K * I ** self.K
        E = np.sqrt(2 * I / (C_LIGHT * EPS0 * self.n0))
        ex = -(4.0 / 3.0) * np.sqrt(2 * M_E) * self.Ui ** 1.5 / (E_CHARGE * HBAR * E + self.tunnel_guard)
        return w_mpi + 1e15 * np.exp(ex)

    def ionization_rate(self, I):
        lg = np.clip(np.log10(I + 1e-30), self.lut_log10_I[0], self.lut_log10_I[-1])
        return np.interp(lg, self.lut_log10_I, self.lut_W)

    def shock(self, omega):
        return (omega + self.omega0) / self.omega0 if self.use_shock else np.ones_like(omega)

    def avalanche_rate(self, I):
        """The gases override the base rate with zero (media_physics.py:431-432 and siblings)."""
        return np.zeros(np.shape(I))

    def response(self, A, dt):
        """(Kerr, Plasma) of media_physics.py:156-205 in double mode."""
        I = np.abs(A) ** 2
        w_ion = self.ionization_rate(I).astype(np.float32)
        w_ava = self.avalanche_rate(I).astype(np.float32)
        rho, gen = _k1.k1(w_ion, w_ava, 1.0 / self.tau_rec, dt, self.rho_nt)
        g = 1j * (self.omega0 / C_LIGHT)
        kerr = g * self.n2 * I * A
        if self.n4 != 0.0:
            kerr = kerr + g * self.n4 * I * I * A
        loss = -(gen * self.Ui) / (2.0 * I + 1e-30) * A
        defo = -(self.sigma_drude / 2.0) * (1.0 + 1j * self.omega0 * self.tau_c) * rho * A
        # both terms in the field's precision (the reference casts to its complex dtype)
        return kerr.astype(A.dtype), (loss + defo).astype(A.dtype)


@dataclass
class FusedSilicaMedium(GasMedium):
    """FusedSilica of media_physics.py:208-250 (no Raman): Sellmeier index clamped to
    0.2-6 um, n0 = 1.45 in the Keldysh field and the Drude cross section, and the base class's
    avalanche rate (sigma_drude / Ui) I (media_physics.py:152-154), which the solid keeps."""
    name: str = "FusedSilica"

    def __post_init__(self):
        self.lambda0 = self.center_wavelength
        self.c = C_LIGHT
        self.omega0 = 2 * np.pi * C_LIGHT / self.lambda0
        self.n0 = 1.45
        self.n2 = 3.0e-20
        self.n4 = 0.0
        self.B = (0.6961663, 0.4079426, 0.8974794)
        self.C = (0.0684043 ** 2, 0.1162414 ** 2, 9.896161 ** 2)
        self.valid_range = (0.2e-6, 6.0e-6)
        self.Ui = 9.0 * E_CHARGE
        self.rho_nt = 2.1e28
        self.K = 8
        self.tau_rec = 150e-15
        self.tau_c = 3.0e-15
        self.sigma_drude = (E_CHARGE ** 2 * self.tau_c) / (
            C_LIGHT * EPS0 * self.n0 * M_E * (1.0 + self.omega0 ** 2 * self.tau_c ** 2))
        self.sigma_K = 2.5e-125
        lg = np.linspace(np.log10(1e12), np.log10(1e20), self.lut_points)
        self.lut_log10_I = lg
        self.lut_W = self.keldysh_rate(10 ** lg)

    def refractive_index(self, omega):
        w = np.asarray(omega, np.float64)
        w = np.where(np.abs(w) < 1e-12, 1e-12, w)
        lam = np.clip(2 * np.pi * C_LIGHT / np.abs(w), *self.valid_range)
        l2 = (lam * 1e6) ** 2
        n2 = np.ones_like(l2)
        for b, c in zip(self.B, self.C):
            n2 = n2 + b * l2 / (l2 - c)
        return np.sqrt(n2)

    def avalanche_rate(self, I):
        return (self.sigma_drude / self.Ui) * np.asarray(I, np.float64)


def dispersion_operator(medium, omega, kperp2, valid_mask=True):
    """D(x-index, y-index, omega) of solver_3d.py:56-87 (FP64)."""
    w0 = medium.omega0
    n_w = medium.refractive_index(omega + w0)
    k = n_w * (omega + w0) / C_LIGHT
    n_c = float(medium.refractive_index(np.array([w0]))[0])
    dw = w0 * 1e-3
    n_p = float(medium.refractive_index(np.array([w0 + dw]))[0])
    k0 = n_c * w0 / C_LIGHT
    k1 = (n_p * (w0 + dw) / C_LIGHT - k0) / dw
    kref = k0 + k1 * omega
    ksq = (k ** 2)[None, None, :]
    kp = kperp2[:, :, None]
    kz = np.sqrt((ksq - kp).astype(np.complex128))
    D = 1j * (((ksq - kref[None, None, :] ** 2) - kp) / (kz + kref[None, None, :]))
    if valid_mask:
        D = D * (omega + w0 > 0)[None, None, :]
    return D, k, kref, k0, k1


class Oracle3D:
    # error norm of the embedded estimate: "linf" (solver_3d.py:151-153) or "l2" (the older
    # solver of src/3+1D-ERK4(3)IP UPPE/solver_3d.py:252-254); valid_mask False drops the
    # negative-absolute-frequency mask from D and N (that older solver has none)
    err_norm = "linf"
    valid_mask = True

    def __init__(self, x, y, t, medium, workers=None, precision="double"):
        """precision "single" restates the reference's default arithmetic (solver_3d.py:14-19,
        35, 57): omega rounded to float32 before the FP64 operator build, operators, field and
        nonlinear terms in complex64 / float32 (the plasma recurrence stays FP64, as in its
        kernel).  It measures what FP32 arithmetic costs the reference itself."""
        self.m = medium
        self.workers = workers
        self.precision = precision
        self.x, self.y, self.t = (np.asarray(a, np.float64) for a in (x, y, t))
        self.Nx, self.Ny, self.Nt = len(x), len(y), len(t)
        self.dx, self.dy = float(x[1] - x[0]), float(y[1] - y[0])
        self.dt = float(t[1] - t[0])
        kx = 2 * np.pi * sfft.fftfreq(self.Nx, self.dx)
        ky = 2 * np.pi * sfft.fftfreq(self.Ny, self.dy)
        self.kperp2 = kx[:, None] ** 2 + ky[None, :] ** 2
        self.omega = 2 * np.pi * sfft.fftfreq(self.Nt, self.dt)
        if precision == "single":
            self.omega = self.omega.astype(np.float32).astype(np.float64)
        self.freq_axis_hz = sfft.fftshift(self.omega) / (2 * np.pi)
        self.D, self.k, self.kref, self.k0, self.k1 = dispersion_operator(
            medium, self.omega, self.kperp2, self.valid_mask)
        X, Y = np.meshgrid(self.x, self.y, indexing="ij")
        R = np.sqrt(X ** 2 + Y ** 2)
        hw = float(np.max(np.abs(self.x)))
        s = np.clip((R - 0.85 * hw) / (0.98 * hw - 0.85 * hw + 1e-30), 0.0, 1.0)
        self.damp = (120.0 * s ** 10)[:, :, None]
        self.shock = medium.shock(self.omega)[None, None, :]
        f = self.omega / (2 * np.pi)
        fmax = float(np.max(np.abs(f))) + 1e-10
        self.filt = np.exp(-(f / (0.85 * fmax)) ** 16)[None, None, :]
        self.mask = (self.omega + medium.omega0 > 0).astype(np.float64)[None, None, :]
        if not self.valid_mask:
            self.mask = np.ones_like(self.mask)
        if precision == "single":
            self.D = self.D.astype(np.complex64)
            for name in ("damp", "shock", "filt", "mask"):
                setattr(self, name, getattr(self, name).astype(np.float32))
        self.A_scale = 1.0

    # FFT helpers (backward norm, as scipy/cupyx)
    def _f(self, a, axes):
        return sfft.fftn(a, axes=axes, workers=self.workers)

    def _fi(self, a, axes):
        return sfft.ifftn(a, axes=axes, workers=self.workers)

    def N(self, Fk):
        At = self._fi(Fk, (0, 1, 2)) * self.A_scale
        kerr, plas = self.m.response(At, self.dt)
        kw = self._f(kerr / self.A_scale, (2,))
        pw = self._f(plas / self.A_scale, (2,))
        return self._f((self.shock * kw + pw) * self.mask, (0, 1))

    def step(self, A_t, h):
        """solver_3d.py:100-160. Returns (A4_t, err)."""
        Ak = self._f(A_t, (0, 1, 2))
        E = np.exp((h / 2) * self.D)
        v = Ak * E
        a1 = self.N(Ak) * E
        a2 = self.N(v + (h / 2) * a1)
        a3 = self.N(v + (h / 2) * a2)
        a4 = self.N((v + h * a3) * E)
        beta = (v + (h / 6) * (a1 + 2 * a2 + 2 * a3)) * E
        A4 = beta + (h / 6) * a4
        a5 = self.N(A4)
        A3 = beta + (h / 30) * (2 * a4 + 3 * a5)
        if self.err_norm == "l2":
            err = float(np.sqrt(np.sum(np.abs(A4 - A3) ** 2)) /
                        (np.sqrt(np.sum(np.abs(A4) ** 2)) + 1e-20))
        else:
            err = float(np.max(np.abs(A4 - A3)) / (np.max(np.abs(A4)) + 1e-20))
        A4w = self._fi(A4, (0, 1)) * self.filt
        A4t = self._fi(A4w, (2,)) * np.exp(-self.damp * h)
        return A4t, err

    def propagate(self, A0, L, tol=1e-4, max_step=1e-2, min_step=1e-9, replay=None,
                  max_attempts=None, return_time=False):
        """solver_3d.py:162-233. Returns (z, freq_axis_hz, A_w, log); log = [(h, err, acc)].

        With ``replay`` the attempts and accept decisions come from another run's log.
        """
        A0 = np.asarray(A0, np.complex64 if self.precision == "single" else np.complex128)
        s = float(np.max(np.abs(A0)))
        self.A_scale = s if s >= 1e-30 else 1.0
        A = A0 / self.A_scale
        z = 0.0
        log = []
        if replay is not None:
            for (h, _, acc) in replay:
                A_new, err = self.step(A, h)
                log.append((h, err, acc))
                if acc:
                    z += h
                    A = A_new
        else:
            h = min(max_step, L / 1000.0)
            fail = 0
            while z < L:
                if max_attempts is not None and len(log) >= max_attempts:
                    break
                if h > L - z:
                    h = L - z
                    if h < 1e-15:
                        break
                A_new, err = self.step(A, h)
                if not np.isfinite(err):
                    raise FloatingPointError("non-finite error norm at z=%g" % z)
                if err < tol or h <= min_step * 1.01:
                    log.append((h, err, 1))
                    fail = fail + 1 if (h <= min_step * 1.01 and err > tol) else 0
                    z += h
                    A = A_new
                    e = max(err, 1e-30)
                    fac = 0.95 * (tol / e) ** 0.2
                    if fail > 0:
                        fac = min(fac, 1.2)
                    h = min(h * min(fac, 2.0), max_step)
                else:
                    log.append((h, err, 0))
                    e = max(err, 1e-30)
                    fac = 0.95 * (tol / e) ** 0.2
                    h = max(h * fac, min_step)
        A = A * self.A_scale
        if return_time:
            return z, A, log
        Aw = sfft.fftshift(self._f(A, (2,)), axes=-1)
        return z, self.freq_axis_hz, Aw, log
