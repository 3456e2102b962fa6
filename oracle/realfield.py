"""ORACLE (test infrastructure only): FP64 NumPy restatement of the real-field (carrier-resolved)
3+1D UPPE of ``src/3+1D-UPPE-ERK4(3)IP-real-field/``.  Never imported by the product package;
only ``tests/`` use it, as the checker.

Follows:

* ``media_physics.py:185-245``  PPT ionisation rate (atomic units, 15 ATI channels)
* ``media_physics.py:167-183``  the rate table: 2000 points of log10(I), 1e12-1e20 W/m^2, kept
                                in float32 as the reference's GPU copy; np.interp in log10(I)
* ``media_physics.py:247-252``  avalanche rate sigma / Ui * I (fused silica only)
* ``media_physics.py:11-118``   plasma density (exponential integrator with depletion, clamp)
                                and free-electron Drude current, ionisation-loss current
                                gen * Ui * E / (E^2 + 1e-20); constants rounded to float32 as
                                the reference passes them to its kernel, arithmetic in FP64
* ``media_physics.py:258-318``  P = eps0 (chi3 E^3 + chi5 E^5), I = 0.5 eps0 c n0 E^2, the two
                                coefficients rounded to float32 as the reference applies them
* ``media_physics.py:321-626``  FusedSilica, Air, He, Ne, Ar, Kr, Xe parameters
* ``solver_3d.py:54-111``       D = -i (kz - k1 omega), evanescent kz = -i sqrt|kz^2|,
                                P / J coefficients with the omega = 0 column zeroed, absorber
* ``solver_3d.py:113-172``      ERK4(3)-IP step with the low-order estimate
                                E_low = L(v + h/2 a1 + h/2 a2), err = max|E4 - E_low| / max|E4|
* ``solver_3d.py:174-241``      propagate: E_scale, controller (0.9 (tol/err)^0.333, fail_count)

Pinned in tests/test_oracle_realfield.py: the density recurrence against its closed form, the
Drude current against the DC conductivity e^2 rho tau_c / m, the ionisation-loss energy
balance, the linear operator (pure diffraction in a constant-index gas, group-velocity frame),
pulse broadening by the Sellmeier GDD, and the self-phase shift k0 n2 I L of the chi3 term.
The PPT rate itself is "parity unpinned" beyond its multiphoton-order slope at low intensity
and its agreement with the product's table.
"""
import numpy as np

C_LIGHT = 299792458.0
E_CHARGE = 1.60217663e-19
EPS0 = 8.85418781e-12
M_E = 9.1093837e-31
F32 = np.float32


def ppt_rate(I, omega0, Ui, n0, Z=1.0, l=0, m=0):
    """media_physics.py:185-245 (I in W/m^2, rate in 1/s)."""
    from scipy.special import gamma
    I = np.maximum(np.asarray(I, np.float64), 1e4)
    Eau = np.sqrt(2 * I / (C_LIGHT * EPS0 * n0)) / 5.142206538e11
    wau = omega0 * 2.418884e-17
    U = Ui / (27.21138 * E_CHARGE)
    m = abs(m)
    g = wau * np.sqrt(2 * U) / Eau
    ns = Z / np.sqrt(2 * U)
    C2 = 2 ** (2 * ns) / (ns * gamma(ns + l + 1) * gamma(ns - l))
    f = (2 * l + 1) * 0.5 * gamma(l + m + 1) / (2 ** m * gamma(m + 1) * gamma(l - m + 1))
    alpha = 2 * g ** 2 / (1 + g ** 2)
    gg = 3 / (2 * g) * ((1 + 1 / (2 * g ** 2)) * np.arcsinh(g) - np.sqrt(1 + g ** 2) / (2 * g))
    F = 2 * (2 * U) ** 1.5 / Eau
    pre = U * C2 * f * np.sqrt(6 / np.pi) * F ** (2 * ns - m - 1.5) * (1 + g ** 2) ** (m / 2 + 0.75 - ns)
    nu = (U / wau) * (1 + 1 / (2 * g ** 2))
    A = np.zeros_like(Eau)
    for k in range(15):
        A = A + np.exp(-alpha * 2 * (np.ceil(nu) + k - nu) * np.sqrt(1 + g ** 2) / (g ** 2 + 1e-20))
    return pre * A * np.exp(-F / 3 * gg) / 2.418884e-17


# name: (index model, n0 or Cauchy (A, B, C), n2 per atm, n4 per atm^2, Ui eV, rho_nt per atm,
#        tau_rec, tau_c * atm, avalanche)
MEDIA = {
    "FusedSilica": ("sellmeier", None, 3.0e-20, 0.0, 9.0, 2.1e28, 150e-15, 3.0e-15, True),
    "Air": ("const", None, 3.0e-23, 0.0, 12.1, 2.7e25, 1.0e-9, 350e-15, False),
    "Helium": ("cauchy", (3.48e-5, 5.4e-8, 1.2e-10), 4.0e-25, -1.0e-38, 24.59, 2.45e25, 1e-9, 500e-15, False),
    "Neon": ("cauchy", (6.66e-5, 2.4e-8, 1.7e-10), 2.0e-24, -5.0e-38, 21.56, 2.45e25, 1e-9, 350e-15, False),
    "Argon": ("cauchy", (27.9e-5, 12.8e-8, 5.2e-10), 1.0e-23, -1.0e-36, 15.76, 2.45e25, 1e-9, 190e-15, False),
    "Krypton": ("cauchy", (41.9e-5, 21.3e-8, 8.4e-10), 2.8e-23, -2.5e-36, 14.00, 2.45e25, 1e-9, 150e-15, False),
    "Xenon": ("cauchy", (68.7e-5, 40.6e-8, 16.3e-10), 6.5e-23, -8.0e-36, 12.13, 2.45e25, 1e-9, 120e-15, False),
}
SELLMEIER_FS = ((0.6961663, 0.4079426, 0.8974794), (0.0684043 ** 2, 0.1162414 ** 2, 9.896161 ** 2))


class RFMedium:
    def __init__(self, name, center_wavelength, pressure_atm=1.0, lut_points=2000):
        kind, cau, n2, n4, ui, rho, trec, tc, aval = MEDIA[name]
        P = 1.0 if name == "FusedSilica" else pressure_atm
        self.name, self.kind, self.P = name, kind, P
        self.lambda0 = center_wavelength
        self.c = C_LIGHT
        self.omega0 = 2 * np.pi * C_LIGHT / center_wavelength
        self.cauchy = cau
        if kind == "sellmeier":
            self.n0 = 1.45
        elif kind == "const":
            self.n0 = 1.0 + 0.00027 * P
        else:
            self.n0 = 1.0 + cau[0] * P
        self.n2 = n2 * P
        self.n4 = n4 * P * P
        self.chi3 = (4.0 / 3.0) * C_LIGHT * EPS0 * self.n0 ** 2 * self.n2
        self.chi5 = (4.0 / 5.0) * C_LIGHT ** 2 * EPS0 ** 2 * self.n0 ** 3 * self.n4
        self.Ui = ui * E_CHARGE
        self.rho_nt = rho * (1.0 if name == "FusedSilica" else P)
        self.tau_rec = trec
        self.tau_c = tc / (1.0 if name == "FusedSilica" else P)
        self.avalanche = aval
        lg = np.linspace(12.0, 20.0, lut_points)
        self.lut_x = lg.astype(F32)
        self.lut_w = ppt_rate(10 ** lg, self.omega0, self.Ui, self.n0).astype(F32)
        self.plasma = True

    def refractive_index(self, omega):
        w = np.asarray(omega, np.float64)
        if self.kind == "const":
            return np.ones_like(w) * self.n0
        w = np.where(np.abs(w) < 1e-12, 1e-12, w)
        if self.kind == "sellmeier":
            lam = np.clip(2 * np.pi * C_LIGHT / np.abs(w), 0.2e-6, 6.0e-6)
            l2 = (lam * 1e6) ** 2
            B, Cc = SELLMEIER_FS
            return np.sqrt(1.0 + sum(b * l2 / (l2 - c) for b, c in zip(B, Cc)))
        l2 = (2 * np.pi * C_LIGHT / np.abs(w) * 1e6) ** 2
        A, B, Cc = self.cauchy
        return 1.0 + (A + B / l2 + Cc / l2 ** 2) * self.P

    def av_coef(self):
        if not self.avalanche:
            return 0.0
        sig = (E_CHARGE ** 2 * self.tau_c) / (
            C_LIGHT * EPS0 * self.n0 * M_E * (1.0 + self.omega0 ** 2 * self.tau_c ** 2))
        return sig / self.Ui

    def rates(self, I):
        """(W_ion, W_ava) at intensities I: float32 results, as the reference casts them."""
        lg = np.log10(I + 1e-30)
        lg = np.clip(lg, float(self.lut_x[0]), float(self.lut_x[-1]))
        wi = np.interp(lg, self.lut_x.astype(np.float64), self.lut_w.astype(np.float64))
        wa = self.av_coef() * I
        return wi.astype(F32), np.asarray(wa).astype(F32)

    def plasma_current(self, E, dt):
        """media_physics.py:11-118 over lines E[..., t] (physical V/m); returns (J, rho)."""
        E = np.asarray(E, np.float64)
        if not self.plasma:     # test switch: no ionisation at all
            return np.zeros_like(E), np.zeros_like(E)
        I = 0.5 * EPS0 * C_LIGHT * self.n0 * E ** 2
        wi, wa = self.rates(I)
        rnt = float(F32(self.rho_nt))
        rrec = float(F32(1.0 / self.tau_rec))
        dtd = float(F32(dt))
        Ui = float(F32(self.Ui))
        e, me, tc = float(F32(E_CHARGE)), float(F32(M_E)), float(F32(self.tau_c))
        nu = 1.0 / tc
        e2m = e * e / me
        nudt = nu * dtd
        enu = np.exp(-nudt)
        sJ = dtd * (1.0 - 0.5 * nudt) if abs(nudt) < 1e-5 else (1.0 - enu) / nu
        rho = np.zeros(E.shape[:-1])
        jf = np.zeros(E.shape[:-1])
        J = np.zeros_like(E)
        R = np.zeros_like(E)
        for t in range(E.shape[-1]):
            Et = E[..., t]
            wpi = wi[..., t].astype(np.float64)
            wav = wa[..., t].astype(np.float64)
            dep = np.maximum((rnt - rho) / rnt, 0.0)
            S = wpi * rnt * dep
            Gm = wav * dep - rrec
            gdt = np.clip(Gm * dtd, -10.0, 10.0)
            eg = np.exp(gdt)
            with np.errstate(divide="ignore", invalid="ignore"):
                sf = np.where(np.abs(gdt) < 1e-5, dtd * (1.0 + 0.5 * gdt), (eg - 1.0) / Gm)
            nr = np.clip(rho * eg + S * sf, 0.0, rnt)
            gen = S + wav * dep * rho
            nj = jf * enu + e2m * rho * Et * sJ
            jl = np.where(np.abs(Et) > 1e-10, gen * Ui * Et / (Et * Et + 1e-20), 0.0)
            J[..., t] = jf + jl
            R[..., t] = rho
            rho, jf = nr, nj
        return J, R

    def response(self, E, dt):
        """(P, J) of media_physics.py:258-318 in FP64, with the coefficients eps0 chi3 and
        eps0 chi5 rounded to float32: the reference multiplies its float32 field by these Python
        floats, which CuPy casts to float32.  For the gases eps0 chi5 (~1e-53) is below the
        float32 range, so the reference's fifth-order term is exactly zero."""
        J, _ = self.plasma_current(E, dt)
        P = float(F32(EPS0 * self.chi3)) * E ** 3
        c5 = float(F32(EPS0 * self.chi5))
        if c5 != 0.0:
            P = P + c5 * E ** 5
        return P, J


class OracleRF:
    def __init__(self, x, y, t, medium, absorber=True):
        self.m = medium
        self.x, self.y, self.t = (np.asarray(a, np.float64) for a in (x, y, t))
        self.Nx, self.Ny, self.Nt = len(x), len(y), len(t)
        self.dx, self.dy = float(x[1] - x[0]), float(y[1] - y[0])
        self.dt = float(t[1] - t[0])
        kx = 2 * np.pi * np.fft.fftfreq(self.Nx, self.dx)
        ky = 2 * np.pi * np.fft.fftfreq(self.Ny, self.dy)
        self.K2 = kx[:, None] ** 2 + ky[None, :] ** 2
        self.omega = 2 * np.pi * np.fft.rfftfreq(self.Nt, self.dt)
        self.freq_axis_hz = self.omega / (2 * np.pi)
        w, c = self.omega, C_LIGHT
        ws = np.where(w == 0, 1e-12, w)
        k = medium.refractive_index(ws) * ws / c
        w0 = medium.omega0
        k0 = medium.refractive_index(np.array([w0]))[0] * w0 / c
        dw = w0 * 1e-3
        self.k1 = (medium.refractive_index(np.array([w0 + dw]))[0] * (w0 + dw) / c - k0) / dw
        kz2 = (k ** 2)[None, None, :] - self.K2[:, :, None]
        self.kz = np.where(kz2 >= 0, np.sqrt(np.abs(kz2)) + 0j, -1j * np.sqrt(np.abs(kz2)))
        self.D = -1j * (self.kz - (self.k1 * w)[None, None, :])
        kzs = np.where(np.abs(self.kz) < 1e-10, 1e-10, self.kz)
        self.Pc = -1j * w[None, None, :] ** 2 / (2 * EPS0 * c ** 2 * kzs)
        self.Jc = -w[None, None, :] / (2 * EPS0 * c ** 2 * kzs)
        self.Pc[:, :, 0] = 0.0
        self.Jc[:, :, 0] = 0.0
        X, Y = np.meshgrid(self.x, self.y, indexing="ij")
        R = np.sqrt(X ** 2 + Y ** 2)
        hw = float(np.max(np.abs(self.x)))
        s = np.clip((R - 0.85 * hw) / (0.98 * hw - 0.85 * hw + 1e-30), 0.0, 1.0)
        self.damp = (120.0 * s ** 10)[:, :, None] * (1.0 if absorber else 0.0)
        f = self.freq_axis_hz
        self.filt = np.exp(-(f / (0.85 * (np.max(f) + 1e-10))) ** 16)[None, None, :]
        self.E_scale = 1.0

    def N(self, Fk):
        Et = np.fft.irfft(np.fft.ifft2(Fk, axes=(0, 1)), n=self.Nt, axis=-1)
        P, J = self.m.response(Et * self.E_scale, self.dt)
        Pw = np.fft.rfft(P / self.E_scale, axis=-1)
        Jw = np.fft.rfft(J / self.E_scale, axis=-1)
        return np.fft.fft2(self.Pc * Pw + self.Jc * Jw, axes=(0, 1))

    def step(self, E_t, h):
        """solver_3d.py:113-172.  Returns (E4_t, err)."""
        Ek = np.fft.fft2(np.fft.rfft(E_t, axis=-1), axes=(0, 1))
        L = np.exp((h / 2) * self.D)
        v = Ek * L
        a1 = self.N(Ek) * L
        a2 = self.N(v + (h / 2) * a1)
        a3 = self.N(v + (h / 2) * a2)
        a4 = self.N((v + h * a3) * L)
        E4 = (v + (h / 6) * (a1 + 2 * a2 + 2 * a3)) * L + (h / 6) * a4
        El = (v + (h / 2) * a1 + (h / 2) * a2) * L
        err = float(np.max(np.abs(E4 - El)) / (np.max(np.abs(E4)) + 1e-20))
        E4w = np.fft.ifft2(E4, axes=(0, 1)) * self.filt
        return np.fft.irfft(E4w, n=self.Nt, axis=-1) * np.exp(-self.damp * h), err

    def propagate(self, E0, L, tol=1e-4, max_step=1e-2, min_step=1e-9, replay=None,
                  max_attempts=None, return_time=False):
        """solver_3d.py:174-241.  Returns (z, freq_axis_hz, E_w, log)."""
        E0 = np.asarray(E0, np.float64)
        s = float(np.max(np.abs(E0)))
        self.E_scale = s if s >= 1e-30 else 1.0
        E = E0 / self.E_scale
        z, log = 0.0, []
        if replay is not None:
            for (h, _, acc) in replay:
                En, err = self.step(E, h)
                log.append((h, err, acc))
                if acc:
                    z += h
                    E = En
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
                En, err = self.step(E, h)
                if err < tol or h <= min_step * 1.01:
                    log.append((h, err, 1))
                    fail = fail + 1 if (h <= min_step * 1.01 and err > tol) else 0
                    z += h
                    E = En
                    e = max(err, 1e-30)
                    fac = 0.9 * (tol / e) ** 0.333
                    if fail > 0:
                        fac = min(fac, 1.2)
                    h = max(min(h * min(fac, 2.0), max_step), min_step)
                else:
                    log.append((h, err, 0))
                    e = max(err, 1e-30)
                    h = max(h * 0.9 * (tol / e) ** 0.333, min_step)
        E = E * self.E_scale
        if return_time:
            return z, E, log
        return z, self.freq_axis_hz, np.fft.rfft(E, axis=-1), log
