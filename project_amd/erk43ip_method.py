"""1D envelope ERK4(3)-IP UPPE solver on MI355X (drop-in for
``src/1D-ERK4(3)IP UPPE gas/erk43ip_method.py:14-863``).

The class keeps the reference's name, constructor keywords, attributes and hook methods
(``_get_sellmeier_coeffs``, ``_compute_refractive_index``, ``_compute_dispersion_operator``,
``_update_raman_response``, the ``_cached_D`` / ``_cached_n_omega`` /
``_cached_raman_response`` caches).  The hooks run on the host in FP64 NumPy, once per omega
grid, exactly where the reference evaluates them, so class- and instance-level monkeypatches
(``gas.py:11-44``, ``test_3d-1d.py:52-61``) keep working.  ``propagate`` runs the whole
adaptive ERK4(3)-IP loop - steps, controller, error norm, history rule - in one HIP kernel
launch (libuppe.so), reading ``self.gamma`` at call time as callers expect
(``gas.py:120-126``).

Extension with no counterpart in the 1D reference: ``plasma=Plasma1D(...)`` adds the
multiphoton-ionisation plasma term of the 3+1D medium model (``media_physics.py:156-205``)
along t, with intensity |A|^2 / area and rate sigma_K I^K.
"""
import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np
from scipy.fft import fft, ifft, fftfreq, fftshift, ifftshift

from . import _lib

E_CHARGE = 1.60217663e-19
EPS0 = 8.85418781e-12
M_E = 9.1093837e-31


@dataclass
class Plasma1D:
    """Plasma model for the 1D solver.  Defaults: the reference Argon class at 1 atm
    (``media_physics.py:396-432``) with multiphoton order K = 8."""
    K: int = 8
    sigma_K: float = 1.0e-130
    Ui: float = 15.76 * E_CHARGE
    rho_nt: float = 2.45e25
    tau_rec: float = 1.0e-9
    tau_c: float = 190e-15
    sigma_drude: Optional[float] = None   # None: Drude formula with n0 = 1
    avalanche: bool = False               # gases: zero avalanche rate
    area: Optional[float] = None          # None: pi * beam_radius**2 of the solver

    def drude(self, omega0, c=299792458.0):
        if self.sigma_drude is not None:
            return self.sigma_drude
        return (E_CHARGE ** 2 * self.tau_c) / (
            c * EPS0 * 1.0 * M_E * (1.0 + omega0 ** 2 * self.tau_c ** 2))


class ERK43IP_UPPE:
    """ERK4(3)-IP UPPE solver (1D envelope), GPU-resident propagation."""

    MAX_HIST = 1502   # the reference history rule records at most 1000 + 500 + 1 snapshots

    def __init__(self, material='fused_silica',
                 gamma=None, n2=None, beam_radius=None,
                 alpha=0.0, center_wavelength=1064e-9,
                 use_raman=True, f_R=0.18, tau1=12.2e-15, tau2=32e-15,
                 use_self_steepening=True, plasma=None):
        self.material = material
        self.alpha = alpha
        self.lambda0 = center_wavelength
        self.c = 299792458.0
        self.omega0 = 2 * np.pi * self.c / self.lambda0
        if gamma is not None:
            self.gamma = gamma
            self.n2 = 0.0
            self.beam_radius = 0.0
        elif n2 is not None and beam_radius is not None:
            self.beam_radius = beam_radius
            self.n2 = n2
            self.beam_area = np.pi * beam_radius ** 2
            self.gamma = (2 * np.pi / self.lambda0) * (n2 / self.beam_area)
        else:
            self.gamma = 0.0
            self.n2 = 0.0
            self.beam_radius = 0.0
        self.use_raman = use_raman
        self.f_R = f_R
        self.tau1 = tau1
        self.tau2 = tau2
        self.use_self_steepening = use_self_steepening
        self.plasma = plasma
        self.sellmeier_coeffs = self._get_sellmeier_coeffs(material)
        self._cached_D = None
        self._cached_omega_grid = None
        self._cached_n_omega = None
        self._cached_raman_response = None
        self.n0 = self._compute_refractive_index(np.array([self.lambda0]))[0]
        self.last_status = None

    # -- material hooks (host, FP64; erk43ip_method.py:72-148) -----------------------------
    def _get_sellmeier_coeffs(self, material):
        coeffs = {
            'fused_silica': {'B': [0.6961663, 0.4079426, 0.8974794],
                             'C': [0.0684043 ** 2, 0.1162414 ** 2, 9.896161 ** 2],
                             'valid_range': (0.21e-6, 6.7e-6)},
            'sapphire': {'B': [1.4313493, 0.65054713, 5.3414021],
                         'C': [0.0726631 ** 2, 0.1193242 ** 2, 18.028251 ** 2],
                         'valid_range': (0.2e-6, 5.5e-6)},
            'yag': {'B': [2.282, 3.27644, 4.44174],
                    'C': [0.01185 ** 2, 0.0282 ** 2, 282.734 ** 2],
                    'valid_range': (0.4e-6, 5e-6)},
        }
        return coeffs.get(material, coeffs['fused_silica'])

    def _compute_refractive_index(self, wavelength):
        c = self.sellmeier_coeffs
        wl_sq = (np.asarray(wavelength, np.float64) * 1e6) ** 2
        n_sq = np.ones_like(wl_sq)
        for b, cc in zip(c['B'], c['C']):
            n_sq = n_sq + b * wl_sq / (wl_sq - cc)
        return np.sqrt(n_sq)

    def _compute_dispersion_operator(self, omega_grid):
        """D(omega) = i (k - k0 - k1 omega) - alpha/2, central-difference k1 (dw = omega0 1e-5);
        caches n(omega) for the UPPE factor."""
        omega_abs = self.omega0 + omega_grid
        omega_safe = np.where(np.abs(omega_abs) < 1e-10, 1e-10, np.abs(omega_abs))
        lo, hi = self.sellmeier_coeffs['valid_range']
        lam = np.clip(2 * np.pi * self.c / omega_safe, lo, hi)
        n_omega = self._compute_refractive_index(lam)
        self._cached_n_omega = n_omega
        k_omega = n_omega * omega_abs / self.c
        dw = self.omega0 * 1e-5
        n_c = self._compute_refractive_index(np.array([self.lambda0]))[0]
        n_p = self._compute_refractive_index(np.array([2 * np.pi * self.c / (self.omega0 + dw)]))[0]
        n_m = self._compute_refractive_index(np.array([2 * np.pi * self.c / (self.omega0 - dw)]))[0]
        k0 = n_c * self.omega0 / self.c
        k1 = (n_p * (self.omega0 + dw) / self.c - n_m * (self.omega0 - dw) / self.c) / (2 * dw)
        return 1j * (k_omega - k0 - k1 * omega_grid) - self.alpha / 2

    def _update_raman_response(self, omega_grid):
        """H_R(omega) normalised by H_R(0) (erk43ip_method.py:150-179, dt floor 1e-15)."""
        N = len(omega_grid)
        dt = 2 * np.pi / (N * np.abs(omega_grid[1] - omega_grid[0])) if N > 1 else 1e-15
        dt = max(dt, 1e-15)
        t = np.arange(N) * dt
        t1, t2 = max(self.tau1, 1e-20), max(self.tau2, 1e-20)
        h = (t1 ** 2 + t2 ** 2) / (t1 * t2 ** 2) * np.exp(-t / t2) * np.sin(t / t1)
        h[0] = 0.0
        H = fft(h) * dt
        self._cached_raman_response = H / H[0] if np.abs(H[0]) > 1e-15 else np.ones_like(H, dtype=complex)

    def get_beta2(self):
        w0, dw = self.omega0, self.omega0 * 1e-4
        n = self._compute_refractive_index
        k = [n(np.array([2 * np.pi * self.c / w]))[0] * w / self.c for w in (w0 - dw, w0, w0 + dw)]
        return (k[2] - 2 * k[1] + k[0]) / dw ** 2

    # -- device propagation ------------------------------------------------------------------
    def _tables(self, omega):
        if self._cached_D is None or self._cached_omega_grid is None or \
                len(omega) != len(self._cached_omega_grid) or \
                not np.allclose(omega, self._cached_omega_grid):
            self._cached_omega_grid = omega.copy()
            self._cached_D = self._compute_dispersion_operator(omega)
            if self.use_raman:
                self._update_raman_response(omega)
        if self.use_raman and self._cached_raman_response is None:
            self._update_raman_response(omega)
        D = np.asarray(self._cached_D, np.complex128)
        if self.use_self_steepening:
            if self._cached_n_omega is None:
                self._compute_dispersion_operator(omega)
            n_w = np.asarray(self._cached_n_omega, np.float64)
            uppe = ((self.omega0 + omega) / self.omega0) * (self.n0 / np.where(n_w < 1.0, 1.0, n_w))
        else:
            uppe = np.ones_like(omega)
        H = np.asarray(self._cached_raman_response, np.complex128) if self.use_raman else None
        return D, uppe, H

    def _desc(self, n, dt, L, tol, max_step, min_step, max_hist, max_attempts=0):
        d = _lib.Uppe1DDesc(nt=n, dt=dt, omega0=self.omega0, gamma=float(self.gamma),
                            use_ss=int(bool(self.use_self_steepening)),
                            use_raman=int(bool(self.use_raman)), f_R=float(self.f_R),
                            L=float(L), tol=float(tol), max_step=float(max_step),
                            min_step=float(min_step), max_hist=int(max_hist),
                            max_attempts=int(max_attempts))
        pl = self.plasma
        if pl is not None:
            area = pl.area if pl.area is not None else getattr(self, 'beam_area', None)
            if not area:
                raise ValueError("plasma model needs an area (Plasma1D.area or n2+beam_radius)")
            d.use_plasma, d.K, d.sigma_K, d.Ui = 1, int(pl.K), float(pl.sigma_K), float(pl.Ui)
            d.rho_nt, d.tau_rec, d.tau_c = float(pl.rho_nt), float(pl.tau_rec), float(pl.tau_c)
            d.sigma_drude, d.area = float(pl.drude(self.omega0, self.c)), float(area)
            d.avalanche = int(bool(pl.avalanche))
        return d

    def propagate(self, A0, t, L, tol=1e-4, max_step=1e-2, min_step=1e-11):
        """Adaptive ERK4(3)-IP propagation (erk43ip_method.py:288-369) on the GPU.

        Returns (z_history, A_history, omega) as the reference: NumPy arrays, A_history
        complex128 (Nz, Nt) in the time domain, history rule "append while fewer than 1000
        entries or when z advanced by more than L/500".
        """
        import torch
        lib = _lib.load()
        t = np.asarray(t, np.float64)
        N = len(t)
        dt = t[1] - t[0]
        omega = 2 * np.pi * fftfreq(N, dt)
        D, uppe, H = self._tables(omega)
        dev = torch.device("cuda")
        mh = self.MAX_HIST
        d = self._desc(N, dt, L, tol, max_step, min_step, mh)
        nbytes = ctypes.c_size_t()
        _lib.check(lib.uppe1d_workspace_size(ctypes.byref(d), ctypes.byref(nbytes)),
                   "uppe1d_workspace_size")
        ws = torch.empty(nbytes.value, dtype=torch.uint8, device=dev)
        a0 = torch.from_numpy(np.ascontiguousarray(A0, np.complex64)).to(dev)
        hist = torch.empty((mh, N), dtype=torch.complex64, device=dev)
        zh = torch.empty(mh, dtype=torch.float64, device=dev)
        Dre, Dim = np.ascontiguousarray(D.real), np.ascontiguousarray(D.imag)
        uf = np.ascontiguousarray(uppe, np.float64)
        Hre = np.ascontiguousarray(H.real) if H is not None else None
        Him = np.ascontiguousarray(H.imag) if H is not None else None
        st = _lib.Uppe1DStatus()
        rc = lib.uppe1d_propagate(
            ctypes.byref(d), _lib.host_ptr(Dre), _lib.host_ptr(Dim), _lib.host_ptr(uf),
            _lib.host_ptr(Hre) if Hre is not None else None,
            _lib.host_ptr(Him) if Him is not None else None,
            _lib.ptr(a0), _lib.ptr(ws), _lib.ptr(hist), _lib.ptr(zh), None, None,
            _lib.stream_handle(), ctypes.byref(st))
        if rc == -3:
            raise FloatingPointError(lib.uppe_last_error().decode())
        _lib.check(rc, "uppe1d_propagate")
        self.last_status = st
        nh = st.nhist
        z_hist = zh[:nh].cpu().numpy()
        A_hist = hist[:nh].cpu().numpy().astype(np.complex128)
        return z_hist, A_hist, omega

    def propagate_slices(self, A0, t, lengths, gammas=None, tol=1e-4, max_step=1e-2,
                         min_step=1e-11, record="all", max_hist=None):
        """A chain of propagations in one kernel launch (the slicing loop of gas.py:113-147 and
        "gas copy.py":115-141, with no host round trip between slices).

        Slice i propagates over lengths[i] with nonlinear coefficient gammas[i] (default
        self.gamma) and its own tol / max_step / min_step (scalars or one value per slice),
        starting like a fresh ``propagate`` call from the field the reference loop hands on:
        A_hist[-1] of slice i-1.  record="all": returns (z_hists, A_hists, omega), one
        (z_hist, A_hist) per slice as ``propagate`` returns them; record="last": returns
        (A_out, omega), A_out (n, Nt) complex128 the field each slice hands on.
        max_hist: history snapshots kept in total for record="all" (default: up to
        1502 per slice within 1 GiB); running out raises.
        """
        import torch
        lib = _lib.load()
        if record not in ("all", "last"):
            raise ValueError("record must be 'all' or 'last'")
        t = np.asarray(t, np.float64)
        N = len(t)
        dt = t[1] - t[0]
        omega = 2 * np.pi * fftfreq(N, dt)
        lengths = np.atleast_1d(np.asarray(lengths, np.float64))
        n = len(lengths)
        if n == 0:
            raise ValueError("no slices")
        per = lambda v: np.broadcast_to(np.asarray(v, np.float64), (n,))
        g = per(self.gamma if gammas is None else gammas)
        tl, mx, mn = per(tol), per(max_step), per(min_step)
        slices = (_lib.Uppe1DSlice * n)()
        for i in range(n):
            slices[i] = _lib.Uppe1DSlice(L=lengths[i], gamma=g[i], tol=tl[i], max_step=mx[i],
                                         min_step=mn[i])
        D, uppe, H = self._tables(omega)
        dev = torch.device("cuda")
        if record == "all":
            if max_hist is None:
                max_hist = int(min(n * self.MAX_HIST, max(64 * n, (1 << 30) // (8 * N))))
        else:
            max_hist = 0
        d = self._desc(N, dt, 0.0, 1.0, 1.0, 0.0, max_hist)
        nbytes = ctypes.c_size_t()
        _lib.check(lib.uppe1d_slices_workspace_size(ctypes.byref(d), n, ctypes.byref(nbytes)),
                   "uppe1d_slices_workspace_size")
        ws = torch.empty(nbytes.value, dtype=torch.uint8, device=dev)
        a0 = torch.from_numpy(np.ascontiguousarray(A0, np.complex64)).to(dev)
        hist = torch.empty((max(max_hist, 1), N), dtype=torch.complex64, device=dev)
        zh = torch.empty(max(max_hist, 1), dtype=torch.float64, device=dev)
        out_t = torch.empty((n, N), dtype=torch.complex64, device=dev) if record == "last" \
            else None
        Dre, Dim = np.ascontiguousarray(D.real), np.ascontiguousarray(D.imag)
        uf = np.ascontiguousarray(uppe, np.float64)
        Hre = np.ascontiguousarray(H.real) if H is not None else None
        Him = np.ascontiguousarray(H.imag) if H is not None else None
        st = (_lib.Uppe1DStatus * n)()
        rc = lib.uppe1d_propagate_slices(
            ctypes.byref(d), n, slices, _lib.host_ptr(Dre), _lib.host_ptr(Dim),
            _lib.host_ptr(uf), _lib.host_ptr(Hre) if Hre is not None else None,
            _lib.host_ptr(Him) if Him is not None else None, _lib.ptr(a0), _lib.ptr(ws),
            _lib.ptr(hist) if max_hist > 0 else None, _lib.ptr(zh) if max_hist > 0 else None,
            _lib.ptr(out_t) if out_t is not None else None, None, _lib.stream_handle(), st)
        if rc == -3:
            raise FloatingPointError(lib.uppe_last_error().decode())
        _lib.check(rc, "uppe1d_propagate_slices")
        self.last_slice_status = list(st)
        self.last_status = st[n - 1]
        if record == "last":
            return out_t.cpu().numpy().astype(np.complex128), omega
        if any(x.overflow & 1 for x in st):
            raise RuntimeError("propagate_slices: history buffer of %d snapshots ran out; pass a "
                               "larger max_hist" % max_hist)
        counts = [x.nhist for x in st]
        tot = sum(counts)
        zall = zh[:tot].cpu().numpy()
        Aall = hist[:tot].cpu().numpy().astype(np.complex128)
        edges = np.concatenate([[0], np.cumsum(counts)])
        z_hists = [zall[edges[i]:edges[i + 1]] for i in range(n)]
        A_hists = [Aall[edges[i]:edges[i + 1]] for i in range(n)]
        return z_hists, A_hists, omega

    # -- input generators (host; erk43ip_method.py:375-446) ----------------------------------
    def generate_gaussian_pulse(self, t, pulse_energy, pulse_fwhm, chirp=0):
        P0 = pulse_energy / (pulse_fwhm * np.sqrt(np.pi / (4 * np.log(2))))
        T0 = pulse_fwhm / (2 * np.sqrt(np.log(2)))
        return np.sqrt(P0) * np.exp(-(1 + 1j * chirp) * (t / T0) ** 2 / 2)

    def generate_dispersed_pulse(self, t, pulse_energy, pulse_fwhm, GDD=0, TOD=0):
        P0 = pulse_energy / (pulse_fwhm * np.sqrt(np.pi / (4 * np.log(2))))
        T0 = pulse_fwhm / (2 * np.sqrt(np.log(2)))
        A_tl = np.sqrt(P0) * np.exp(-(t / T0) ** 2 / 2)
        if GDD == 0 and TOD == 0:
            return A_tl
        omega = 2 * np.pi * fftfreq(len(t), t[1] - t[0])
        phase = 0.5 * GDD * omega ** 2 + (1.0 / 6.0) * TOD * omega ** 3
        return fftshift(ifft(fft(ifftshift(A_tl)) * np.exp(1j * phase)))


def _attach_spectrum_methods():
    from . import spectrum_io

    def generate_pulse_from_spectrum(self, t, wavelengths_nm, intensities, pulse_energy, GDD=0,
                                     TOD=0, use_jacobian=True, interp_kind='cubic',
                                     exact_fourier=False, time_window_auto=True,
                                     N_time_points=4096):
        """Pulse from a measured spectrum (erk43ip_method.py:599-764); host FP64, see
        project_amd.spectrum_io.pulse_from_spectrum."""
        return spectrum_io.pulse_from_spectrum(
            t, wavelengths_nm, intensities, pulse_energy, self.lambda0, GDD=GDD, TOD=TOD,
            use_jacobian=use_jacobian, interp_kind=interp_kind, exact_fourier=exact_fourier,
            time_window_auto=time_window_auto, N_time_points=N_time_points, c=self.c)

    def _calculate_fwhm_precise(self, x, y):
        return spectrum_io.fwhm_precise(x, y)

    ERK43IP_UPPE.generate_pulse_from_spectrum = generate_pulse_from_spectrum
    ERK43IP_UPPE._calculate_fwhm_precise = _calculate_fwhm_precise


_attach_spectrum_methods()


class SimResultAdapter:
    """Result adapter of erk43ip_method.py:827-860 (plotting helpers consume it)."""

    def __init__(self, solver, z_array, t, A_evolution):
        self.material = solver.material
        self.lambda0 = solver.lambda0
        self.omega0 = solver.omega0
        self.c = solver.c
        br = getattr(solver, 'beam_radius', None)
        self.beam_radius = br if br is not None else 0.0
        n2 = getattr(solver, 'n2', None)
        self.n2 = n2 if n2 is not None else 0.0
        self.gamma = solver.gamma
        self.t = t
        self.Nt = len(t)
        self.dt = t[1] - t[0]
        self.T = self.Nt * self.dt
        self.z = z_array
        self.L = z_array[-1]
        self.Nz = len(z_array)
        self.dz = z_array[1] - z_array[0] if len(z_array) > 1 else 0
        try:
            self.beta2 = solver.get_beta2()
        except Exception:
            self.beta2 = 0.0
        self.A_evolution = np.array(A_evolution)


ERK43IP_FullDispersion = ERK43IP_UPPE
