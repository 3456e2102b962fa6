"""Spectrum input and output formats of the 1D drivers (host, FP64).

* ``load_spectrum_csv``: the measured-spectrum CSVs the drivers read
  (``src/1D-ERK4(3)IP UPPE/test_spectrum_pulse-out.py:23-52``): rows of at least two numbers,
  wavelength (nm) then intensity; empty cells are ignored, a cell is read after its last ``]``,
  and a row with a cell that is not a number (a header, or a UTF-8 byte-order mark, which the
  reference's ``utf-8`` decoding keeps) is skipped.
* ``write_spectrum_csv``: the output spectrum file
  (``test_spectrum_pulse-coupling_efficiency-out.py:336-383``): header
  ``Wavelength (nm),Normalized Intensity``, rows ``%.6f,%.8e`` sorted by wavelength, intensity
  |FFT(A)|^2 * f_abs^2 normalised to its maximum (the Jacobian that puts a frequency spectrum
  on a wavelength axis), positive absolute frequencies only, wavelengths in [wl_min, wl_max].
* ``pulse_from_spectrum``: ``ERK43IP_UPPE.generate_pulse_from_spectrum``
  (``erk43ip_method.py:448-497, 499-597, 599-764``), FFT and exact-Fourier variants.
"""
import csv

import numpy as np
from scipy.fft import fftfreq, fftshift, ifft
from scipy.integrate import simpson
from scipy.interpolate import interp1d

C0 = 299792458.0


def load_spectrum_csv(filename):
    wl, inten = [], []
    with open(filename, "r", encoding="utf-8") as f:
        for row in csv.reader(f):
            vals = []
            try:
                for cell in row:
                    s = cell.strip()
                    if "]" in s:
                        s = s.rsplit("]", 1)[1].strip()
                    if s:
                        vals.append(float(s))
            except ValueError:
                continue
            if len(vals) >= 2:
                wl.append(vals[0])
                inten.append(vals[1])
    return np.array(wl), np.array(inten)


def spectrum_rows(A_out, t, center_wavelength, wl_min=1060.0, wl_max=1070.0, c=C0):
    """(wavelength nm, normalised intensity) rows of write_spectrum_csv, sorted, in range."""
    n = len(t)
    f = fftshift(fftfreq(n, t[1] - t[0]))
    S = np.abs(fftshift(np.fft.fft(np.asarray(A_out)))) ** 2
    f_abs = f + c / center_wavelength
    keep = f_abs > 0
    f_abs, S = f_abs[keep], S[keep]
    wl = c / f_abs * 1e9
    I = S * f_abs ** 2
    peak = np.max(I)
    if peak > 0:
        I = I / peak
    order = np.argsort(wl)
    wl, I = wl[order], I[order]
    sel = (wl >= wl_min) & (wl <= wl_max)
    return wl[sel], I[sel]


def write_spectrum_csv(filename, A_out, t, center_wavelength, wl_min=1060.0, wl_max=1070.0,
                       c=C0):
    wl, I = spectrum_rows(A_out, t, center_wavelength, wl_min, wl_max, c)
    with open(filename, "w", newline="", encoding="utf-8") as f:
        w = csv.writer(f)
        w.writerow(["Wavelength (nm)", "Normalized Intensity"])
        for a, b in zip(wl, I):
            w.writerow(["%.6f" % a, "%.8e" % b])
    return len(wl)


def fwhm_precise(x, y):
    """Half-maximum crossings by linear interpolation (erk43ip_method.py:448-497): the first
    and last sign changes of y - max(y)/2; with fewer than two crossings, the first and last
    samples at or above half maximum; (None, None, None) when fewer than two such samples."""
    half = np.max(y) / 2
    above = (y >= half).astype(int)
    cross = np.nonzero(np.diff(above))[0]
    if len(cross) < 2:
        idx = np.nonzero(y >= half)[0]
        if len(idx) < 2:
            return None, None, None
        return x[idx[0]], x[idx[-1]], x[idx[-1]] - x[idx[0]]

    def at(i):
        return x[i] + (half - y[i]) * (x[i + 1] - x[i]) / (y[i + 1] - y[i])

    xl, xr = at(cross[0]), at(cross[-1])
    return xl, xr, xr - xl


def _spectrum_in_frequency(wavelengths_nm, intensities, use_jacobian, c):
    # sort by wavelength, normalise, to frequency (Jacobian lambda^2 / c), drop duplicate
    # frequencies (first occurrence in ascending-frequency order), order by descending frequency
    order = np.argsort(wavelengths_nm)
    wl = np.asarray(wavelengths_nm, np.float64)[order]
    I = np.asarray(intensities, np.float64)[order]
    if np.max(I) > 0:
        I = I / np.max(I)
    nu = c / (wl * 1e-9)
    if use_jacobian:
        I = I * (wl * 1e-9) ** 2 / c
        if np.max(I) > 0:
            I = I / np.max(I)
    u, first = np.unique(nu, return_index=True)
    if len(u) < len(nu):
        nu, I = u, I[first]
    if nu[0] < nu[-1]:
        nu, I = nu[::-1], I[::-1]
    return nu, I


def _interpolator(nu, I, kind):
    if kind in ("cubic", "quadratic") and len(nu) < 4:
        kind = "linear"
    return interp1d(nu, I, kind=kind, bounds_error=False, fill_value=0.0)


def pulse_from_spectrum(t, wavelengths_nm, intensities, pulse_energy, lambda0, GDD=0.0, TOD=0.0,
                        use_jacobian=True, interp_kind="cubic", exact_fourier=False,
                        time_window_auto=True, N_time_points=4096, c=C0):
    """Field A(t) whose spectrum is the measured one, with spectral phase
    GDD w^2 / 2 + TOD w^3 / 6, scaled to pulse_energy (trapezoid rule on |A|^2).

    FFT variant: the spectral density interpolated onto the simulation's absolute frequency
    grid c / lambda0 + fftfreq(N, dt), A = fftshift(ifft(sqrt(S) e^{i phase})).
    Exact variant: S on a dense uniform grid of min(4096, 4 M) frequencies, centred on the
    middle of its half-maximum band (or its centroid), and A(t) = Simpson integral of
    sqrt(S) e^{i phase} e^{-i w t} over the centred frequency, then fftshift.  With t None the
    exact variant builds its own window: 1 / (input frequency spacing), N_time_points."""
    wavelengths_nm = np.asarray(wavelengths_nm, np.float64)
    if t is None and exact_fourier and time_window_auto:
        nu_in = c / (wavelengths_nm * 1e-9)
        span = nu_in.max() - nu_in.min()
        if len(nu_in) > 1:
            tmax = 1 / np.abs(nu_in[1] - nu_in[0])
        else:
            tmax = 1 / span if span > 0 else 1e-12
        t = np.linspace(-tmax / 2, tmax / 2, N_time_points)
    t = np.asarray(t, np.float64)
    N = len(t)
    dt = t[1] - t[0]
    nu, I = _spectrum_in_frequency(wavelengths_nm, intensities, use_jacobian, c)
    f = _interpolator(nu, I, interp_kind)
    if exact_fourier:
        m = min(4096, len(nu) * 4)
        nu_d = np.linspace(nu.min(), nu.max(), m)
        S = f(nu_d)
        S[S < 0] = 0
        fl, fr, dnu = fwhm_precise(nu_d, S)
        if dnu:
            nu_c = (fl + fr) / 2
        else:
            nu_c = np.sum(nu_d * S) / np.sum(S) if np.sum(S) > 0 else c / lambda0
        df = nu_d - nu_c
        w = 2 * np.pi * df
        amp = np.sqrt(S) * np.exp(1j * (0.5 * GDD * w ** 2 + TOD * w ** 3 / 6.0))
        A = np.empty(N, complex)
        for i in range(N):   # one Simpson integral per time sample
            A[i] = simpson(amp * np.exp(-1j * w * t[i]), x=df)
        A = fftshift(A)
    else:
        w_rel = 2 * np.pi * fftfreq(N, dt)
        S = f(w_rel / (2 * np.pi) + c / lambda0)
        S[S < 0] = 0
        S[np.isnan(S)] = 0
        phase = 0.5 * GDD * w_rel ** 2 + TOD * w_rel ** 3 / 6.0
        A = fftshift(ifft(np.sqrt(S) * np.exp(1j * phase)))
    E = np.trapezoid(np.abs(A) ** 2, t) if hasattr(np, "trapezoid") else np.trapz(np.abs(A) ** 2, t)
    if E > 0:
        A = A * np.sqrt(pulse_energy / E)
    return A
