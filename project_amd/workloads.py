"""Synthetic inputs of the benchmark and parity workloads (no datasets: everything is
generated from seeds).  The generators follow the reference's conventions:

* grids ``t = linspace(-T/2, T/2, N)`` (dt = T/(N-1)), as in the reference harnesses;
* 1D pulses of ``erk43ip_method.py:375-446`` (P0 = E / (FWHM sqrt(pi/(4 ln 2))));
* 3D beams normalised to a pulse energy sum |A|^2 dx dy dt
  (``src/3+1D-ERK4(3)IP UPPE/sim.py:98-118``).
"""
import numpy as np


# -- 3+1D: Gaussian beam w0 = 1 mm, 40 fs intensity FWHM at 800 nm, 0.8 mJ in argon -------
def grid3d(nx=128, nt=1024, half_width=4e-3, window=400e-15):
    x = np.linspace(-half_width, half_width, nx)
    t = np.linspace(-window / 2, window / 2, nt)
    return x, x.copy(), t


def beam3d(x, y, t, energy=0.8e-3, w0=1e-3, tau_fwhm=40e-15, noise=1e-4, seed=1):
    """complex128 (nx, ny, nt): Gaussian in space and time times a seeded complex spatial
    noise factor (1 + noise (n_r + i n_i)/sqrt 2), scaled to `energy`."""
    X, Y = np.meshgrid(x, y, indexing="ij")
    sp = np.exp(-(X ** 2 + Y ** 2) / w0 ** 2)
    rng = np.random.default_rng(seed)
    nz = (rng.standard_normal(sp.shape) + 1j * rng.standard_normal(sp.shape)) / np.sqrt(2)
    sp = sp * (1.0 + noise * nz)
    tm = np.exp(-2.0 * np.log(2.0) * t ** 2 / tau_fwhm ** 2)
    A = sp[:, :, None] * tm[None, None, :]
    dx, dy, dt = x[1] - x[0], y[1] - y[0], t[1] - t[0]
    A *= np.sqrt(energy / (np.sum(np.abs(A) ** 2) * dx * dy * dt))
    return A


# -- 1D: 800 nm, 30 fs FWHM, 1 uJ Gaussian on a 400 fs window with seeded white noise ------
def pulse1d(nt=4096, window=400e-15, energy=1e-6, fwhm=30e-15, noise=1e-6, seed=0, GDD=0.0):
    t = np.linspace(-window / 2, window / 2, nt)
    p0 = energy / (fwhm * np.sqrt(np.pi / (4 * np.log(2))))
    t0 = fwhm / (2 * np.sqrt(np.log(2)))
    a = np.sqrt(p0) * np.exp(-(t / t0) ** 2 / 2)
    if GDD != 0.0:
        w = 2 * np.pi * np.fft.fftfreq(nt, t[1] - t[0])
        a = np.fft.fftshift(np.fft.ifft(np.fft.fft(np.fft.ifftshift(a)) * np.exp(0.5j * GDD * w ** 2)))
    if noise:
        rng = np.random.default_rng(seed)
        a = a + noise * np.abs(a).max() * (rng.standard_normal(nt)
                                           + 1j * rng.standard_normal(nt)) / np.sqrt(2)
    return t, a.astype(np.complex128)
