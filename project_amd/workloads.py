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


# -- batched 1D sweep: N_t ~ U{1024..8192}, E ~ logU[0.1, 10] uJ, GDD ~ U[-500, 500] fs^2,
#    P ~ U[0.5, 3] bar, argon, 400 fs window, L = 1 m -------------------------------------
def batch1d_params(n=256, seed=2, nts=(1024, 2048, 4096, 8192)):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        out.append(dict(nt=int(rng.choice(nts)),
                        energy=float(10 ** rng.uniform(np.log10(0.1e-6), np.log10(10e-6))),
                        gdd=float(rng.uniform(-500e-30, 500e-30)),
                        pressure_bar=float(rng.uniform(0.5, 3.0))))
    return out


def argon_solver_1d(pressure_bar=1.0, beam_radius=20e-6, sigma8_per_atm=1e-130, plasma=True):
    """The 1D argon setup of the sweep: n2 and the plasma constants scale with pressure
    (n2 = 1e-23 P[bar], as the reference gas driver scales n2, gas.py:69-71)."""
    from . import gas  # noqa: F401  (argon refractive-index patch)
    from .erk43ip_method import ERK43IP_UPPE, Plasma1D
    p_atm = pressure_bar * 1e5 / 101325.0
    pl = Plasma1D(K=8, sigma_K=sigma8_per_atm * p_atm, rho_nt=2.45e25 * p_atm,
                  tau_c=190e-15 / p_atm) if plasma else None
    return ERK43IP_UPPE(material="argon_gas", n2=1e-23 * pressure_bar, beam_radius=beam_radius,
                        center_wavelength=800e-9, use_raman=False, use_self_steepening=True,
                        gas_P=pressure_bar * 1e5, gas_T=293.15, plasma=pl)


def batch1d_inputs(params, seed=2):
    inputs = []
    for k, p in enumerate(params):
        t, a = pulse1d(p["nt"], energy=p["energy"], GDD=p["gdd"], seed=seed * 1000 + k)
        inputs.append((t, a))
    return inputs
