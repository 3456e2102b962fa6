"""Gas / solid media for the 3+1D envelope solver: parameter carriers for the fused kernels.

Drop-in for ``src/3+1D-UPPE-ERK4(3)IP-envelope/media_physics.py`` (reference paths are
relative to the reference root).  The classes keep the reference names, constructor
signatures and attributes (``c, lambda0, omega0, n0, n2, n4, Ui, rho_nt, tau_rec, tau_c,
sigma_drude, sigma_K, K``, the ionisation LUT ``log10_I_cpu`` / ``W_cpu``) and the methods
``refractive_index``, ``get_shock_operator``, ``calculate_keldysh_rate_cpu``,
``get_avalanche_rate``.  Host-side arithmetic is FP64 NumPy; it only builds tables once per
solver.  The nonlinear response itself is computed by the HIP t-pass kernel from these
parameters (Kerr + HOKE, LUT ionisation rate cast to float32, FP64 plasma recurrence, loss and
defocus terms) - a subclass that overrides ``calc_nonlinear_response`` is rejected by the
solver with a clear error rather than silently ignored.

Differences from the reference, both opt-in and off by default:

* ``tunnel_guard`` (default 1e-20, the reference's value in ``media_physics.py:141-143``):
  the reference adds 1e-20 to e*hbar*E (~1e-44..1e-41), so its tunnelling term is
  1e15 * exp(-~1e-21) = 1e15 s^-1 at every intensity.  With that LUT the plasma loss term
  -gen*Ui/(2I + 1e-30)*A is enormous where |A| is small and the first ERK attempt of a
  Gaussian input produces a non-finite error norm (the reference's loop would then never
  terminate).  ``tunnel_guard=0.0`` evaluates the intended Keldysh exponent.
* ``precision`` is accepted for signature compatibility; the device path is always complex64
  with FP64 plasma recurrence and error norm.
"""
import numpy as np

C_LIGHT = 299792458.0


class OpticalMedium:
    """Base medium (media_physics.py:70-205): Keldysh MPI + tunnel LUT, plasma constants."""

    def __init__(self, center_wavelength, precision="single", use_shock=False,
                 tunnel_guard=1e-20):
        self.lambda0 = center_wavelength
        self.c = C_LIGHT
        self.omega0 = 2 * np.pi * self.c / self.lambda0
        self.use_shock = use_shock
        self.precision = precision
        self.tunnel_guard = tunnel_guard
        self.e_charge = 1.60217663e-19
        self.eps0 = 8.85418781e-12
        self.m_e = 9.1093837e-31
        self.hbar = 1.054571817e-34
        self.n0 = 1.0
        self.n2 = 0.0
        self.n4 = 0.0
        self.Ui = 0.0
        self.rho_nt = 0.0
        self.tau_c = 0.0
        self.tau_rec = 1e-9
        self.sigma_drude = 0.0
        self.sigma_K = 0.0
        self.K = 1

    def _drude(self, n0):
        return (self.e_charge ** 2 * self.tau_c) / (
            self.c * self.eps0 * n0 * self.m_e * (1.0 + self.omega0 ** 2 * self.tau_c ** 2))

    def refractive_index(self, omega_array):
        raise NotImplementedError

    def get_shock_operator(self, omega_array):
        """(omega + omega0) / omega0 if use_shock else 1, shape (1, 1, Nt)."""
        w = np.asarray(omega_array, np.float64)
        f = (w + self.omega0) / self.omega0 if self.use_shock else np.ones_like(w)
        return f[None, None, :]

    def _init_ionization_lut(self, I_min_W_cm2=1e8, I_max_W_cm2=1e16, points=2000):
        self.log10_I_cpu = np.linspace(np.log10(I_min_W_cm2 * 1e4),
                                       np.log10(I_max_W_cm2 * 1e4), points)
        self.I_grid_cpu = 10 ** self.log10_I_cpu
        self.W_cpu = self.calculate_keldysh_rate_cpu(self.I_grid_cpu)

    def calculate_keldysh_rate_cpu(self, I_grid):
        """sigma_K I^K + 1e15 exp(-(4/3) sqrt(2 m_e) Ui^1.5 / (e hbar E + guard))."""
        I_grid = np.asarray(I_grid, np.float64)
        w_mpi = self.sigma_K * I_grid ** self.K
        E = np.sqrt(2 * I_grid / (self.c * self.eps0 * self.n0))
        with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
            ex = -(4.0 / 3.0) * np.sqrt(2 * self.m_e) * self.Ui ** 1.5 / (
                self.e_charge * self.hbar * E + self.tunnel_guard)
            return w_mpi + 1e15 * np.exp(ex)

    def get_ionization_rate(self, Intensity):
        """Host evaluation of the LUT (media_physics.py:131-137), for inspection only."""
        lg = np.clip(np.log10(np.asarray(Intensity, np.float64) + 1e-30),
                     self.log10_I_cpu[0], self.log10_I_cpu[-1])
        return np.interp(lg, self.log10_I_cpu, self.W_cpu)

    def get_effective_intensity(self, Intensity):
        return Intensity

    def get_avalanche_rate(self, Intensity):
        """Base avalanche rate (sigma_drude / Ui) * I (media_physics.py:152-154)."""
        return (self.sigma_drude / self.Ui) * np.asarray(Intensity)

    # -- what the fused kernel consumes ------------------------------------------------
    def avalanche_coefficient(self):
        """av_coef with W_ava = av_coef * I: 0 for media that override the rate to zero."""
        probe = np.asarray(self.get_avalanche_rate(np.array([1.0, 2.0])), np.float64)
        if not np.any(probe):
            return 0.0
        coef = float(probe[0])
        if not np.allclose(probe, coef * np.array([1.0, 2.0]), rtol=1e-12, atol=0):
            raise NotImplementedError("avalanche rate must be linear in I for the fused kernel")
        return coef


class FusedSilica(OpticalMedium):
    """Fused silica (media_physics.py:208-250): Sellmeier index clamped to 0.2-6 um, n0 = 1.45
    in the Keldysh field and the Drude cross section, K = 8 multiphoton ionisation and the base
    class's avalanche rate (sigma_drude / Ui) I, which the fused kernel evaluates with the
    general (non-affine) FP64 density recurrence.  ``use_raman=True`` is accepted here and
    rejected by the solver (the Raman response, ``:252-278``, is not in the HIP kernels)."""

    def __init__(self, center_wavelength, precision="single", use_shock=False, use_raman=False,
                 tunnel_guard=1e-20):
        super().__init__(center_wavelength, precision, use_shock, tunnel_guard)
        self.n0 = 1.45
        self.n2 = 3.0e-20
        self.use_raman = use_raman
        self.B = [0.6961663, 0.4079426, 0.8974794]
        self.C = [0.0684043 ** 2, 0.1162414 ** 2, 9.896161 ** 2]
        self.valid_range = (0.2e-6, 6.0e-6)
        self.f_R = 0.18
        self.tau1 = 12.2e-15
        self.tau2 = 32.0e-15
        self.H_raman = None
        self.Ui = 9.0 * self.e_charge
        self.rho_nt = 2.1e28
        self.K = 8
        self.tau_rec = 150e-15
        self.tau_c = 3.0e-15
        self.sigma_drude = self._drude(self.n0)
        self.sigma_K = 2.5e-125
        self._init_ionization_lut()

    def refractive_index(self, omega_array):
        w = np.asarray(omega_array, np.float64)
        w = np.where(np.abs(w) < 1e-12, 1e-12, w)
        lam = np.clip(2 * np.pi * self.c / np.abs(w), self.valid_range[0], self.valid_range[1])
        l2 = (lam * 1e6) ** 2
        n_sq = np.ones_like(l2)
        for b, c in zip(self.B, self.C):
            n_sq = n_sq + b * l2 / (l2 - c)
        return np.sqrt(n_sq)


class _CauchyGas(OpticalMedium):
    """Noble gas with Cauchy dispersion (media_physics.py:317-510)."""
    _params = None   # (A, B, C, n2/atm, n4/atm^2, Ui eV, K, sigma_K/atm, tau_c*atm)

    def __init__(self, center_wavelength, pressure_atm=1.0, precision="single",
                 use_shock=False, tunnel_guard=1e-20):
        super().__init__(center_wavelength, precision, use_shock, tunnel_guard)
        A, B, Cc, n2, n4, ui, K, sk, tc = self._params
        P = pressure_atm
        self.pressure = P
        self.cauchy_A, self.cauchy_B, self.cauchy_C = A, B, Cc
        self.n2 = n2 * P
        self.n4 = n4 * (P ** 2)
        self.Ui = ui * self.e_charge
        self.rho_nt = 2.45e25 * P
        self.K = K
        self.tau_rec = 1.0e-9
        self.tau_c = tc / P
        self.sigma_drude = self._drude(1.0)
        self.sigma_K = sk * P
        self._init_ionization_lut()

    def refractive_index(self, omega_array):
        w = np.asarray(omega_array, np.float64)
        w = np.where(np.abs(w) < 1e-12, 1e-12, w)
        l2 = (2 * np.pi * self.c / np.abs(w) * 1e6) ** 2
        return 1.0 + (self.cauchy_A + self.cauchy_B / l2 + self.cauchy_C / l2 ** 2) * self.pressure

    def get_avalanche_rate(self, Intensity):
        return np.zeros_like(np.asarray(Intensity, np.float64))


class Helium(_CauchyGas):
    _params = (3.48e-5, 5.4e-8, 1.2e-10, 4.0e-25, -1.0e-38, 24.59, 17, 0.0, 500e-15)


class Neon(_CauchyGas):
    _params = (6.66e-5, 2.4e-8, 1.7e-10, 2.0e-24, -5.0e-38, 21.56, 15, 0.0, 350e-15)


class Argon(_CauchyGas):
    _params = (27.9e-5, 12.8e-8, 5.2e-10, 1.0e-23, -1.0e-36, 15.76, 11, 1.0e-188, 190e-15)


class Krypton(_CauchyGas):
    _params = (41.9e-5, 21.3e-8, 8.4e-10, 2.8e-23, -2.5e-36, 14.00, 10, 1.0e-168, 150e-15)


class Xenon(_CauchyGas):
    _params = (68.7e-5, 40.6e-8, 16.3e-10, 6.5e-23, -8.0e-36, 12.13, 8, 1.0e-130, 120e-15)


class Air(OpticalMedium):
    """Air (media_physics.py:282-313): constant index, O2 ionisation, no avalanche."""

    def __init__(self, center_wavelength, pressure_atm=1.0, precision="single",
                 use_shock=False, tunnel_guard=1e-20):
        super().__init__(center_wavelength, precision, use_shock, tunnel_guard)
        self.pressure = pressure_atm
        self.n0 = 1.0 + 0.00027 * pressure_atm
        self.n2 = 3.0e-23 * pressure_atm
        self.Ui = 12.1 * self.e_charge
        self.rho_nt = 2.7e25 * pressure_atm
        self.K = 11
        self.tau_rec = 1.0e-9
        self.tau_c = 350e-15 / pressure_atm
        self.sigma_drude = self._drude(self.n0)
        self.sigma_K = 2.0e-183 * pressure_atm
        self._init_ionization_lut()

    def refractive_index(self, omega_array):
        return np.ones_like(np.asarray(omega_array, np.float64)) * self.n0

    def get_avalanche_rate(self, Intensity):
        return np.zeros_like(np.asarray(Intensity, np.float64))
