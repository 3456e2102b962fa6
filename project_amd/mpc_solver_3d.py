"""The MPC cavity runner's 3+1D ERK4(3)-IP solver on MI355X (drop-in for
``src/3+1D-ERK4(3)IP UPPE/solver_3d.py:6-343``).

That solver is the envelope solver's algorithm with a built-in medium instead of a plugin:
a Sellmeier index for fused silica and sapphire, the constant 1.00027 for air
(``solver_3d.py:96-129``), a Kerr-only nonlinearity i (omega0 / c) n2 |A|^2 A with an optional
shock factor (``:41,171-173,202-219``), no negative-frequency mask, and the relative L2 error
norm sqrt(sum|A4 - A3|^2) / (sqrt(sum|A4|^2) + 1e-20) (``:252-254``).  The step, absorber,
anti-alias filter, A_scale normalisation and controller are the envelope solver's, so this
class runs on the same HIP plan (``project_amd.solver_3d``) with the L2 norm switched on
(``uppe3d_set_error_norm``) and a medium whose ionisation table is zero (the plasma code is
never instantiated).

Constructor kept: ``(x_grid, y_grid, t_grid, material='fused_silica', n2=3e-20,
center_wavelength=800e-9, use_raman=True, use_shock=True, f_R=0.18, tau1=12.2e-15,
tau2=32e-15, precision='single')``.  Differences: the Raman response (``:187-200,205-210``) is
not implemented and ``use_raman=True`` raises (the runner passes ``use_raman=False``,
``sim.py:61-68``); the t axis takes 128 to 2048 points with prime factors of at most 31 (the
runner's own N_t = 2046 included).
"""
import numpy as np

from . import solver_3d as _env

C_LIGHT = 299792458.0

# solver_3d.py:96-106: (B terms, C terms [um^2], valid wavelength range [m]); unknown names
# fall back to fused silica
MATERIALS = {
    "fused_silica": ((0.6961663, 0.4079426, 0.8974794),
                     (0.0684043 ** 2, 0.1162414 ** 2, 9.896161 ** 2), (0.2e-6, 6.0e-6)),
    "sapphire": ((1.4313493, 0.65054713, 5.3414021),
                 (0.0726631 ** 2, 0.1193242 ** 2, 18.028251 ** 2), (0.2e-6, 5.5e-6)),
    "air": (None, None, None),
}


class BuiltinMaterial:
    """The solver's built-in medium, in the parameter form the HIP plan reads: Kerr (n2), no
    HOKE, an all-zero ionisation table and no avalanche, so the plasma term is identically 0."""

    def __init__(self, material="fused_silica", n2=3.0e-20, center_wavelength=800e-9,
                 use_shock=True):
        self.material = material
        self.sellmeier_coeffs = MATERIALS.get(material, MATERIALS["fused_silica"])
        self.c = C_LIGHT
        self.lambda0 = center_wavelength
        self.omega0 = 2 * np.pi * C_LIGHT / center_wavelength
        self.n2 = float(n2)
        self.n4 = 0.0
        self.use_shock = bool(use_shock)
        self.use_raman = False
        self.Ui = 0.0
        self.rho_nt = 0.0
        self.tau_rec = 1.0
        self.tau_c = 0.0
        self.sigma_drude = 0.0
        self.log10_I_cpu = np.array([8.0, 20.0])
        self.W_cpu = np.zeros(2)

    def refractive_index(self, omega_array):
        """solver_3d.py:108-129 (FP64)."""
        w = np.asarray(omega_array, np.float64)
        if self.material == "air":
            return np.ones_like(w) * 1.00027
        B, Cc, rng = self.sellmeier_coeffs
        w = np.where(np.abs(w) < 1e-12, 1e-12, w)
        lam = 2 * np.pi * C_LIGHT / np.abs(w)
        if rng:
            lam = np.clip(lam, rng[0], rng[1])
        l2 = (lam * 1e6) ** 2
        nsq = np.ones_like(lam)
        for b, c in zip(B, Cc):
            nsq = nsq + b * l2 / (l2 - c)
        return np.sqrt(nsq)

    def get_shock_operator(self, omega_array):
        w = np.asarray(omega_array, np.float64)
        return (w + self.omega0) / self.omega0 if self.use_shock else np.ones_like(w)

    def avalanche_coefficient(self):
        return 0.0


class ERK43IP_UPPE_3D_Optimized(_env.ERK43IP_UPPE_3D_Optimized):
    error_norm = "l2"
    valid_mask = False

    def __init__(self, x_grid, y_grid, t_grid, material="fused_silica", n2=3.0e-20,
                 center_wavelength=800e-9, use_raman=True, use_shock=True, f_R=0.18,
                 tau1=12.2e-15, tau2=32e-15, precision="single"):
        if use_raman:
            raise NotImplementedError(
                "the Raman response is not implemented in the HIP kernels; pass "
                "use_raman=False (as the MPC runner does)")
        medium = BuiltinMaterial(material, n2, center_wavelength, use_shock)
        super().__init__(x_grid, y_grid, t_grid, medium, precision)
        self.material = material
        self.sellmeier_coeffs = medium.sellmeier_coeffs
        self.n2 = float(n2)
        self.use_raman = False
        self.use_shock = bool(use_shock)
        self.f_R, self.tau1, self.tau2 = f_R, tau1, tau2
        self.gamma_base = 1j * (self.omega0 / self.c) * self.n2

    def _compute_refractive_index(self, omega_array):
        return self.medium.refractive_index(omega_array)
