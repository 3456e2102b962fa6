"""ORACLE (test infrastructure only): FP64 NumPy restatement of the older 3+1D solver and the
MPC cavity runner of ``src/3+1D-ERK4(3)IP UPPE/``.  Never imported by the product package;
only ``tests/`` use it, as the checker.

Follows:

* ``solver_3d.py:96-106``   material table (fused silica, sapphire Sellmeier terms; air)
* ``solver_3d.py:108-129``  refractive index: air 1.00027, else Sellmeier with |omega| guard
                            and the wavelength clipped to the valid range
* ``solver_3d.py:41,171-173``  Kerr coefficient i (omega0 / c) n2, shock (omega + omega0) / omega0
* ``solver_3d.py:202-219``  nonlinear operator (Raman off: the runner builds both solvers with
                            ``use_raman=False``)
* ``solver_3d.py:221-265``  the same ERK4(3)-IP step as the envelope solver, with the
                            relative L2 error norm (:252-254) and no negative-frequency mask
* ``solver_3d.py:267-343``  propagate (A_scale, controller with fail_count)
* ``sim.py:120-137``        propagate_segment (air: max_step min(d, 5 mm), tol 1e-4; fused
                            silica: max_step 20 um, tol 1e-5)
* ``sim.py:82-96``          curved mirror (``oracle.diagnostics.mirror``)
* ``sim.py:319-334``        one cavity cycle: forward segments, curved mirror, the segments in
                            reverse order, flat-mirror loss sqrt(T)

The envelope-solver machinery (``oracle.uppe3d.Oracle3D``) is reused with ``err_norm = "l2"``
and ``valid_mask = False``, which are exactly the two algorithmic differences of the older
solver.  ``SellmeierKerr.refractive_index`` is pinned against published fused-silica indices
in tests/test_oracle_mpc.py; the L2 norm against the L-infinity norm on single-mode fields and
against a brute-force sum.
"""
from dataclasses import dataclass

import numpy as np

from . import diagnostics as _diag
from .uppe3d import C_LIGHT, Oracle3D

# solver_3d.py:96-106: (B terms, C terms [um^2], valid wavelength range [m])
MATERIALS = {
    "fused_silica": ((0.6961663, 0.4079426, 0.8974794),
                     (0.0684043 ** 2, 0.1162414 ** 2, 9.896161 ** 2), (0.2e-6, 6.0e-6)),
    "sapphire": ((1.4313493, 0.65054713, 5.3414021),
                 (0.0726631 ** 2, 0.1193242 ** 2, 18.028251 ** 2), (0.2e-6, 5.5e-6)),
    "air": (None, None, None),
}


@dataclass
class SellmeierKerr:
    """The older solver's built-in medium: Sellmeier (or constant, for air) index and Kerr."""
    material: str = "fused_silica"
    n2: float = 3.0e-20
    center_wavelength: float = 800e-9
    use_shock: bool = True

    def __post_init__(self):
        self.lambda0 = self.center_wavelength
        self.c = C_LIGHT
        self.omega0 = 2 * np.pi * C_LIGHT / self.lambda0
        # unknown names fall back to fused silica (solver_3d.py:106)
        self.coeffs = MATERIALS.get(self.material, MATERIALS["fused_silica"])

    def refractive_index(self, omega):
        w = np.asarray(omega, np.float64)
        if self.material == "air":
            return np.ones_like(w) * 1.00027
        B, Cc, rng = self.coeffs
        w = np.where(np.abs(w) < 1e-12, 1e-12, w)
        lam = 2 * np.pi * C_LIGHT / np.abs(w)
        if rng:
            lam = np.clip(lam, rng[0], rng[1])
        l2 = (lam * 1e6) ** 2
        n2 = np.ones_like(lam)
        for b, c in zip(B, Cc):
            n2 = n2 + b * l2 / (l2 - c)
        return np.sqrt(n2)

    def shock(self, omega):
        return (omega + self.omega0) / self.omega0 if self.use_shock else np.ones_like(omega)

    def response(self, A, dt):
        kerr = 1j * (self.omega0 / C_LIGHT) * self.n2 * np.abs(A) ** 2 * A
        return kerr, np.zeros_like(A)


class OracleMPC3D(Oracle3D):
    err_norm = "l2"
    valid_mask = False


# sim.py:128-132: per-material controller settings of propagate_segment
def segment_settings(distance, material):
    if material == "air":
        return dict(max_step=min(distance, 5e-3), tol=1e-4)
    return dict(max_step=20e-6, tol=1e-5)


def propagate_segment(solver, A_norm, A_scale, distance, material):
    """sim.py:120-137 with the result kept in the time domain (the reference's
    ifft(ifftshift(A_spec)) of the shifted output spectrum is the same field)."""
    A_norm = np.nan_to_num(A_norm)
    z, A_phys, log = solver.propagate(A_norm * A_scale, distance, return_time=True,
                                      **segment_settings(distance, material))
    return A_phys / A_scale, log


def cavity_cycle(solvers, A_norm, A_scale, seq_forward, R_concave, T_round, omega, omega0,
                 x, y):
    """One round trip of sim.py:322-334; ``solvers`` maps 'air' / 'fs' to OracleMPC3D."""
    logs = []
    for dist, mat in seq_forward:
        A_norm, lg = propagate_segment(solvers[mat], A_norm, A_scale, dist, mat)
        logs.append(lg)
    A_norm = _diag.mirror(A_norm, x, y, omega, omega0, R_concave)
    for dist, mat in seq_forward[::-1]:
        A_norm, lg = propagate_segment(solvers[mat], A_norm, A_scale, dist, mat)
        logs.append(lg)
    return A_norm * np.sqrt(T_round), logs
