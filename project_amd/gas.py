"""Argon patch of the 1D solver (drop-in for ``src/1D-ERK4(3)IP UPPE gas/gas.py:11-44``).

Importing this module patches :class:`project_amd.erk43ip_method.ERK43IP_UPPE` exactly as the
reference patches its class at import time:

* ``__init__`` accepts ``gas_P`` (Pa, default 101325) and ``gas_T`` (K, default 293.15);
* ``_compute_refractive_index`` uses the Peck & Fisher (1964) argon formula at 15 C / 1 atm,
  (n - 1) 1e8 = 1205582.27 / (130 - s^2) + 317222.05 / (38.9 - s^2), s = 1/lambda[um],
  scaled by the ideal-gas density ratio (P / 101325) (288.15 / T) when
  ``material == 'argon_gas'``; other materials keep the Sellmeier table.

The formula's constants are the published Peck & Fisher values the reference uses.
``calculate_cavity_mode`` (gas.py:49-57) is re-exported from :mod:`project_amd.cavity`, which
also holds the slicing drivers of gas.py's ``main``.
"""
from .erk43ip_method import ERK43IP_UPPE, SimResultAdapter  # noqa: F401
from .cavity import calculate_cavity_mode  # noqa: F401  (gas.py:49-57)

_orig_compute_ri = ERK43IP_UPPE._compute_refractive_index
_orig_init = ERK43IP_UPPE.__init__


def patched_init(self, *args, **kwargs):
    self.gas_P = kwargs.pop('gas_P', 101325.0)
    self.gas_T = kwargs.pop('gas_T', 293.15)
    _orig_init(self, *args, **kwargs)


def patched_compute_ri(self, wavelength):
    if getattr(self, 'material', '') == 'argon_gas':
        import numpy as np
        s2 = 1.0 / (np.asarray(wavelength, np.float64) * 1e6) ** 2
        nm1 = (1205582.27 / (130.0 - s2) + 317222.05 / (38.9 - s2)) * 1e-8
        return 1.0 + nm1 * (self.gas_P / 101325.0) * (288.15 / self.gas_T)
    return _orig_compute_ri(self, wavelength)


if not getattr(ERK43IP_UPPE, '_argon_patched', False):
    ERK43IP_UPPE.__init__ = patched_init
    ERK43IP_UPPE._compute_refractive_index = patched_compute_ri
    ERK43IP_UPPE._argon_patched = True
