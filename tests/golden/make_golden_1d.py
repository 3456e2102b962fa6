"""Writes golden vectors of the reference 1D ``propagate`` trajectory (run in the build
container only, where the reference is mounted read-only).

The components of the 1D solver are pinned by closed forms in ``tests/test_oracle_1d.py``;
the adaptive end-to-end trajectory (accepted-step sequence, FSAL rule, history sampling and
the FP64 final field) is pinned only by running the reference itself.  This script imports
``src/1D-ERK4(3)IP UPPE gas/erk43ip_method.py`` with the argon patch of ``gas.py:11-44``
applied (``gas.py`` is imported for its module-level patch; ``main()`` is not run) and saves
small cases to ``golden_1d.npz``: inputs, parameters, z history and the final field.

Usage (build container only):  python tests/golden/make_golden_1d.py /root/reference
"""
import contextlib
import io
import os
import sys

import numpy as np

CASES = {
    # name: (constructor kwargs, Nt, window, energy, fwhm, GDD, L, tol, max_step, min_step)
    "argon_ss": (dict(material="argon_gas", n2=1e-23, beam_radius=20e-6,
                      center_wavelength=800e-9, use_raman=False, use_self_steepening=True,
                      gas_P=1e5, gas_T=293.15),
                 512, 400e-15, 1e-6, 30e-15, 0.0, 0.05, 1e-6, 1e-2, 1e-11),
    "argon_7bar_chirp": (dict(material="argon_gas", n2=7e-23, beam_radius=50e-6,
                              center_wavelength=1030e-9, use_raman=False,
                              use_self_steepening=True, gas_P=7e5, gas_T=293.15),
                         1024, 2e-12, 20e-6, 150e-15, 2000e-30, 0.08, 1e-5, 8e-3, 1e-8),
    "silica_raman": (dict(material="fused_silica", n2=2.6e-20, beam_radius=30e-6,
                          center_wavelength=1064e-9, use_raman=True,
                          use_self_steepening=True),
                     512, 2e-12, 2e-9, 200e-15, 0.0, 0.01, 1e-5, 1e-3, 1e-11),
    "gnlse_gamma": (dict(material="sapphire", gamma=5e-3, center_wavelength=1550e-9,
                         use_raman=False, use_self_steepening=False),
                    256, 4e-12, 2e-10, 400e-15, 0.0, 0.5, 1e-6, 5e-2, 1e-11),
    "forced_min_step": (dict(material="fused_silica", n2=2.6e-20, beam_radius=10e-6,
                             center_wavelength=800e-9, use_raman=False,
                             use_self_steepening=True),
                        256, 1e-12, 5e-9, 100e-15, 0.0, 2e-4, 1e-16, 1e-4, 2e-5),
}


def main(ref_root):
    gas_dir = os.path.join(ref_root, "src", "1D-ERK4(3)IP UPPE gas")
    sys.path.insert(0, gas_dir)
    with contextlib.redirect_stdout(io.StringIO()):
        import matplotlib
        matplotlib.use("Agg")
        import gas  # noqa: F401  (applies the argon patch to ERK43IP_UPPE at import)
        from erk43ip_method import ERK43IP_UPPE
    out = {}
    for name, (kw, nt, win, energy, fwhm, gdd, L, tol, mx, mn) in CASES.items():
        with contextlib.redirect_stdout(io.StringIO()):
            s = ERK43IP_UPPE(**kw)
            t = np.linspace(-win / 2, win / 2, nt)
            A0 = s.generate_dispersed_pulse(t, energy, fwhm, GDD=gdd)
            A0 = np.asarray(A0, complex)
            z, Ah, w = s.propagate(A0, t, L, tol=tol, max_step=mx, min_step=mn)
        out[name + "__t"] = t
        out[name + "__A0"] = A0
        out[name + "__z"] = z
        out[name + "__A_final"] = Ah[-1]
        out[name + "__A_mid"] = Ah[len(Ah) // 2]
        out[name + "__nhist"] = np.array(len(Ah))
        out[name + "__omega"] = w
        out[name + "__scalars"] = np.array([L, tol, mx, mn, energy, fwhm, gdd, s.gamma, s.n0])
        print(name, "history", len(z), "final |A|max", np.abs(Ah[-1]).max())
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_1d.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
