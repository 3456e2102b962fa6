"""Writes golden vectors of the reference ``generate_pulse_from_spectrum``
(``src/1D-ERK4(3)IP UPPE gas/erk43ip_method.py:599-764``), run in the build container only,
where the reference is mounted read-only.

Nothing but the reference itself pins this function end to end (cubic interpolation of a
measured spectrum onto the simulation grid, the duplicate-frequency rule, the exact-Fourier
Simpson variant and its centre frequency).  Inputs are the reference's own measured-spectrum
CSVs, and the wavelength column of its output spectrum file is kept for the writer; the spectra are parsed here with project_amd.spectrum_io.load_spectrum_csv and stored in
the fixture with the outputs, so the test does not read the reference.

Usage (build container only):  python tests/golden/make_golden_spectrum.py /root/reference
"""
import contextlib
import io
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

CASES = {
    # name: (csv file, lambda0, Nt, window, energy, GDD, TOD, jacobian, kind, exact)
    "input_spectrum_cubic": ("input_spectrum.csv", 1064e-9, 2048, 120e-12, 10e-6,
                             2.36e-23, 0.0, True, "cubic", False),
    "input_linear_nojac": ("input.csv", 1064e-9, 1024, 60e-12, 1e-6, 0.0, 0.0, False,
                           "linear", False),
    "sheet1_quadratic_tod": ("Sheet1.csv", 1030e-9, 4096, 4e-12, 50e-6, -200e-30, 3000e-45,
                             True, "quadratic", False),
    "input_spectrum_exact": ("input_spectrum.csv", 1064e-9, 128, 60e-12, 10e-6, 1e-24, 0.0,
                             True, "cubic", True),
}


def main(ref_root):
    sys.path.insert(0, ROOT)
    from project_amd.spectrum_io import load_spectrum_csv
    sys.path.insert(0, os.path.join(ref_root, "src", "1D-ERK4(3)IP UPPE gas"))
    with contextlib.redirect_stdout(io.StringIO()):
        from erk43ip_method import ERK43IP_UPPE
    out = {}
    for name, (fn, lam0, nt, win, E, gdd, tod, jac, kind, exact) in CASES.items():
        wl, I = load_spectrum_csv(os.path.join(ref_root, fn))
        s = ERK43IP_UPPE(material="fused_silica", n2=2.7e-20, beam_radius=1e-4,
                         center_wavelength=lam0, use_raman=False)
        t = np.linspace(-win / 2, win / 2, nt)
        with contextlib.redirect_stdout(io.StringIO()):
            A = s.generate_pulse_from_spectrum(t, wl, I, E, GDD=gdd, TOD=tod, use_jacobian=jac,
                                               interp_kind=kind, exact_fourier=exact)
        out[name + "__wl"] = wl
        out[name + "__I"] = I
        out[name + "__t"] = t
        out[name + "__A"] = np.asarray(A, complex)
        print(name, len(wl), np.abs(A).max())
    # the reference's own output spectrum file (written by
    # test_spectrum_pulse-coupling_efficiency-out.py:336-383 on Nt 2048, 120 ps, 1064 nm):
    # its wavelength column pins the writer's grid, range, order and format
    rows = np.loadtxt(os.path.join(ref_root, "output_final_spectrum_coupling.csv"),
                      delimiter=",", skiprows=1, dtype=str)
    out["ref_output__wl_text"] = rows[:, 0]
    out["ref_output__I"] = rows[:, 1].astype(float)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                     "golden_spectrum.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
