"""Spectrum input / output formats (host): the CSV reader's row rules, the output writer
against the reference's own output file and the Gaussian's closed-form spectrum, and
generate_pulse_from_spectrum against golden vectors written from the reference
(tests/golden/make_golden_spectrum.py)."""
import os

import numpy as np
import pytest

from project_amd import gas  # noqa: F401
from project_amd import spectrum_io as sio
from project_amd.erk43ip_method import ERK43IP_UPPE

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "golden_spectrum.npz"))


def test_reader_row_rules(tmp_path):
    f = tmp_path / "s.csv"
    f.write_text("\ufeff1063.9,0\n,PickedData1,\n1064.0,0.5,\n[3] 1064.1, 1.0\n\n1064.2,x\n"
                 "1064.3,0.25,7\n1064.4\n", encoding="utf-8")
    wl, I = sio.load_spectrum_csv(str(f))
    # the byte-order-mark row and the header are not numbers; a single value is not a row
    np.testing.assert_array_equal(wl, [1064.0, 1064.1, 1064.3])
    np.testing.assert_array_equal(I, [0.5, 1.0, 0.25])


def test_writer_grid_matches_reference_output(tmp_path):
    t = np.linspace(-60e-12, 60e-12, 2048)
    A = np.exp(-(t / 5e-12) ** 2) * np.exp(1j * 1e23 * t ** 2)
    f = tmp_path / "out.csv"
    n = sio.write_spectrum_csv(str(f), A, t, 1064e-9)
    lines = f.read_text(encoding="utf-8").splitlines()
    assert lines[0] == "Wavelength (nm),Normalized Intensity"
    wl_text = [ln.split(",")[0] for ln in lines[1:]]
    assert n == len(G["ref_output__wl_text"]) == 317
    assert wl_text == list(G["ref_output__wl_text"])
    assert all(len(ln.split(",")[1].split("e")[0].split(".")[1]) == 8 for ln in lines[1:])


def test_writer_intensity_closed_form():
    # A = exp(-t^2 / (2 T0^2)) has |FFT|^2 proportional to exp(-(2 pi f)^2 T0^2); the writer
    # multiplies by f_abs^2 and normalises by the maximum over all positive frequencies
    T0, lam0, c = 300e-15, 1064e-9, sio.C0
    t = np.linspace(-20e-12, 20e-12, 8192)
    A = np.exp(-t ** 2 / (2 * T0 ** 2))
    wl, I = sio.spectrum_rows(A, t, lam0, wl_min=1000, wl_max=1130)
    f_abs = c / (wl * 1e-9)
    f_rel = f_abs - c / lam0
    f_all = np.fft.fftfreq(len(t), t[1] - t[0]) + c / lam0
    f_all = f_all[f_all > 0]
    S_all = np.exp(-(2 * np.pi * (f_all - c / lam0)) ** 2 * T0 ** 2) * f_all ** 2
    expect = np.exp(-(2 * np.pi * f_rel) ** 2 * T0 ** 2) * f_abs ** 2 / S_all.max()
    np.testing.assert_allclose(I, expect, rtol=1e-9, atol=1e-12)
    assert np.all(np.diff(wl) > 0)


@pytest.mark.parametrize("name", ["input_spectrum_cubic", "input_linear_nojac",
                                  "sheet1_quadratic_tod", "input_spectrum_exact"])
def test_pulse_from_spectrum_matches_reference(name):
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "mgs", os.path.join(HERE, "golden", "make_golden_spectrum.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    fn, lam0, nt, win, E, gdd, tod, jac, kind, exact = m.CASES[name]
    s = ERK43IP_UPPE(material="fused_silica", n2=2.7e-20, beam_radius=1e-4,
                     center_wavelength=lam0, use_raman=False)
    t = G[name + "__t"]
    A = s.generate_pulse_from_spectrum(t, G[name + "__wl"], G[name + "__I"], E, GDD=gdd, TOD=tod,
                                       use_jacobian=jac, interp_kind=kind, exact_fourier=exact)
    ref = G[name + "__A"]
    np.testing.assert_allclose(A, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
    E_out = np.trapezoid(np.abs(A) ** 2, t)
    assert E_out == pytest.approx(E, rel=1e-12)
