"""The MPC diagnostics oracle (oracle/diagnostics.py) against closed forms: Gaussian energy,
fluence centroid and rms width, temporal FWHM, the SMF overlap of a beam in the fibre mode,
Parseval, and the curved mirror on a single-frequency field."""
import numpy as np
import pytest

from oracle import diagnostics as od


def grid(n=64, nt=256, hw=3e-3, T=400e-15):
    x = np.linspace(-hw, hw, n)
    t = np.linspace(-T / 2, T / 2, nt)
    return x, x.copy(), t


def test_energy_moments_fwhm_closed_forms():
    x, y, t = grid(96, 512)
    w, tau, xc, yc = 0.7e-3, 40e-15, 0.2e-3, -0.1e-3
    X, Y = np.meshgrid(x, y, indexing="ij")
    sp = np.exp(-((X - xc) ** 2 + (Y - yc) ** 2) / w ** 2)
    f = sp[:, :, None] * np.exp(-t ** 2 / tau ** 2)[None, None, :]
    dx, dt = x[1] - x[0], t[1] - t[0]
    # |A|^2 integrates to (pi w^2 / 2) (tau sqrt(pi/2))
    assert od.energy(f, dx, dx, dt) == pytest.approx(np.pi * w ** 2 / 2 * tau * np.sqrt(np.pi / 2),
                                                     rel=1e-9)
    s = od.snapshot(f, x, y, t, 2.0, 1e-3, 800e-9, np.fft.fftshift(np.fft.fftfreq(len(t), dt)))
    assert s["energy_total"] == pytest.approx(4 * od.energy(f, dx, dx, dt), rel=1e-12)
    # intensity exp(-2 r^2 / w^2): centroid (xc, yc), rms width w / 2
    assert s["x_mean_m"] == pytest.approx(xc, abs=1e-12)
    assert s["y_mean_m"] == pytest.approx(yc, abs=1e-12)
    assert s["wx_rms_m"] == pytest.approx(w / 2, rel=1e-9)
    assert s["wy_rms_m"] == pytest.approx(w / 2, rel=1e-9)
    # intensity exp(-2 t^2 / tau^2): FWHM tau sqrt(2 ln 2) (linear interpolation on the grid)
    assert s["fwhm_center"] == pytest.approx(tau * np.sqrt(2 * np.log(2)), rel=2e-3)
    assert s["it_peak"] in (len(t) // 2 - 1, len(t) // 2)
    assert s["P_t"].sum() * dt == pytest.approx(s["energy_total"], rel=1e-12)


def test_smf_overlap_and_parseval():
    x, y, t = grid(64, 128)
    w0 = 0.8e-3
    X, Y = np.meshgrid(x, y, indexing="ij")
    mode = np.exp(-(X ** 2 + Y ** 2) / w0 ** 2)
    rng = np.random.default_rng(5)
    g = np.exp(-t ** 2 / (60e-15) ** 2) * np.exp(1j * rng.standard_normal(len(t)) * 0.3)
    f = mode[:, :, None] * g[None, None, :]
    dt = t[1] - t[0]
    s = od.snapshot(f, x, y, t, 1.0, w0, 800e-9, np.fft.fftshift(np.fft.fftfreq(len(t), dt)))
    ov = np.sum(mode ** 2)
    np.testing.assert_allclose(s["spec_smf"], s["spec_full_incoh"] * ov, rtol=1e-10,
                               atol=1e-12 * s["spec_smf"].max())
    # Parseval for the unnormalised FFT
    assert s["spec_full_incoh"].sum() == pytest.approx(len(t) * np.sum(np.abs(f) ** 2), rel=1e-12)
    np.testing.assert_allclose(s["I_smf_t"], np.abs(g) ** 2 * ov ** 2, rtol=1e-10)


def test_mirror_on_single_frequency():
    x, y, t = grid(32, 64)
    nt = len(t)
    omega = 2 * np.pi * np.fft.fftfreq(nt, t[1] - t[0])
    X, Y = np.meshgrid(x, y, indexing="ij")
    g = np.exp(-(X ** 2 + Y ** 2) / (1e-3) ** 2)
    m = 5
    f = g[:, :, None] * np.exp(2j * np.pi * m * np.arange(nt) / nt)[None, None, :]
    omega0, R = 2 * np.pi * od.C0 / 1064e-9, 0.3
    out = od.mirror(f, x, y, omega, omega0, R)
    k = (omega0 + omega[m]) / od.C0
    expect = f * np.exp(-1j * k * (X ** 2 + Y ** 2) / R)[:, :, None]
    np.testing.assert_allclose(out, expect, rtol=0, atol=1e-12)
