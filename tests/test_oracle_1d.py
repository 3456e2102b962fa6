"""Pins the 1D FP64 oracle (oracle/uppe1d.py): against the reference's own trajectories
(golden vectors written by tests/golden/make_golden_1d.py from the reference solver), and
against closed forms and invariants that a dropped term, wrong sign or index would break."""
import importlib.util
import os

import numpy as np
import pytest
from scipy.fft import fft, ifft

from oracle import k1, uppe1d as o1

HERE = os.path.dirname(os.path.abspath(__file__))


def _cases():
    spec = importlib.util.spec_from_file_location(
        "make_golden_1d", os.path.join(HERE, "golden", "make_golden_1d.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m.CASES


CASES = _cases()


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_matches_reference_golden(name):
    g = np.load(os.path.join(HERE, "golden", "golden_1d.npz"))
    kw, nt, win, energy, fwhm, gdd, L, tol, mx, mn = CASES[name]
    s = o1.Oracle1D(o1.Params1D(**kw))
    t = g[name + "__t"]
    A0 = o1.dispersed_pulse(t, energy, fwhm, GDD=gdd)
    np.testing.assert_array_equal(A0, g[name + "__A0"])
    z, Ah, w, log = s.propagate(A0, t, L, tol=tol, max_step=mx, min_step=mn)
    np.testing.assert_array_equal(z, g[name + "__z"])
    assert len(Ah) == int(g[name + "__nhist"])
    np.testing.assert_array_equal(Ah[-1], g[name + "__A_final"])
    np.testing.assert_array_equal(Ah[len(Ah) // 2], g[name + "__A_mid"])
    if name == "forced_min_step":
        assert any(a == 2 for (_, _, a) in log)


def grid(n=512, win=2e-12):
    return np.linspace(-win / 2, win / 2, n)


def test_linear_propagation_is_exact():
    s = o1.Oracle1D(o1.Params1D(material="fused_silica", gamma=0.0, use_raman=False,
                                use_self_steepening=False, alpha=0.3))
    t = grid()
    A0 = o1.gaussian_pulse(t, 1e-9, 100e-15, chirp=2.0)
    L = 0.05
    z, Ah, w, log = s.propagate(A0, t, L, tol=1e-8, max_step=1e-2)
    exact = ifft(fft(A0) * np.exp(L * s.D))
    assert np.linalg.norm(Ah[-1] - exact) / np.linalg.norm(exact) < 1e-12
    assert all(e == 0.0 for (_, e, _) in log)


def test_spm_closed_form():
    s = o1.Oracle1D(o1.Params1D(material="fused_silica", gamma=2e-3, use_raman=False,
                                use_self_steepening=False))
    t = grid()
    A0 = o1.gaussian_pulse(t, 1e-9, 100e-15)
    s.setup(2 * np.pi * np.fft.fftfreq(len(t), t[1] - t[0]))
    setup = s.setup

    def no_dispersion(omega):          # the D = 0 monkeypatch of test_3d-1d.py:52-61
        setup(omega)
        s.D = np.zeros_like(s.D)
    s.setup = no_dispersion
    L = 0.5
    z, Ah, w, log = s.propagate(A0, t, L, tol=1e-11, max_step=0.1)
    # past 1000 entries the history only records when z advanced by > L/500, so the last
    # snapshot need not sit at L: compare at its own z
    exact = A0 * np.exp(1j * 2e-3 * np.abs(A0) ** 2 * z[-1])
    assert np.abs(2e-3 * np.abs(A0).max() ** 2 * L) > 1.0        # a real nonlinear phase
    assert np.linalg.norm(Ah[-1] - exact) / np.linalg.norm(exact) < 1e-8


def test_energy_conserved_without_shock():
    s = o1.Oracle1D(o1.Params1D(material="argon_gas", n2=1e-23, beam_radius=20e-6,
                                center_wavelength=800e-9, use_raman=False,
                                use_self_steepening=False, gas_P=1e5))
    t = grid(1024, 400e-15)
    A0 = o1.gaussian_pulse(t, 1e-6, 30e-15)
    z, Ah, w, log = s.propagate(A0, t, 0.5, tol=1e-10)
    e0, e1 = np.sum(np.abs(A0) ** 2), np.sum(np.abs(Ah[-1]) ** 2)
    assert abs(e1 / e0 - 1) < 1e-9


def test_weighted_photon_number_conserved_with_uppe_factor():
    """With self-steepening and the n0/n(w) factor, sum n(w)|A_w|^2/w_abs is invariant."""
    s = o1.Oracle1D(o1.Params1D(material="argon_gas", n2=1e-23, beam_radius=20e-6,
                                center_wavelength=800e-9, use_raman=False,
                                use_self_steepening=True, gas_P=5e5))
    t = grid(1024, 400e-15)
    A0 = o1.gaussian_pulse(t, 1e-6, 30e-15)
    z, Ah, w, log = s.propagate(A0, t, 0.5, tol=1e-10)
    wabs = s.omega0 + w
    q = lambda A: np.sum(s.n_w * np.abs(fft(A)) ** 2 / wabs)
    e = lambda A: np.sum(np.abs(A) ** 2)
    assert abs(q(Ah[-1]) / q(A0) - 1) < 1e-10


def test_fourth_order_convergence():
    s = o1.Oracle1D(o1.Params1D(material="fused_silica", n2=2.6e-20, beam_radius=10e-6,
                                center_wavelength=800e-9, use_raman=False,
                                use_self_steepening=True))
    t = grid(512, 1e-12)
    A0 = o1.gaussian_pulse(t, 1e-8, 100e-15)
    L = 0.05                                   # ~3 rad of nonlinear phase

    def run(nsteps):
        return s.propagate(A0, t, L, replay=[(L / nsteps, 0.0, 1)] * nsteps)[1][-1]
    ref = run(512)       # < 1000 steps: every step is recorded, so A_hist[-1] is at z = L
    e = [np.linalg.norm(run(n) - ref) for n in (16, 32, 64)]
    orders = np.log2(np.array(e[:-1]) / np.array(e[1:]))
    assert np.all(np.abs(orders - 4.0) < 0.35), orders


def test_plasma_term_matches_k1_and_formula():
    pl = o1.Plasma1D(K=8, sigma_K=1e-121, area=np.pi * (20e-6) ** 2)
    s = o1.Oracle1D(o1.Params1D(material="argon_gas", gamma=0.0, use_raman=False,
                                use_self_steepening=False, plasma=pl,
                                center_wavelength=800e-9))
    t = grid(1024, 400e-15)
    s.dt = t[1] - t[0]
    A = o1.gaussian_pulse(t, 1e-6, 30e-15)
    P = s.plasma_term(A, s.dt)
    I = np.abs(A) ** 2 / pl.area
    w = (pl.sigma_K * I ** 8).astype(np.float32)
    rho, gen = k1.k1_numpy(w, np.zeros_like(w), 1 / pl.tau_rec, s.dt, pl.rho_nt)
    sig = pl.drude(s.omega0)
    ref = -(gen * pl.Ui / (2 * I + 1e-30)) * A - sig / 2 * (1 + 1j * s.omega0 * pl.tau_c) * rho * A
    np.testing.assert_allclose(P, ref, rtol=1e-12, atol=1e-300)
    # density builds up during the pulse, then recombines at rate 1/tau_rec
    assert rho.max() > 1e18
    tail = rho[-200:]
    r = tail[1:] / tail[:-1]
    np.testing.assert_allclose(r, np.exp(-float(np.float32(1 / pl.tau_rec)) *
                                         float(np.float32(s.dt))), rtol=1e-9)
