"""Pins the 3+1D FP64 oracle (oracle/uppe3d.py, a restatement of solver_3d.py and
media_physics.py, which need cupy and cannot run here) by closed forms and invariants."""
import numpy as np
import pytest
import scipy.fft as sfft

from oracle import uppe3d as o3
from project_amd import workloads as wl


def small(nx=16, nt=64, energy=0.8e-3, n4=True, shock=False, guard=0.0, hw=4e-3):
    x, y, t = wl.grid3d(nx, nt, half_width=hw)
    A0 = wl.beam3d(x, y, t, energy=energy)
    m = o3.GasMedium("Argon", 800e-9, 1.0, use_shock=shock, n4_enabled=n4, tunnel_guard=guard)
    return A0, o3.Oracle3D(x, y, t, m)


def strip_boundaries(o):
    o.filt = np.ones_like(o.filt)
    o.damp = np.zeros_like(o.damp)


def test_linear_propagation_exact_in_k_space():
    A0, o = small()
    o.m.n2 = 0.0
    o.m.n4 = 0.0
    o.m.lut_W = np.zeros_like(o.m.lut_W)
    strip_boundaries(o)
    L = 0.3
    z, f, Aw, log = o.propagate(A0, L, tol=1e-6, max_step=0.05)
    Ak0 = sfft.fftn(A0 / o.A_scale) * o.A_scale
    exact = sfft.fftshift(sfft.ifftn(Ak0 * np.exp(L * o.D), axes=(0, 1)), axes=-1)
    assert np.linalg.norm(Aw - exact) / np.linalg.norm(exact) < 1e-12
    assert max(e for (_, e, _) in log) == 0.0


def test_dispersion_operator_equals_kz_minus_kref():
    """The difference-of-squares form of solver_3d.py:80-84 equals i (kz - k_ref)."""
    A0, o = small(nx=32, nt=128)
    kz = np.sqrt((o.k[None, None, :] ** 2 - o.kperp2[:, :, None]).astype(complex))
    direct = 1j * (kz - o.kref[None, None, :]) * o.mask
    np.testing.assert_allclose(o.D, direct, rtol=1e-9, atol=1e-9 * np.abs(direct).max())
    # forward-difference k1 with dw = omega0 1e-3 is the group index to O(dw): 1/v_g ~ n_g / c
    assert 1.0 < o.k1 * o3.C_LIGHT < 1.001


def test_energy_conserved_kerr_hoke_no_plasma():
    A0, o = small(energy=0.3e-3)
    o.m.lut_W = np.zeros_like(o.m.lut_W)
    strip_boundaries(o)
    z, At, log = o.propagate(A0, 0.01, tol=1e-9, return_time=True)
    e = lambda A: np.sum(np.abs(A) ** 2)
    assert abs(e(At) / e(A0) - 1) < 1e-8
    assert len(log) > 3


def test_fourth_order_convergence_3d():
    """solver README claim err(2h)/err(h/2) = 16: global error ~ h^4 at fixed steps."""
    A0, o = small(energy=0.3e-3)
    o.m.lut_W = np.zeros_like(o.m.lut_W)
    strip_boundaries(o)
    L = 2e-3

    def run(n):
        return o.propagate(A0, L, replay=[(L / n, 0.0, 1)] * n, return_time=True)[1]
    ref = run(256)
    e = [np.linalg.norm(run(n) - ref) for n in (4, 8, 16)]
    orders = np.log2(np.array(e[:-1]) / np.array(e[1:]))
    assert np.all(np.abs(orders - 4.0) < 0.4), orders


def test_lut_interpolation_and_guard_quirk():
    m = o3.GasMedium("Argon", 800e-9, 1.0, tunnel_guard=0.0)
    I = 10 ** m.lut_log10_I[[0, 500, 1999]]
    np.testing.assert_allclose(m.ionization_rate(I), m.keldysh_rate(I), rtol=1e-12)
    # far below the tunnelling regime the rate is the multiphoton term sigma_K I^K
    assert abs(m.keldysh_rate(np.array([1e16]))[0] / (m.sigma_K * 1e16 ** 11) - 1) < 1e-12
    # the reference's "+1e-20" guard dominates e*hbar*E: its tunnelling term is ~1e15 s^-1
    g = o3.GasMedium("Argon", 800e-9, 1.0)
    low = 10 ** g.lut_log10_I < 1e16
    assert np.allclose(g.lut_W[low], 1e15, rtol=1e-6) and np.all(g.lut_W >= 1e15 * (1 - 1e-6))


def test_response_matches_k1_with_float32_rates():
    A0, o = small(nx=16, nt=64, energy=30e-3, hw=1e-3)
    At = A0 * 5
    kerr, plas = o.m.response(At, o.dt)
    I = np.abs(At) ** 2
    g = o.m.omega0 / o3.C_LIGHT
    np.testing.assert_allclose(kerr, 1j * g * (o.m.n2 * I + o.m.n4 * I * I) * At, rtol=1e-12)
    assert np.abs(plas).max() > 0


# -- invariants that pin the assembly of the nonlinear operator N (media_physics.py:156-205,
#    solver_3d.py:107-132) -------------------------------------------------------------------
def _rates(o, A, weight=None):
    """d/dz of sum_k w_k |A_k|^2 / Ntot from N alone (A in the time domain, A_scale = 1)."""
    Ak = sfft.fftn(A)
    Nk = o.N(Ak)
    w = 1.0 if weight is None else weight
    return 2.0 * np.real(np.sum(np.conj(Ak) * Nk * w)) / A.size, np.sum(np.abs(Ak * Nk * w)) / A.size


def _plasma_case(shock=False):
    x, y, t = wl.grid3d(16, 128, half_width=1e-3)
    A = wl.beam3d(x, y, t, energy=30e-3, w0=0.3e-3)      # peak ~5e18 W/m^2: strong plasma
    m = o3.GasMedium("Argon", 800e-9, 1.0, use_shock=shock, tunnel_guard=0.0)
    o = o3.Oracle3D(x, y, t, m)
    assert np.all(o.mask == 1.0)                          # no negative absolute frequency
    return A, o


def _balance(o, A):
    """(rate of sum |A|^2 from N, -sum(gen Ui + sigma rho I) from the K1 recurrence)."""
    I = np.abs(A) ** 2
    w_ion = o.m.ionization_rate(I).astype(np.float32)
    rho, gen = o3._k1.k1(w_ion, np.zeros_like(w_ion), 1.0 / o.m.tau_rec, o.dt, o.m.rho_nt)
    return _rates(o, A)[0], -np.sum(gen * o.m.Ui + o.m.sigma_drude * rho * I)


def test_energy_balance_with_plasma():
    """Kerr and HOKE are energy-neutral (i real I A), so d/dz sum|A|^2 = -sum(gen Ui + sigma rho
    I): the loss term removes the ionisation energy, the defocus term's real part the Drude
    absorption.  A flipped sign, a factor of 2 or a lost 1/2 in either term breaks it."""
    A, o = _plasma_case()
    got, want = _balance(o, A)
    assert want < 0 and abs(got - want) < 1e-9 * abs(want), (got, want)

    class Mutant(o3.GasMedium):
        def response(self, A, dt):
            k, p = super().response(A, dt)
            return k, self.scale * p
    for scale in (-1.0, 2.0, 0.5):
        mm = Mutant("Argon", 800e-9, 1.0, tunnel_guard=0.0)
        mm.scale = scale
        om = o3.Oracle3D(o.x, o.y, o.t, mm)
        g2, w2 = _balance(om, A)
        assert abs(g2 - w2) > 0.1 * abs(w2), scale


def test_photon_number_conserved_with_shock():
    """With the shock operator (omega + omega0)/omega0 on the Kerr and HOKE terms, N conserves
    the photon number sum |A_w|^2 / omega_abs (the envelope form of the Kerr response's
    Manley-Rowe relation); without the factor, or with it inverted, it does not."""
    x, y, t = wl.grid3d(16, 128)
    # a chirped, asymmetric pulse (a real, even envelope makes every variant conserve it)
    A = wl.beam3d(x, y, t, energy=0.8e-3, tau_fwhm=10e-15) * np.exp(
        1j * (t / 20e-15) ** 2 + 1j * (t / 30e-15) ** 3)[None, None, :]
    for variant in ("reference", "no shock factor", "inverted"):
        m = o3.GasMedium("Argon", 800e-9, 1.0, use_shock=True, tunnel_guard=0.0)
        m.lut_W = np.zeros_like(m.lut_W)
        if variant == "no shock factor":
            m.shock = lambda w: np.ones_like(w)
        elif variant == "inverted":
            m.shock = lambda w, m=m: m.omega0 / (w + m.omega0)
        o = o3.Oracle3D(x, y, t, m)
        wabs = (o.omega + m.omega0)[None, None, :]
        rate, scale = _rates(o, A, 1.0 / wabs)
        if variant == "reference":
            assert abs(rate) < 1e-12 * scale, (rate, scale)
        else:
            assert abs(rate) > 1e-6 * scale, (variant, rate, scale)


def test_shock_applies_to_kerr_only():
    """With plasma and shock on, the photon-number rate is the plasma term's alone, taken
    without the shock factor (solver_3d.py:128: shock * Kerr_w + Plasma_w)."""
    A, o = _plasma_case(shock=True)
    wabs = (o.omega + o.m.omega0)[None, None, :]
    rate, _ = _rates(o, A, 1.0 / wabs)
    Ak = sfft.fftn(A)
    _, plas = o.m.response(A, o.dt)
    Pk = sfft.fftn(plas) * o.mask
    want = 2.0 * np.real(np.sum(np.conj(Ak) * Pk / wabs)) / A.size
    assert abs(rate - want) < 1e-9 * abs(want)
    # the shock factor on the plasma term would weight it by omega_abs / omega0
    wrong = 2.0 * np.real(np.sum(np.conj(Ak) * Pk * (wabs / o.m.omega0) / wabs)) / A.size
    assert abs(wrong - want) > 1e-4 * abs(want)


def test_uniform_transverse_field_reduces_to_1d():
    """A field uniform in x and y (k_perp = 0 only) obeys the 1D equation: the 3D D at
    k_perp = 0 differs from the 1D D of the same medium only by the k1 term (forward vs
    central difference, i (k1_1D - k1_3D) omega), and one ERK4(3)-IP step of the 3D stepper
    equals one step of the 1D stepper (pinned bit-exactly by the reference's own trajectories)
    with that D, gamma = omega0 n2 / c and no self-steepening, to rounding."""
    from oracle import uppe1d as o1
    x, y, t = wl.grid3d(8, 256)
    m = o3.GasMedium("Argon", 800e-9, 1.0, n4_enabled=False, tunnel_guard=0.0)
    m.lut_W = np.zeros_like(m.lut_W)
    o = o3.Oracle3D(x, y, t, m)
    strip_boundaries(o)
    o.damp = np.zeros_like(o.damp)
    a = np.sqrt(1e17) * np.exp(-2 * np.log(2) * t ** 2 / 40e-15 ** 2) * np.exp(
        1j * (t / 50e-15) ** 2)                               # 1e17 W/m^2 peak, chirped
    A = np.broadcast_to(a, (8, 8, len(t))).copy()
    p = o1.Params1D(material="argon_gas", gamma=m.omega0 / o3.C_LIGHT * m.n2,
                    center_wavelength=800e-9, use_raman=False, use_self_steepening=False)
    o1d = o1.Oracle1D(p)
    o1d.index = lambda lam: m.refractive_index(2 * np.pi * o3.C_LIGHT / np.asarray(lam))
    o1d.dt = t[1] - t[0]
    o1d.setup(o.omega)
    dk1 = (o.D[0, 0, 1:] - o1d.D[1:]) / (1j * o.omega[1:])
    np.testing.assert_allclose(dk1, o1d.k1 - o.k1, rtol=1e-4)   # FP64 cancellation in D3 - D1
    o1d.D = o.D[0, 0, :].copy()
    h = 2e-3
    A3, _ = o.step(A, h)
    A1, _, _ = o1d.step(a, h)
    assert np.abs(A3 - A3[0, 0]).max() < 1e-12 * np.abs(A3).max()
    assert np.linalg.norm(A3[0, 0] - A1) / np.linalg.norm(A1) < 1e-12
    assert np.linalg.norm(A1 - a) / np.linalg.norm(a) > 1e-3     # the step did something
