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
