"""Device-resident slice chain (ERK43IP_UPPE.propagate_slices and the cavity drivers) against
the FP64 oracle driven slice by slice the way gas.py:113-147 drives the reference (a fresh
propagate per slice with a new gamma, input A_hist[-1] of the previous slice), and against
repeated single propagate calls of this library.  Tolerance: relative L2 of the fields."""
import numpy as np
import pytest

from conftest import announce
from oracle import uppe1d as o1
from project_amd import cavity, gas  # noqa: F401  (argon patch)
from project_amd.erk43ip_method import ERK43IP_UPPE

pytestmark = pytest.mark.gpu

KW = dict(material="argon_gas", n2=1e-23 * 7, beam_radius=60e-6, center_wavelength=1030e-9,
          use_raman=False, use_self_steepening=True, gas_P=7e5, gas_T=293.15)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def pulse(nt, window=2e-12, energy=40e-6, fwhm=150e-15):
    t = np.linspace(-window / 2, window / 2, nt)
    return t, ERK43IP_UPPE(**KW).generate_gaussian_pulse(t, energy, fwhm)


def oracle_chain(A0, t, lengths, gammas, tol, max_steps, min_step):
    orc = o1.Oracle1D(o1.Params1D(**KW))
    A = A0.copy()
    out = []
    for L, g, mx in zip(lengths, gammas, max_steps):
        orc.gamma = g
        z, Ah, _, _ = orc.propagate(A, t, L, tol=tol, max_step=mx, min_step=min_step)
        out.append((z, Ah))
        A = Ah[-1]
    return out


def test_chain_vs_oracle_loop(gpu):
    t, A0 = pulse(1024)
    s = ERK43IP_UPPE(**KW)
    z_start, dz, w, gam = cavity.cavity_slices(s, 0.2865, 0.3, 6)
    ref = oracle_chain(A0, t, dz, gam, 1e-5, dz / 10, 1e-8)
    announce("cavity chain 6 slices nt=1024 vs oracle loop")
    zh, Ah, _ = s.propagate_slices(A0, t, dz, gam, tol=1e-5, max_step=dz / 10, min_step=1e-8)
    A_last, _ = s.propagate_slices(A0, t, dz, gam, tol=1e-5, max_step=dz / 10, min_step=1e-8,
                                   record="last")
    for i, (zo, Ao) in enumerate(ref):
        e = rel_l2(Ah[i][-1], Ao[-1])
        print("slice", i, "hist", len(zh[i]), len(zo), "rel L2", e)
        assert len(zh[i]) == len(zo) and abs(zh[i][-1] - zo[-1]) < 1e-12
        assert rel_l2(Ah[i][0], ref[i - 1][1][-1] if i else A0) < 2e-5
        assert e < 2e-5
        assert rel_l2(A_last[i], Ao[-1]) < 2e-5
    # the pulse broadened: the check is not trivial
    assert rel_l2(np.abs(np.fft.fft(Ah[-1][-1])), np.abs(np.fft.fft(A0))) > 1e-2


def test_chain_vs_single_calls_and_unrecorded_last_step(gpu):
    """Slice 1 takes more than 1000 accepted steps, so the history rule skips steps and the
    reference hands on the last recorded snapshot, not the final field."""
    t, A0 = pulse(256, window=1e-12)
    s = ERK43IP_UPPE(**KW)
    lengths = np.array([0.01, 0.05, 0.02])
    gam = np.array([s.gamma, 0.5 * s.gamma, 2.0 * s.gamma])
    mx = np.array([1e-3, 0.05 / 1202.0, 2e-3])
    ref = oracle_chain(A0, t, lengths, gam, 1e-6, mx, 1e-11)
    assert ref[1][0][-1] < 0.05 - 1e-9, "test setup: slice 1 must end on an unrecorded step"
    announce("chain vs single calls, nt=256, slice of >1000 steps")
    zh, Ah, _ = s.propagate_slices(A0, t, lengths, gam, tol=1e-6, max_step=mx, min_step=1e-11)
    A = A0
    for i in range(3):
        s.gamma = gam[i]
        z1, Ah1, _ = s.propagate(A, t, lengths[i], tol=1e-6, max_step=mx[i], min_step=1e-11)
        print("slice", i, "hist", len(zh[i]), len(z1), len(ref[i][0]),
              "chain vs calls", rel_l2(Ah[i][-1], Ah1[-1]),
              "chain vs oracle", rel_l2(Ah[i][-1], ref[i][1][-1]))
        assert len(zh[i]) == len(z1) == len(ref[i][0])
        # steps capped at max_step: the recorded positions agree to rounding
        np.testing.assert_allclose(zh[i], ref[i][0], rtol=1e-9, atol=1e-15)
        assert rel_l2(Ah[i][-1], Ah1[-1]) < 1e-5
        assert rel_l2(Ah[i][-1], ref[i][1][-1]) < 2e-5
        A = Ah1[-1]


def test_multipass_driver(gpu):
    t, A0 = pulse(1024)
    s = ERK43IP_UPPE(**KW)
    announce("multipass 3 passes x 4 slices nt=1024")
    z, A = cavity.run_multipass(s, A0, t, 0.2865, 0.3, 3, 4)
    z_start, dz, w, gam = cavity.cavity_slices(s, 0.2865, 0.3, 4)
    ref = oracle_chain(A0, t, np.tile(dz, 3), np.tile(gam, 3), 1e-5, np.tile(dz, 3) / 5, 1e-8)
    assert len(z) == 4 and z[-1] == pytest.approx(3 * 0.2865)
    for p in range(3):
        e = rel_l2(A[p + 1], ref[4 * p + 3][1][-1])
        print("pass", p, "rel L2", e)
        assert e < 2e-5
