"""Cavity slicing drivers, host logic (no GPU): the cavity eigenmode against the resonator
closed forms, the per-slice gamma rule, and the assembly of per-slice histories against the
reference loops (gas.py:113-147, "gas copy.py":115-141) run over a synthetic propagator."""
import numpy as np
import pytest

from project_amd import cavity


def test_cavity_mode_closed_forms():
    lam = 1030e-9
    for L, R in [(0.2865, 0.3), (0.5, 1.0), (0.3, 0.3), (1.2, 0.7)]:
        w0, zR = cavity.calculate_cavity_mode(L, R, lam)
        # symmetric resonator: zR = sqrt(L (2R - L)) / 2, and the Gaussian beam's wavefront
        # radius at the mirrors (z = L/2) equals the mirror radius
        assert zR == pytest.approx(np.sqrt(L * (2 * R - L)) / 2, rel=1e-12)
        z = L / 2
        assert z * (1 + (zR / z) ** 2) == pytest.approx(R, rel=1e-12)
        assert zR == pytest.approx(np.pi * w0 ** 2 / lam, rel=1e-12)
    # confocal cavity: zR = L / 2
    assert cavity.calculate_cavity_mode(0.3, 0.3, lam)[1] == pytest.approx(0.15, rel=1e-12)
    with pytest.raises(ValueError):
        cavity.calculate_cavity_mode(0.7, 0.3, lam)   # g = -4/3: unstable


class FakeSolver:
    """Deterministic stand-in for the GPU solver: a 'propagation' that scales the field by
    (1 + i gamma L) and returns a history whose length depends on the slice."""

    def __init__(self, n2=1e-23, lambda0=1030e-9):
        self.n2, self.lambda0 = n2, lambda0
        self.gamma = 0.0
        self.beam_radius = self.beam_area = None

    def propagate(self, A0, t, L, tol=1e-4, max_step=1e-2, min_step=1e-11):
        m = 2 + int(round(L / max_step)) % 5
        z = np.linspace(0.0, L, m)
        A = np.array([A0 * (1 + 1j * self.gamma * zz) for zz in z])
        return z, A, None

    def propagate_slices(self, A0, t, lengths, gammas, tol, max_step, min_step, record):
        A = np.asarray(A0, np.complex128)
        zs, As, outs = [], [], []
        mx = np.broadcast_to(max_step, np.shape(lengths))
        for L, g, m in zip(lengths, gammas, mx):
            self.gamma = g
            z, Ah, _ = self.propagate(A, t, L, tol, m, min_step)
            zs.append(z)
            As.append(Ah)
            A = Ah[-1]
            outs.append(A)
        if record == "last":
            return np.array(outs), None
        return zs, As, None


def reference_pass(solver, A0, t, L_cavity, R_mirror, N_slices):
    # the slicing loop of gas.py:103-147, over the solver's propagate
    w0, zR = cavity.calculate_cavity_mode(L_cavity, R_mirror, solver.lambda0)
    z_edges = np.linspace(-L_cavity / 2, L_cavity / 2, N_slices + 1)
    A_current = A0.copy()
    master_z, master_A = [], []
    for i in range(N_slices):
        z_start, z_end = z_edges[i], z_edges[i + 1]
        dz = z_end - z_start
        z_mid = (z_start + z_end) / 2
        w = w0 * np.sqrt(1 + (z_mid / zR) ** 2)
        solver.beam_radius = w
        solver.beam_area = np.pi * w ** 2
        solver.gamma = (2 * np.pi / solver.lambda0) * (solver.n2 / solver.beam_area)
        z_hist, A_hist, _ = solver.propagate(A_current, t, L=dz, tol=1e-5, max_step=dz / 10,
                                             min_step=1e-8)
        gz = z_hist + (z_start + L_cavity / 2)
        if i == 0:
            master_z.extend(gz[:-1])
            master_A.extend(A_hist[:-1])
        else:
            master_z.extend(gz[1:-1])
            master_A.extend(A_hist[1:-1])
        A_current = A_hist[-1]
    master_z.append(L_cavity)
    master_A.append(A_current)
    return np.array(master_z), np.array(master_A), solver.gamma, solver.beam_radius


def test_pass_assembly_matches_reference_loop():
    t = np.linspace(-1, 1, 8)
    A0 = np.exp(-t ** 2) + 0j
    ref = FakeSolver()
    z_r, A_r, g_r, w_r = reference_pass(ref, A0, t, 0.2865, 0.3, 7)
    s = FakeSolver()
    z, A = cavity.run_cavity_pass(s, A0, t, 0.2865, 0.3, 7)
    np.testing.assert_allclose(z, z_r, rtol=0, atol=1e-15)
    np.testing.assert_allclose(A, A_r, rtol=1e-15)
    assert s.gamma == pytest.approx(g_r, rel=1e-15) and s.beam_radius == pytest.approx(w_r)


def test_multipass_assembly_matches_reference_loop():
    t = np.linspace(-1, 1, 8)
    A0 = np.exp(-t ** 2) + 0j
    L, R, P, S = 0.2865, 0.3, 3, 4
    # the per-pass loop of "gas copy.py":115-141
    ref = FakeSolver()
    w0, zR = cavity.calculate_cavity_mode(L, R, ref.lambda0)
    z_edges = np.linspace(-L / 2, L / 2, S + 1)
    A_cur = A0.copy()
    zr, Ar, gz = [0.0], [A0.copy()], 0.0
    for _ in range(P):
        for i in range(S):
            dz = z_edges[i + 1] - z_edges[i]
            zm = (z_edges[i] + z_edges[i + 1]) / 2
            w = w0 * np.sqrt(1 + (zm / zR) ** 2)
            ref.gamma = (2 * np.pi / ref.lambda0) * (ref.n2 / (np.pi * w ** 2))
            _, Ah, _ = ref.propagate(A_cur, t, L=dz, tol=1e-5, max_step=dz / 5, min_step=1e-8)
            A_cur = Ah[-1]
            gz += dz
        zr.append(gz)
        Ar.append(A_cur.copy())
    z, A = cavity.run_multipass(FakeSolver(), A0, t, L, R, P, S)
    np.testing.assert_allclose(z, zr, rtol=1e-15)
    np.testing.assert_allclose(A, np.array(Ar), rtol=1e-15)
