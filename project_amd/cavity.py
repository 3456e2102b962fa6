"""Cavity / multipass slicing drivers on the device-resident slice chain.

Replaces the host loops of ``src/1D-ERK4(3)IP UPPE gas/gas.py:49-57,103-147`` (one pass of a
symmetric two-mirror cavity cut into slices, every slice's history kept) and
``"gas copy.py":104-141`` (N passes, one field kept per pass).  Those loops call
``solver.propagate`` once per slice after setting the slice's beam radius, area and gamma
from the Gaussian cavity mode; here all slices run in one kernel launch
(``ERK43IP_UPPE.propagate_slices``) and the results are assembled exactly as the reference
loops assemble them.
"""
import numpy as np


def calculate_cavity_mode(L_cavity, R_mirror, lambda0):
    """Waist and Rayleigh length of the eigenmode of a symmetric two-mirror cavity
    (gas.py:49-57): g = 1 - L/R must satisfy 0 <= g^2 <= 1;
    w0 = sqrt(L lambda / (2 pi) sqrt((1 + g) / (1 - g))), zR = pi w0^2 / lambda."""
    g = 1 - L_cavity / R_mirror
    if not (0 <= g ** 2 <= 1):
        raise ValueError("cavity parameters are outside the stable region")
    w0 = np.sqrt((L_cavity * lambda0) / (2 * np.pi) * np.sqrt((1 + g) / (1 - g)))
    zR = np.pi * w0 ** 2 / lambda0
    return w0, zR


def cavity_slices(solver, L_cavity, R_mirror, n_slices):
    """Per-slice (z_start, dz, w, gamma) of one pass, as gas.py:113-126 computes them:
    z_edges = linspace(-L/2, L/2, n + 1), w = w0 sqrt(1 + (z_mid / zR)^2),
    gamma = (2 pi / lambda0) n2 / (pi w^2)."""
    w0, zR = calculate_cavity_mode(L_cavity, R_mirror, solver.lambda0)
    z_edges = np.linspace(-L_cavity / 2, L_cavity / 2, n_slices + 1)
    z_start, z_end = z_edges[:-1], z_edges[1:]
    dz = z_end - z_start
    z_mid = (z_start + z_end) / 2
    w = w0 * np.sqrt(1 + (z_mid / zR) ** 2)
    gamma = (2 * np.pi / solver.lambda0) * (solver.n2 / (np.pi * w ** 2))
    return z_start, dz, w, gamma


def _set_last_slice(solver, w, gamma):
    # the reference loop leaves the solver holding the last slice's beam (SimResultAdapter)
    solver.beam_radius = w
    solver.beam_area = np.pi * w ** 2
    solver.gamma = gamma


def run_cavity_pass(solver, A0, t, L_cavity, R_mirror, n_slices, tol=1e-5, max_step_div=10,
                    min_step=1e-8):
    """One pass through the cavity in n_slices slices (gas.py:103-147).  Slice i propagates
    dz with tol, max_step = dz / max_step_div, min_step.  Returns (z, A_evolution) assembled
    as the reference: slice 0's history without its last entry, every later slice's history
    without its first and last entries (z shifted to the global coordinate), then
    (L_cavity, A_hist[-1] of the last slice)."""
    z_start, dz, w, gamma = cavity_slices(solver, L_cavity, R_mirror, n_slices)
    z_hists, A_hists, _ = solver.propagate_slices(A0, t, dz, gamma, tol=tol,
                                                  max_step=dz / max_step_div,
                                                  min_step=min_step, record="all")
    zs, As = [], []
    for i in range(n_slices):
        zg = z_hists[i] + (z_start[i] + L_cavity / 2)
        sl = slice(None, -1) if i == 0 else slice(1, -1)
        zs.extend(zg[sl])
        As.extend(A_hists[i][sl])
    zs.append(L_cavity)
    As.append(A_hists[-1][-1])
    _set_last_slice(solver, w[-1], gamma[-1])
    return np.array(zs), np.array(As)


def run_multipass(solver, A0, t, L_cavity, R_mirror, n_passes, n_slices_per_pass, tol=1e-5,
                  max_step_div=5, min_step=1e-8):
    """n_passes passes of n_slices_per_pass slices ("gas copy.py":104-141), all in one
    launch.  Returns (z, A_evolution): z = 0 and A0 first, then the accumulated length and the
    field after every pass."""
    z_start, dz, w, gamma = cavity_slices(solver, L_cavity, R_mirror, n_slices_per_pass)
    dzs = np.tile(dz, n_passes)
    A_out, _ = solver.propagate_slices(A0, t, dzs, np.tile(gamma, n_passes), tol=tol,
                                       max_step=dzs / max_step_div, min_step=min_step,
                                       record="last")
    ends = np.arange(1, n_passes + 1) * n_slices_per_pass - 1
    z = [0.0]
    zacc = 0.0
    for p in range(n_passes):
        for d in dz:
            zacc += d
        z.append(zacc)
    As = [np.asarray(A0, np.complex128).copy()] + [A_out[e] for e in ends]
    _set_last_slice(solver, w[-1], gamma[-1])
    return np.array(z), np.array(As)
