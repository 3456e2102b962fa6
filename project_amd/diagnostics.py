"""MPC runner diagnostics and mirror on the device (``src/3+1D-ERK4(3)IP UPPE/sim.py``).

``snapshot(solver, w0_mode)`` returns the quantities ``save_snapshot`` stores
(``sim.py:182-227``: total energy, beam centroid and rms widths, P(t), fluence, the field at the
peak-power time, the full-beam / centre / SMF-coupled spectra on the wavelength axis, the centre
and SMF-coupled temporal intensities and their FWHM) for the current field of a 3+1D solver
(after ``propagate`` or ``load``).  Every sum over the field runs in one device pass
(``uppe3d_diagnostics``, FP64 accumulation); the host only applies the unit factors, the
fftshift and the FWHM search on the 1D results.  ``save_snapshot_npz`` writes the reference's
``*_DIAG.npz`` keys.

``apply_mirror(solver, R, amp)`` applies the curved mirror of ``sim.py:82-96``
(A_w *= exp(-i (omega0 + omega) r^2 / (c R))) and an amplitude factor (the flat-mirror loss
sqrt(T) of ``sim.py:331``) to the solver's field in place, on the device.
"""
import ctypes

import numpy as np
import torch

from . import _lib

C0 = 299792458.0


def _fwhm(t, I):
    # half-maximum width around the peak with linear interpolation (sim.py:159-179)
    I = np.asarray(I, np.float64)
    t = np.asarray(t, np.float64)
    if I.size < 4:
        return np.nan
    imax = np.max(I)
    if not np.isfinite(imax) or imax <= 0:
        return np.nan
    half = 0.5 * imax
    ip = int(np.argmax(I))
    lo, hi = ip, ip
    while lo > 0 and I[lo] >= half:
        lo -= 1
    while hi < I.size - 1 and I[hi] >= half:
        hi += 1
    if hi <= lo + 1:
        return np.nan

    def at(i, j):
        return t[i] + (half - I[i]) * (t[j] - t[i]) / (I[j] - I[i] + 1e-30)

    w = (at(hi - 1, hi) if hi - 1 >= 0 else t[hi]) - (at(lo, lo + 1) if lo + 1 < I.size else t[lo])
    return float(w) if (w > 0 and np.isfinite(w)) else np.nan


def _geometry(solver):
    return float(solver.x[0]), solver.dx, float(solver.y[0]), solver.dy


def snapshot(solver, w0_mode, ix0=None, iy0=None):
    if solver._plan is None:
        raise RuntimeError("the solver holds no field (call propagate or load first)")
    lib = _lib.load()
    nx, ny, nt = solver.Nx, solver.Ny, solver.Nt
    ix0 = nx // 2 if ix0 is None else ix0
    iy0 = ny // 2 if iy0 is None else iy0
    dev = solver.device
    nb = ctypes.c_size_t()
    _lib.check(lib.uppe3d_diag_workspace_size(ctypes.byref(solver._desc), ctypes.byref(nb)),
               "uppe3d_diag_workspace_size")
    work = torch.empty(nb.value, dtype=torch.uint8, device=dev)
    out = torch.empty(8 * nt + nx * ny + 5, dtype=torch.float64, device=dev)
    x0, dx, y0, dy = _geometry(solver)
    s = _lib.stream_handle()
    _lib.check(lib.uppe3d_diagnostics(solver._plan, float(w0_mode), x0, dx, y0, dy, ix0, iy0,
                                      _lib.ptr(work), _lib.ptr(out), s), "uppe3d_diagnostics")
    o = out.cpu().numpy()
    sI = solver.A_scale ** 2
    dt = solver.dt
    q = o[:6 * nt].reshape(6, nt)
    spec_full = np.fft.fftshift(q[0]) * sI
    spec_smf = np.fft.fftshift(q[1] ** 2 + q[2] ** 2) * sI
    P_t = q[3] * dx * dy * sI
    I_smf = (q[4] ** 2 + q[5] ** 2) * sI
    spec_center = np.fft.fftshift(o[6 * nt:7 * nt]) * sI
    I_center = o[7 * nt:8 * nt] * sI
    flu = o[8 * nt:8 * nt + nx * ny].reshape(nx, ny) * dt * sI
    mom = o[8 * nt + nx * ny:]
    E = mom[0] * dx * dy * dt * sI
    it = int(np.argmax(P_t))
    axy = torch.empty((nx, ny), dtype=torch.complex64, device=dev)
    _lib.check(lib.uppe3d_field_at_time(solver._plan, it, _lib.ptr(axy), s),
               "uppe3d_field_at_time")
    A_xy = axy.cpu().numpy().astype(np.complex128) * solver.A_scale
    f_abs = np.asarray(solver.freq_axis_hz, np.float64) + C0 / solver.lambda0
    valid = f_abs > 1e9
    wl = np.zeros_like(f_abs)
    wl[valid] = C0 / f_abs[valid] * 1e9
    t = np.asarray(solver.t, np.float64)
    return dict(energy_total=float(E), x_mean_m=float(mom[1]), y_mean_m=float(mom[2]),
                wx_rms_m=float(mom[3]), wy_rms_m=float(mom[4]), t_s=t, P_t=P_t,
                I_center_t=I_center, I_smf_t=I_smf, wl_nm=wl, freq_abs_Hz=f_abs,
                spec_smf=spec_smf, spec_full_incoh=spec_full, spec_center=spec_center,
                fluence=flu, it_peak=it, A_xy=A_xy, fwhm_center=_fwhm(t, I_center),
                fwhm_smf=_fwhm(t, I_smf))


def save_snapshot_npz(path, snap, cycle, label):
    """The reference's ``cycle_XX_<label>_DIAG.npz`` keys (sim.py:222-227)."""
    keys = ["energy_total", "x_mean_m", "y_mean_m", "wx_rms_m", "wy_rms_m", "t_s", "P_t",
            "I_center_t", "I_smf_t", "wl_nm", "freq_abs_Hz", "spec_smf", "spec_full_incoh",
            "spec_center", "fluence", "it_peak"]
    np.savez(path, cycle=cycle, label=label, **{k: snap[k] for k in keys})


def apply_mirror(solver, R, amp=1.0):
    """In place on the solver's field: A_w *= amp exp(-i (omega0 + omega) r^2 / (c R));
    R = 0 applies amp only (a flat mirror with loss)."""
    if solver._plan is None:
        raise RuntimeError("the solver holds no field (call propagate or load first)")
    lib = _lib.load()
    x0, dx, y0, dy = _geometry(solver)
    _lib.check(lib.uppe3d_apply_mirror(solver._plan, float(R), float(amp), x0, dx, y0, dy,
                                       _lib.stream_handle()), "uppe3d_apply_mirror")
