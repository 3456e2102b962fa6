"""Test infrastructure (the oracle): FP64 NumPy restatement of the 3+1D MPC runner's snapshot
diagnostics and mirror, used only by tests/ as the checker of the device kernels.

Follows ``src/3+1D-ERK4(3)IP UPPE/sim.py``:
* ``mirror``        :82-96   curved mirror, A_w *= exp(-i (omega0 + omega)/c r^2 / R)
* ``energy``        :140-143 sum |A|^2 dx dy dt (FP64)
* ``beam_moments``  :145-157 centroid and rms widths of the fluence
* ``fwhm``          :159-179 half-maximum width around the peak, linear interpolation
* ``snapshot``      :182-220 P(t), fluence, peak-time slice, full / centre / SMF-coupled
                             spectra, centre and SMF-coupled temporal intensities
The field is normalised (A / A_scale, layout [x][y][t]); ``scale_I`` = A_scale^2 restores
physical units, as in the reference.  Every function here is pinned by closed forms in
tests/test_oracle_diagnostics.py.
"""
import numpy as np

C0 = 299792458.0


def mirror(field_norm, x, y, omega, omega0, R):
    X, Y = np.meshgrid(x, y, indexing="ij")
    k = (omega0 + omega) / C0
    ph = np.exp(-1j * k[None, None, :] * (X ** 2 + Y ** 2)[:, :, None] / R)
    return np.fft.ifft(np.fft.fft(field_norm, axis=-1) * ph, axis=-1)


def energy(field_norm, dx, dy, dt):
    return float(np.sum(np.abs(field_norm).astype(np.float64) ** 2) * dx * dy * dt)


def beam_moments(fluence, x, y, dx, dy):
    X, Y = np.meshgrid(x, y, indexing="ij")
    f = fluence.astype(np.float64)
    tot = np.sum(f) * dx * dy + 1e-30
    xm = np.sum(X * f) * dx * dy / tot
    ym = np.sum(Y * f) * dx * dy / tot
    x2 = np.sum((X - xm) ** 2 * f) * dx * dy / tot
    y2 = np.sum((Y - ym) ** 2 * f) * dx * dy / tot
    return float(xm), float(ym), float(np.sqrt(x2)), float(np.sqrt(y2))


def fwhm(t, I):
    I = np.asarray(I, np.float64)
    t = np.asarray(t, np.float64)
    if I.size < 4:
        return np.nan
    imax = np.max(I)
    if not np.isfinite(imax) or imax <= 0:
        return np.nan
    half = 0.5 * imax
    ip = int(np.argmax(I))
    il = ip
    while il > 0 and I[il] >= half:
        il -= 1
    ir = ip
    while ir < I.size - 1 and I[ir] >= half:
        ir += 1
    if ir <= il + 1:
        return np.nan

    def cross(i1, i2):
        return t[i1] + (half - I[i1]) * (t[i2] - t[i1]) / (I[i2] - I[i1] + 1e-30)

    tl = cross(il, il + 1) if il + 1 < I.size else t[il]
    tr = cross(ir - 1, ir) if ir - 1 >= 0 else t[ir]
    w = tr - tl
    return float(w) if (w > 0 and np.isfinite(w)) else np.nan


def snapshot(field_norm, x, y, t, A_scale, w0_mode, lambda0, freq_axis_hz):
    """The quantities save_snapshot stores (sim.py:182-227), FP64 sums."""
    dx, dy, dt = x[1] - x[0], y[1] - y[0], t[1] - t[0]
    f = np.asarray(field_norm, np.complex128)
    sI = A_scale ** 2
    nx, ny, nt = f.shape
    ix0, iy0 = nx // 2, ny // 2
    E = energy(f, dx, dy, dt) * sI
    a2 = np.abs(f) ** 2
    P_t = np.sum(a2, axis=(0, 1)) * dx * dy * sI
    flu = np.sum(a2, axis=-1) * dt * sI
    xm, ym, wx, wy = beam_moments(flu, x, y, dx, dy)
    it = int(np.argmax(P_t))
    A_xy = f[:, :, it] * A_scale
    fw = np.fft.fftshift(np.fft.fft(f, axis=-1), axes=-1)
    X, Y = np.meshgrid(x, y, indexing="ij")
    mode = np.exp(-(X ** 2 + Y ** 2) / w0_mode ** 2)[:, :, None]
    spec_smf = np.abs(np.sum(fw * mode, axis=(0, 1))) ** 2 * sI
    spec_full = np.sum(np.abs(fw) ** 2, axis=(0, 1)) * sI
    spec_center = np.abs(fw[ix0, iy0, :]) ** 2 * sI
    f_abs = np.asarray(freq_axis_hz) + C0 / lambda0
    valid = f_abs > 1e9
    wl = np.zeros_like(f_abs)
    wl[valid] = C0 / f_abs[valid] * 1e9
    I_c = np.abs(f[ix0, iy0, :] * A_scale) ** 2
    I_smf = np.abs(np.sum(f * mode, axis=(0, 1)) * A_scale) ** 2
    return dict(energy_total=E, x_mean_m=xm, y_mean_m=ym, wx_rms_m=wx, wy_rms_m=wy,
                t_s=np.asarray(t), P_t=P_t, I_center_t=I_c, I_smf_t=I_smf, wl_nm=wl,
                freq_abs_Hz=f_abs, spec_smf=spec_smf, spec_full_incoh=spec_full,
                spec_center=spec_center, fluence=flu, it_peak=it, A_xy=A_xy,
                fwhm_center=fwhm(t, I_c), fwhm_smf=fwhm(t, I_smf))
