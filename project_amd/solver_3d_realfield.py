"""Real-field (carrier-resolved) 3+1D UPPE solver on MI355X (drop-in for
``src/3+1D-UPPE-ERK4(3)IP-real-field/solver_3d.py:7-241``).

Host side (FP64 NumPy, once per solver): grids, the propagator exponent
D = -i (kz - k1 omega) with evanescent kz = -i sqrt|kz^2| (``solver_3d.py:54-89``), the P and J
coupling coefficients -i omega^2 / (2 eps0 c^2 kz) and -omega / (2 eps0 c^2 kz) with the
omega = 0 column zeroed (``:91-100``), the absorber (``:102-111``) and anti-alias filter
(``:48-52``), cast to the reference's float32 / complex64 tables (the full D and coefficient
arrays stay on the host for inspection; the device forms the same values per element from
FP64 tables of k(omega), k1 omega, omega, kx^2 and ky^2).  Device side (libuppe.so,
``uppe_rf3d_*``): every attempt of the ERK4(3)-IP step with its low-order error estimate and
the medium response (PPT table lookup, plasma density and Drude current in FP64, chi3 / chi5
polarisation).  The accept / reject controller is the reference's (``:189-223``, exponent
0.333, safety 0.9) on the host, one synchronisation per attempt, as in the reference.

Call contract kept: ``propagate(E0, L, tol=1e-4, max_step=1e-2, min_step=1e-9,
return_gpu=False) -> (z, freq_axis_hz, E_w)`` with E_w = rFFT_t(E(L)) of shape
(Nx, Ny, Nt/2 + 1) (not shifted).  ``return_gpu=True`` returns a torch tensor on the GPU.
Differences: a non-finite error raises ``FloatingPointError`` (the reference's loop would not
end); KeyboardInterrupt returns the partial field, as the reference; ``precision='double'``
(float64 / complex128 in the reference, ``:15-19``) warns and runs the float32 device path.
"""
import ctypes
import warnings

import numpy as np

from . import _lib


class ERK43_RealField_UPPE_3D:
    def __init__(self, x_grid, y_grid, t_grid, medium, precision="single"):
        if precision not in ("single", "double"):
            raise ValueError("precision must be 'single' or 'double'")
        if precision == "double":
            warnings.warn("ERK43_RealField_UPPE_3D: precision='double' is not implemented on the "
                          "device; the field runs in float32 / complex64 with FP64 plasma "
                          "density and current", RuntimeWarning, stacklevel=2)
        if getattr(medium, "use_raman", False):
            raise NotImplementedError("the Raman response is not implemented in the HIP kernels")
        self.medium = medium
        self.c = medium.c
        self.lambda0 = medium.lambda0
        self.omega0 = medium.omega0
        self.eps0 = 8.85418781e-12
        self.precision = precision
        self.float_dtype, self.complex_dtype = np.float32, np.complex64
        x = np.asarray(x_grid, np.float64)
        y = np.asarray(y_grid, np.float64)
        t = np.asarray(t_grid, np.float64)
        self.x, self.y, self.t = x, y, t
        self.Nx, self.Ny, self.Nt = len(x), len(y), len(t)
        self.dx, self.dy, self.dt = float(x[1] - x[0]), float(y[1] - y[0]), float(t[1] - t[0])
        kx = 2 * np.pi * np.fft.fftfreq(self.Nx, self.dx)
        ky = 2 * np.pi * np.fft.fftfreq(self.Ny, self.dy)
        self.kx2, self.ky2 = kx ** 2, ky ** 2
        self.K2_perp_64 = kx[:, None] ** 2 + ky[None, :] ** 2
        self.omega_64 = 2 * np.pi * np.fft.rfftfreq(self.Nt, self.dt)
        self.freq_axis_hz = self.omega_64 / (2 * np.pi)
        self._precompute_operators()
        self._init_absorber()
        self.E_scale = 1.0
        fmax = float(np.max(self.freq_axis_hz)) + 1e-10
        self.anti_alias_filter = np.exp(-(self.freq_axis_hz / (0.85 * fmax)) ** 16).astype(np.float32)
        import torch
        self.device = torch.device("cuda")
        self._plan = None
        self._ws = None
        self.step_log = None
        self.z = 0.0

    # -- host setup --------------------------------------------------------------------------
    def _precompute_operators(self):
        """solver_3d.py:54-100 (FP64, then complex64)."""
        w = self.omega_64
        c = float(self.c)
        ws = np.where(w == 0, 1e-12, w)
        k = np.asarray(self.medium.refractive_index(ws), np.float64) * ws / c
        w0 = self.omega0
        k0 = float(np.asarray(self.medium.refractive_index(np.array([w0])))[0]) * w0 / c
        dw = w0 * 1e-3
        kp = float(np.asarray(self.medium.refractive_index(np.array([w0 + dw])))[0]) * (w0 + dw) / c
        self.k1 = (kp - k0) / dw
        kref = self.k1 * w
        self.k_w, self.kref_w = k, kref
        kz2 = (k ** 2)[None, None, :] - self.K2_perp_64[:, :, None]
        kz = np.where(kz2 >= 0, np.sqrt(np.abs(kz2)).astype(np.complex128),
                      -1j * np.sqrt(np.abs(kz2)))
        self.D_operator = (-1j * (kz - kref[None, None, :])).astype(np.complex64)
        kzs = np.where(np.abs(kz) < 1e-10, 1e-10, kz)
        P = -1j * w[None, None, :] ** 2 / (2 * self.eps0 * c ** 2 * kzs)
        J = -w[None, None, :] / (2 * self.eps0 * c ** 2 * kzs)
        P[:, :, 0] = 0.0
        J[:, :, 0] = 0.0
        self.UPPE_P_coeff = P.astype(np.complex64)
        self.UPPE_J_coeff = J.astype(np.complex64)

    def _init_absorber(self):
        """solver_3d.py:102-111, on the float32 grids."""
        X, Y = np.meshgrid(self.x.astype(np.float32), self.y.astype(np.float32), indexing="ij")
        R = np.sqrt(X ** 2 + Y ** 2)
        hw = float(np.max(np.abs(self.x.astype(np.float32))))
        r_edge, r_start = 0.98 * hw, 0.85 * hw
        s = np.clip((R - r_start) / (r_edge - r_start + 1e-30), 0.0, 1.0)
        self.damp_space = (120.0 * s ** 10).astype(np.float32)

    def _desc(self):
        m = self.medium
        return _lib.UppeRf3DDesc(
            nx=self.Nx, ny=self.Ny, nt=self.Nt, lut_n=int(m.log10_I_f32.size), dt=self.dt,
            i_coef=0.5 * m.eps0 * m.c * m.n0, chi3_coef=m.eps0 * m.chi3,
            chi5_coef=m.eps0 * m.chi5 if m.chi5 != 0.0 else 0.0,
            av_coef=float(m.avalanche_coefficient()), rate_rec=1.0 / m.tau_rec,
            rho_nt=m.rho_nt, Ui=m.Ui, e_charge=m.e_charge, m_e=m.m_e, tau_c=m.tau_c,
            coupling=2 * self.eps0 * float(self.c) ** 2)

    def _make_plan(self):
        import torch
        lib = _lib.load()
        d = self._desc()
        nb = ctypes.c_size_t(0)
        _lib.check(lib.uppe_rf3d_workspace_size(ctypes.byref(d), ctypes.byref(nb)),
                   "uppe_rf3d_workspace_size")
        self._ws = torch.empty(nb.value, dtype=torch.uint8, device=self.device)
        tabs = [np.ascontiguousarray(a, np.float64) for a in (
            self.k_w, self.kref_w, self.omega_64, self.kx2, self.ky2)]
        tabs += [np.ascontiguousarray(a, np.float32) for a in (
            self.anti_alias_filter, self.damp_space, self.medium.log10_I_f32, self.medium.W_f32)]
        plan = ctypes.c_void_p()
        _lib.check(lib.uppe_rf3d_plan_create(
            ctypes.byref(d), _lib.ptr(self._ws), nb, *[_lib.host_ptr(a) for a in tabs],
            _lib.stream_handle(), ctypes.byref(plan)), "uppe_rf3d_plan_create")
        self._plan = plan
        self._desc_v = d

    def __del__(self):
        try:
            if self._plan is not None:
                _lib.load().uppe_rf3d_plan_destroy(self._plan)
        except Exception:
            pass

    # -- propagation -------------------------------------------------------------------------
    def load(self, E0):
        """Uploads the real field E0 (x, y, t) and normalises it by E_scale = max|E0|."""
        import torch
        lib = _lib.load()
        if self._plan is None:
            self._make_plan()
        if isinstance(E0, torch.Tensor):
            e = E0.to(device=self.device, dtype=torch.float32)
        else:
            e = torch.from_numpy(np.ascontiguousarray(E0, np.float32)).to(self.device)
        if tuple(e.shape) != (self.Nx, self.Ny, self.Nt):
            raise ValueError("E0 must have shape (Nx, Ny, Nt) = %s" % ((self.Nx, self.Ny, self.Nt),))
        e = e.contiguous()
        s = float(e.abs().max().item())
        self.E_scale = s if s >= 1e-30 else 1.0
        _lib.check(lib.uppe_rf3d_load(self._plan, _lib.ptr(e), self.E_scale, _lib.stream_handle()),
                   "uppe_rf3d_load")
        return e

    def attempt(self, h):
        err = ctypes.c_double()
        _lib.check(_lib.load().uppe_rf3d_attempt(self._plan, float(h), _lib.stream_handle(),
                                                 ctypes.byref(err)), "uppe_rf3d_attempt")
        return err.value

    def accept(self):
        _lib.check(_lib.load().uppe_rf3d_accept(self._plan), "uppe_rf3d_accept")

    def propagate(self, E0, L, tol=1e-4, max_step=1e-2, min_step=1e-9, return_gpu=False):
        self.load(E0)
        z = 0.0
        h = min(max_step, L / 1000.0)
        fail = 0
        log = []
        try:
            while z < L:
                if h > L - z:
                    h = L - z
                    if h < 1e-15:
                        break
                err = self.attempt(h)
                if not np.isfinite(err):
                    raise FloatingPointError("error norm became non-finite at z = %.9g" % z)
                if err < tol or h <= min_step * 1.01:
                    log.append((h, err, 1))
                    fail = fail + 1 if (h <= min_step * 1.01 and err > tol) else 0
                    z += h
                    self.accept()
                    e = max(err, 1e-30)
                    fac = 0.9 * (tol / e) ** 0.333
                    if fail > 0:
                        fac = min(fac, 1.2)
                    h = min(h * min(fac, 2.0), max_step)
                    h = max(h, min_step)
                else:
                    log.append((h, err, 0))
                    e = max(err, 1e-30)
                    h = max(h * 0.9 * (tol / e) ** 0.333, min_step)
        except KeyboardInterrupt:
            print("\n[Solver] Interrupted at z=%.3f mm" % (z * 1000))
        self.step_log = log
        self.z = z
        return self._finish(return_gpu)

    def propagate_replay(self, E0, steps, return_gpu=False):
        """Replays a step sequence [(h, err, accepted), ...] (e.g. an oracle log)."""
        self.load(E0)
        z, errs = 0.0, []
        for (h, _, acc) in steps:
            errs.append(self.attempt(h))
            if acc:
                self.accept()
                z += h
        self.replay_errors = errs
        self.z = z
        return self._finish(return_gpu)

    def _finish(self, return_gpu):
        import torch
        out = torch.empty((self.Nx, self.Ny, self.Nt // 2 + 1), dtype=torch.complex64,
                          device=self.device)
        _lib.check(_lib.load().uppe_rf3d_store_spectrum(self._plan, _lib.ptr(out),
                                                        _lib.stream_handle()),
                   "uppe_rf3d_store_spectrum")
        if return_gpu:
            return self.z, self.freq_axis_hz, out
        return self.z, self.freq_axis_hz, out.cpu().numpy()

    def field(self):
        """Current field E(x, y, t) (V/m) as a float32 torch tensor on the GPU."""
        import torch
        out = torch.empty((self.Nx, self.Ny, self.Nt), dtype=torch.float32, device=self.device)
        _lib.check(_lib.load().uppe_rf3d_store_field(self._plan, _lib.ptr(out),
                                                     _lib.stream_handle()),
                   "uppe_rf3d_store_field")
        return out
