"""MPC runner diagnostics and mirror on the device (project_amd.diagnostics) against the FP64
oracle (oracle/diagnostics.py) on the same field.  Tolerances: relative 1e-5 on sums (the
device field is complex64, the sums are FP64), exact peak index, 1e-5 on the FWHM."""
import numpy as np
import pytest

from conftest import announce
from oracle import diagnostics as od
from project_amd import diagnostics as dg
from project_amd import media_physics as mp
from project_amd import workloads as wl
from project_amd.solver_3d import ERK43IP_UPPE_3D_Optimized

pytestmark = pytest.mark.gpu


def field(nx, nt, seed):
    x, y, t = wl.grid3d(nx, nt)
    A = wl.beam3d(x, y, t)
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(x, y, indexing="ij")
    # off-centre, astigmatic, chirped: every moment and spectrum is non-trivial
    A = A * np.exp(-((X - 3e-4) ** 2) / (1.6e-3) ** 2)[:, :, None]
    A = A * np.exp(1j * 2e26 * t ** 2)[None, None, :]
    A = A * (1 + 0.05 * rng.standard_normal(A.shape))
    return x, y, t, A.astype(np.complex64)


@pytest.mark.parametrize("nx,nt", [(32, 256), (16, 1024), (64, 128), (16, 2046), (32, 1001)])
def test_snapshot_vs_oracle(gpu, nx, nt):
    x, y, t, A = field(nx, nt, 7)
    s = ERK43IP_UPPE_3D_Optimized(x, y, t, mp.Argon(800e-9, 1.0, tunnel_guard=0.0))
    announce("diagnostics %dx%dx%d" % (nx, nx, nt))
    s.load(A)
    w0 = 1.1e-3
    g = dg.snapshot(s, w0)
    o = od.snapshot(A.astype(np.complex128) / s.A_scale, x, y, t, s.A_scale, w0, s.lambda0,
                    s.freq_axis_hz)
    for k in ["energy_total", "x_mean_m", "y_mean_m", "wx_rms_m", "wy_rms_m"]:
        assert g[k] == pytest.approx(o[k], rel=1e-5, abs=1e-12), k
    for k in ["P_t", "I_center_t", "I_smf_t", "spec_smf", "spec_full_incoh", "spec_center",
              "fluence"]:
        np.testing.assert_allclose(g[k], o[k], rtol=1e-5, atol=1e-6 * np.abs(o[k]).max(),
                                   err_msg=k)
    assert g["it_peak"] == o["it_peak"]
    np.testing.assert_allclose(g["A_xy"], o["A_xy"], rtol=1e-5, atol=1e-6 * np.abs(o["A_xy"]).max())
    np.testing.assert_allclose(g["wl_nm"], o["wl_nm"])
    assert g["fwhm_center"] == pytest.approx(o["fwhm_center"], rel=1e-5)
    assert g["fwhm_smf"] == pytest.approx(o["fwhm_smf"], rel=1e-5)


@pytest.mark.parametrize("nx,nt", [(32, 256), (16, 1001)])
def test_mirror_vs_oracle(gpu, nx, nt):
    x, y, t, A = field(nx, nt, 3)
    s = ERK43IP_UPPE_3D_Optimized(x, y, t, mp.Argon(800e-9, 1.0, tunnel_guard=0.0))
    announce("mirror %dx%dx%d" % (nx, nx, nt))
    s.load(A)
    # load (forward t transform) and store (inverse) round trip
    back = s.field().cpu().numpy()
    rt = np.linalg.norm(back - A) / np.linalg.norm(A)
    print("load/store round trip rel L2", rt)
    assert rt < 1e-6
    amp = np.sqrt(0.539 ** (1 / 33))
    dg.apply_mirror(s, 0.3, amp)
    got = s.field().cpu().numpy().astype(np.complex128)
    ref = od.mirror(A.astype(np.complex128), x, y, s.omega, s.omega0, 0.3) * amp
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("mirror rel L2", err)
    assert err < 2e-6
    dg.apply_mirror(s, 0.0, 2.0)   # flat, amplitude only
    np.testing.assert_allclose(s.field().cpu().numpy(), 2.0 * got, rtol=1e-5,
                               atol=1e-6 * np.abs(got).max())
