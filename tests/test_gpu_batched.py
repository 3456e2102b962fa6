"""Batched 1D sweep: every instance of one persistent launch agrees with the FP64 oracle run
on its own, for every launch plan the autotuner may pick, and a captured HIP graph replays to
the same result."""
import numpy as np
import pytest

from conftest import announce
from oracle import uppe1d as o1
from project_amd import workloads as wl
from project_amd.batched import BatchedSweep1D

pytestmark = pytest.mark.gpu


def oracle_for(p):
    p_atm = p["pressure_bar"] * 1e5 / 101325.0
    pl = o1.Plasma1D(K=8, sigma_K=1e-130 * p_atm, rho_nt=2.45e25 * p_atm, tau_c=190e-15 / p_atm)
    return o1.Oracle1D(o1.Params1D(material="argon_gas", n2=1e-23 * p["pressure_bar"],
                                   beam_radius=20e-6, center_wavelength=800e-9, use_raman=False,
                                   use_self_steepening=True, gas_P=p["pressure_bar"] * 1e5,
                                   gas_T=293.15, plasma=pl))


def test_batched_sweep_matches_oracle_and_graph_replay(gpu):
    params = wl.batch1d_params(n=10, seed=5, nts=(256, 512, 1024, 2048))
    inputs = wl.batch1d_inputs(params, seed=5)
    solvers = [wl.argon_solver_1d(p["pressure_bar"]) for p in params]
    L = 0.05
    announce("batched 1d: 10 instances, nt 256..2048, L=%g" % L)
    sw = BatchedSweep1D(solvers, inputs, L, tol=1e-6)
    ref = []
    for p, (t, a) in zip(params, inputs):
        z, Ah, w, log = oracle_for(p).propagate(a, t, L, tol=1e-6)
        ref.append((np.fft.fft(Ah[-1]), len(log)))
    for plan in sw.candidate_plans():
        st, secs = sw.run(plan)
        spec = sw.spectra()
        for k in range(len(params)):
            assert st[k].done == 1 and abs(st[k].z - L) < 1e-12
            e = np.linalg.norm(spec[k] - ref[k][0]) / np.linalg.norm(ref[k][0])
            assert e < 2e-5, (plan.groups, k, e)
            assert abs(st[k].attempts - ref[k][1]) <= 2
    best, secs = sw.autotune()
    st, _ = sw.run()
    eager = [s.copy() for s in sw.spectra()]
    g = sw.capture()
    for _ in range(3):
        sw.out.zero_()
        g.replay()
        import torch
        torch.cuda.synchronize()
        for k, s in enumerate(sw.spectra()):
            np.testing.assert_array_equal(s, eager[k])
