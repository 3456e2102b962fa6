"""FP64 oracle solves for the GPU parity tests, run in a pool of fresh (spawned) processes
that import only NumPy/SciPy and the oracle, never the GPU stack.  Test infrastructure."""
import multiprocessing as mp
import os

import numpy as np


def sweep_oracle(p, precision="double"):
    """The 1D argon model of the batched sweep (workloads.argon_solver_1d) as an oracle;
    precision "single" is the same algorithm in complex64 arithmetic."""
    from oracle import uppe1d as o1
    p_atm = p["pressure_bar"] * 1e5 / 101325.0
    pl = o1.Plasma1D(K=8, sigma_K=1e-130 * p_atm, rho_nt=2.45e25 * p_atm, tau_c=190e-15 / p_atm)
    return o1.Oracle1D(o1.Params1D(material="argon_gas", n2=1e-23 * p["pressure_bar"],
                                   beam_radius=20e-6, center_wavelength=800e-9, use_raman=False,
                                   use_self_steepening=True, gas_P=p["pressure_bar"] * 1e5,
                                   gas_T=293.15, plasma=pl), precision=precision)


def sweep_solve_job(job):
    """One oracle solve of an instance of the batched sweep; kind:
    "free" (adaptive at tol), "free10" (adaptive at tol/10: the truncation error of "free"),
    "free32" (adaptive at tol, the same algorithm in complex64 arithmetic: how far rounding
    alone moves the attempt count), "replay" (the GPU's accepted steps) or "replay_c64_input"
    (the same steps from the input rounded to complex64: how much the problem amplifies one
    FP32 rounding of the state).  Returns (final spectrum, attempts)."""
    kind, p, t, a, L, tol, hs = job
    o = sweep_oracle(p, "single" if kind == "free32" else "double")
    if kind.startswith("free"):
        _, Ah, _, log = o.propagate(a, t, L, tol=tol / 10 if kind == "free10" else tol)
    else:
        if kind == "replay_c64_input":
            a = np.asarray(a, np.complex64).astype(np.complex128)
        _, Ah, _, log = o.propagate(a, t, L, replay=[(h, 0.0, 1) for h in hs])
    return np.fft.fft(Ah[-1]), len(log)


def run_pool(fn, jobs, workers=None):
    """fn over jobs in spawned worker processes (a worker that fails to start raises
    BrokenProcessPool instead of hanging); results in job order."""
    from concurrent.futures import ProcessPoolExecutor
    workers = workers or min(len(jobs), 16, os.cpu_count() or 1)
    with ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(fn, jobs, chunksize=1))
