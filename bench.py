#!/usr/bin/env python
"""Benchmark of the 3+1D envelope ERK4(3)-IP UPPE propagator on one MI355X.

Workload: argon at 1 bar (the reference Argon medium: Kerr n2, HOKE n4, LUT ionisation,
FP64 plasma recurrence, with the tunnel exponent evaluated without the reference's 1e-20
guard), Gaussian beam w0 = 1 mm, 40 fs, 800 nm, 0.8 mJ with seeded 1e-4 spatial noise, grid
128 x 128 x 1024 complex64, adaptive tolerance 1e-5 over L = 0.5 m.  A "step" is one
ERK4(3)-IP attempt (accepted or rejected) of the adaptive propagation: 5 nonlinear
evaluations, the L-infinity error norm and the on-device controller.  The field is resident in
HBM when the timed region starts.

Prints one JSON line: z-steps/s, per-kernel HBM roofline (HIP events around every launch of
the timed attempts), and the FP64 oracle timed on the host cores on a bounded sample.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# launch slots of one attempt (uppe3d_attempt_timed) and their algorithmic complex64-field
# transfers.  The 11-launch attempt (stage combinations in (kx, ky, omega), 31 field passes):
SLOTS = ["T1", "P2", "T2", "P3", "T3", "P4", "T4", "P5", "T5", "P6", "CTL"]
FIELD_PASSES = {"T1": 2, "P2": 5, "T2": 2, "P3": 5, "T3": 2, "P4": 5, "T4": 2, "P5": 3,
                "T5": 4, "P6": 1, "CTL": 0}
# on 128 x 128 x 1024 the final t-pass leaves Y5 (1 F read, 2 F written) and the error pass
# reads Y4 and Y5
FIELD_PASSES_SPLIT = dict(FIELD_PASSES, T5=3, P6=2)
GROUPS = {"tpass": ["T1", "T2", "T3", "T4"], "tpass_final": ["T5"], "xy_stage": ["P2", "P4"],
          "xy_stage_noexp": ["P3"], "xy_final": ["P5"], "xy_err": ["P6"], "ctl": ["CTL"]}
# the restructured attempt (stage fields in (x, y, omega), fused t-passes, 21 field passes):
# T1 reads B, writes Y1; XY-A reads B, Y1, writes X2, Zh; T23 reads X2, Zh, writes two
# increments; XY-B reads B and the increments, writes X4, R'; T45 reads X4, R', writes B', D;
# XY-E reads B', D
SLOTS_R = ["T1", "XYA", "T23", "XYB", "T45", "XYE", "CTL", "-", "-", "-", "-"]
FIELD_PASSES_R = {"T1": 2, "XYA": 4, "T23": 4, "XYB": 5, "T45": 4, "XYE": 2, "CTL": 0}
GROUPS_R = {"tpass": ["T1"], "tpass_fused": ["T23", "T45"], "xy_a": ["XYA"], "xy_b": ["XYB"],
            "xy_err": ["XYE"], "ctl": ["CTL"]}


def attempt_tables(path, nx, nt):
    """(slots, field passes per slot, kernel groups) of the attempt dataflow `path`."""
    if path == "restructured":
        return SLOTS_R, FIELD_PASSES_R, GROUPS_R
    return SLOTS, (FIELD_PASSES_SPLIT if (nx, nt) == (128, 1024) else FIELD_PASSES), GROUPS


def measured_traffic(nx, nt, kernel="tpass"):
    """HBM bytes per launch of one kernel group (FETCH_SIZE + WRITE_SIZE, calibrated against a
    copy kernel) from the committed rocprofv3 counter summary of this grid, or None."""
    path = os.path.join(ROOT, "profiles", "uppe3d_%dx%dx%d_hbm_traffic.md" % (nx, nx, nt))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("%s_bytes_per_launch =" % kernel):
            return int(line.split("=")[1])
    return None


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo), or the platform's processor string."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(nx, nt, attempts, workers):
    """FP64 oracle (NumPy/SciPy restatement of the reference) on the same workload."""
    from oracle import uppe3d as o3
    from project_amd import workloads as wl
    x, y, t = wl.grid3d(nx, nt)
    A0 = wl.beam3d(x, y, t)
    med = o3.GasMedium("Argon", 800e-9, 1e5 / 101325, tunnel_guard=0.0)
    orc = o3.Oracle3D(x, y, t, med, workers=workers)
    t0 = time.perf_counter()
    z, f, Aw, log = orc.propagate(A0, 0.5, tol=1e-5, max_attempts=attempts)
    dt = time.perf_counter() - t0
    return len(log) / dt, len(log), dt


def oracle_argon_1d(pressure_bar):
    """The FP64 oracle of the 1D argon model (project_amd.workloads.argon_solver_1d: K = 8
    plasma with sigma_8 = 1e-130 per atm, rho_nt and tau_c scaled with pressure)."""
    from oracle import uppe1d as o1
    p_atm = pressure_bar * 1e5 / 101325.0
    pl = o1.Plasma1D(K=8, sigma_K=1e-130 * p_atm, rho_nt=2.45e25 * p_atm, tau_c=190e-15 / p_atm)
    return o1.Oracle1D(o1.Params1D(material="argon_gas", n2=1e-23 * pressure_bar,
                                   beam_radius=20e-6, center_wavelength=800e-9, use_raman=False,
                                   use_self_steepening=True, gas_P=pressure_bar * 1e5,
                                   gas_T=293.15, plasma=pl))


def bench_config0(cpu=True):
    """1D config 0: N_t = 4096 argon, 1 uJ 30 fs 800 nm + 1e-6 noise, K = 8 plasma, 1 m,
    tol 1e-6: wall time of one full solve on the GPU and with the FP64 oracle (1 core)."""
    import torch
    from project_amd import workloads as wl
    t, a0 = wl.pulse1d(4096)
    s = wl.argon_solver_1d(1e5 / 1e5)
    s.propagate(a0, t, 1.0, tol=1e-6)            # warm-up (tables, code objects)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        z, Ah, w = s.propagate(a0, t, 1.0, tol=1e-6)
    torch.cuda.synchronize()
    gpu = (time.perf_counter() - t0) / reps
    st = s.last_status
    out = {"workload": "1D argon 1 bar, Nt 4096, 1 uJ 30 fs, K=8 plasma, L 1 m, tol 1e-6",
           "gpu_s_per_solve": round(gpu, 5), "attempts": st.attempts,
           "accepted": st.accepted, "z_steps_per_s": round(st.attempts / gpu, 1)}
    if cpu:
        orc = oracle_argon_1d(1.0)
        t0 = time.perf_counter()
        zo, Aho, wo, log = orc.propagate(a0, t, 1.0, tol=1e-6)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value_s_per_solve": round(cpu_s, 4), "cores": 1, "kind": "port",
                               "sample": "one full solve with the FP64 oracle"}
        out["rel_l2_vs_oracle"] = float(np.linalg.norm(Ah[-1] - Aho[-1]) / np.linalg.norm(Aho[-1]))
    return out


def batched_counters():
    """FP32-VALU and LDS figures of the batched sweep kernel from the committed rocprofv3
    counter summary, or None."""
    path = os.path.join(ROOT, "profiles", "uppe1d_batch256_counters.md")
    if not os.path.exists(path):
        return None
    out = {}
    for line in open(path):
        if line.startswith("bench:"):
            k, v = line[6:].split("=")
            out[k.strip()] = float(v)
    return out or None


def bench_batched(n=256, L=1.0, cpu=True):
    """Config 2: 256 instances, N_t ~ U{1024..8192}, E ~ logU[0.1, 10] uJ, GDD ~ U[+-500] fs^2,
    P ~ U[0.5, 3] bar, L 1 m, tol 1e-6, one persistent launch (autotuned), graph replay.
    One workgroup per instance, all resident at once, so the sweep cannot end before its
    longest instance: that instance (largest attempts x N_t) is also timed alone.  The CPU
    baseline is the FP64 oracle's time per attempt at each N_t (one full solve per N_t on one
    host core) applied to the attempts of every instance."""
    import torch
    from project_amd import workloads as wl
    from project_amd.batched import BatchedSweep1D
    params = wl.batch1d_params(n)
    inputs = wl.batch1d_inputs(params)
    solvers = [wl.argon_solver_1d(p["pressure_bar"]) for p in params]
    sw = BatchedSweep1D(solvers, inputs, L, tol=1e-6)
    best, tune_s = sw.autotune()
    g = sw.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.replay()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    st = sw.status()
    att = sum(x.attempts for x in st)
    nts = np.array([p["nt"] for p in params])
    done = sum(x.done for x in st)
    # streaming model: 18 complex64 field transfers of N_t points per attempt
    alg_bytes = sum(x.attempts * 18 * 8 * p["nt"] for x, p in zip(st, params))
    atts = np.array([x.attempts for x in st])
    crit = int(np.argmax(atts * nts))
    one = BatchedSweep1D([solvers[crit]], [inputs[crit]], L, tol=1e-6)
    one.run()
    st1, crit_s = one.run()
    out = {"workload": "batched 1D argon sweep, %d instances, Nt in {1024..8192}, L %g m, "
                       "tol 1e-6" % (n, L),
           "solves_per_s": round(n / secs, 2), "seconds": round(secs, 4),
           "z_steps_per_s": round(att / secs, 1), "attempts_total": att,
           "instances_done": done, "plan": [list(g_) for g_ in best.groups],
           "autotune": [[list(map(list, p)), round(t_, 4)] for (p, t_) in sw.tuning],
           "critical_path": {"instance": crit, "nt": int(nts[crit]),
                             "attempts": int(st1[0].attempts), "alone_s": round(crit_s, 4),
                             "sweep_over_alone": round(secs / crit_s, 3)},
           "hbm_model_GBps": round(alg_bytes / secs / 1e9, 1),
           "hbm_model_frac": round(alg_bytes / secs / 1e9 / HBM_PEAK_GBS, 4),
           "counters": batched_counters(),
           "nt_counts": {str(k): int((nts == k).sum()) for k in sorted(set(nts))}}
    if cpu:
        per_att, sample = {}, []
        for nt in sorted(set(nts)):
            k = int(np.nonzero(nts == nt)[0][0])
            t0 = time.perf_counter()
            o = oracle_argon_1d(params[k]["pressure_bar"])
            _, _, _, log = o.propagate(inputs[k][1], inputs[k][0], L, tol=1e-6)
            dt = time.perf_counter() - t0
            per_att[nt] = dt / len(log)
            sample.append("instance %d (Nt %d, %d attempts, %.2f s)" % (k, nt, len(log), dt))
        proj = float(sum(a * per_att[nt] for a, nt in zip(atts, nts)))
        out["cpu_baseline"] = {"value": round(n / proj, 4), "unit": "solves/s", "cores": 1,
                               "kind": "port", "projected_sweep_s": round(proj, 1),
                               "sample": "FP64 oracle, one full solve per Nt: " + "; ".join(sample)
                                         + "; per-attempt time x the GPU's attempts of every "
                                           "instance"}
        out["speedup_vs_cpu"] = round(proj / secs, 1)
    return out


def bench_cavity(passes=2, slices=20, cpu=True):
    """Multipass cavity of "gas copy.py" (argon 7 bar, 1030 nm, 160 uJ 275 fs, N_t 8192 on
    10 ps, L 0.2865 m, R 0.3 m, 20 slices per pass with the mode's gamma, tol 1e-5,
    max_step dz/5, min_step 1e-8), `passes` passes: the device-resident slice chain (one
    launch) against the reference's driver pattern (one propagate call per slice) on the same
    GPU, and the FP64 oracle slice loop on one host core."""
    import torch
    from project_amd import cavity, gas  # noqa: F401
    from project_amd.erk43ip_method import ERK43IP_UPPE
    kw = dict(material="argon_gas", n2=1e-23 * 7, beam_radius=1e-4, center_wavelength=1030e-9,
              use_raman=False, use_self_steepening=True, gas_P=7e5, gas_T=293.15)
    s = ERK43IP_UPPE(**kw)
    t = np.linspace(-5e-12, 5e-12, 8192)
    A0 = s.generate_gaussian_pulse(t, 160e-6, 275e-15)
    _, dz, _, gam = cavity.cavity_slices(s, 0.2865, 0.3, slices)
    dzs, gs = np.tile(dz, passes), np.tile(gam, passes)
    s.propagate_slices(A0, t, dzs[:2], gs[:2], tol=1e-5, max_step=dzs[:2] / 5, min_step=1e-8,
                       record="last")   # warm-up (tables, code objects)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A_out, _ = s.propagate_slices(A0, t, dzs, gs, tol=1e-5, max_step=dzs / 5, min_step=1e-8,
                                  record="last")
    chain_s = time.perf_counter() - t0
    att = sum(x.attempts for x in s.last_slice_status)
    t0 = time.perf_counter()
    A = A0
    for L, g in zip(dzs, gs):
        s.gamma = g
        _, Ah, _ = s.propagate(A, t, L, tol=1e-5, max_step=L / 5, min_step=1e-8)
        A = Ah[-1]
    loop_s = time.perf_counter() - t0
    out = {"workload": "multipass cavity (gas copy.py): argon 7 bar, 160 uJ 275 fs 1030 nm, "
                       "Nt 8192, %d passes x %d slices, tol 1e-5" % (passes, slices),
           "slices": int(len(dzs)), "attempts": int(att),
           "chain_s": round(chain_s, 4), "per_slice_calls_s": round(loop_s, 4),
           "chain_slices_per_s": round(len(dzs) / chain_s, 1),
           "chain_z_steps_per_s": round(att / chain_s, 1),
           "chain_vs_per_slice_calls": round(loop_s / chain_s, 2),
           "rel_l2_chain_vs_calls": float(np.linalg.norm(A_out[-1] - A) / np.linalg.norm(A))}
    if cpu:
        from oracle import uppe1d as o1
        orc = o1.Oracle1D(o1.Params1D(**kw))
        n_cpu = slices   # one pass on the host
        t0 = time.perf_counter()
        Ao = A0
        for L, g in zip(dzs[:n_cpu], gs[:n_cpu]):
            orc.gamma = g
            _, Aho, _, _ = orc.propagate(Ao, t, L, tol=1e-5, max_step=L / 5, min_step=1e-8)
            Ao = Aho[-1]
        cs = time.perf_counter() - t0
        out["cpu_baseline"] = {"value_s_per_slice": round(cs / n_cpu, 5), "cores": 1,
                               "kind": "port", "sample": "first pass (%d slices) with the FP64 "
                                                         "oracle" % n_cpu}
        out["rel_l2_vs_oracle_pass1"] = float(np.linalg.norm(A_out[n_cpu - 1] - Ao)
                                              / np.linalg.norm(Ao))
    return out


def bench_mpc(steps=60, nts=(2046, 2048)):
    """The MPC runner's fused-silica segment (src/3+1D-ERK4(3)IP UPPE/sim.py) at the runner's
    grid, 256 x 256 x N_t over 2 mm and 80 ps, the runner's 100 uJ 10.7 ps eigenmode pulse at
    1064 nm: `steps` steps of the 20 um maximum step (L2 error norm, tol 1e-5) on the device,
    and the time of one full cavity cycle projected from it.  N_t 2046 is the runner's own t
    axis (mixed-radix line transforms); 2048 the power-of-two one beside it."""
    import torch
    from project_amd import mpc
    out = {}
    for nt in nts:
        r = mpc.MPC_Final_Runner(output_dir="", verbose=False, Nt=nt)
        A = r.generate_initial_field()
        L = steps * r.d_fs / 600.0    # 600 steps of 20 um per 12 mm plate
        r.propagate_segment(A, 4 * 20e-6, "fs")   # warm-up (plan, tables, code objects)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r.propagate_segment(A, L, "fs")
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        st = r.solver_fs.last_status
        per = wall / max(st.attempts, 1)
        plates = 2 * r.n_plates * r.d_fs / 20e-6
        res = {"workload": "MPC runner fused-silica segment (sim.py), 256x256x%d, 1064 nm, "
                           "100 uJ 10.7 ps, max_step 20 um, L2 norm, tol 1e-5" % nt,
               "attempts": int(st.attempts), "accepted": int(st.accepted),
               "wall_s": round(wall, 4), "ms_per_attempt": round(1e3 * per, 3),
               "projected_cycle_s_plates": round(plates * per, 1)}
        if not out:
            out = res
        else:
            out["nt_%d" % nt] = res
        del r, A
        torch.cuda.empty_cache()
    return out


def bench_realfield(nx=128, nt=1024, attempts=10, cpu=True):
    """Real-field solver (src/3+1D-UPPE-ERK4(3)IP-real-field/): argon 1 atm, 800 nm, 30 fs,
    3e14 W/cm^2 peak, nx x nx x nt over 2 mm and 300 fs; `attempts` attempts of a fixed step."""
    import torch
    from project_amd import media_realfield as mr
    from project_amd.solver_3d_realfield import ERK43_RealField_UPPE_3D
    x = np.linspace(-1e-3, 1e-3, nx)
    t = np.linspace(-150e-15, 150e-15, nt)
    m = mr.Argon(800e-9)
    X, Y = np.meshgrid(x, x, indexing="ij")
    E0 = np.sqrt(3e18 / (0.5 * 8.85418781e-12 * 299792458.0 * m.n0))
    w = 2 * np.pi * 299792458.0 / 800e-9
    E = (E0 * np.exp(-(X ** 2 + Y ** 2) / (0.3e-3) ** 2)[:, :, None]
         * (np.exp(-2 * np.log(2) * t ** 2 / (30e-15) ** 2) * np.cos(w * t))).astype(np.float32)
    s = ERK43_RealField_UPPE_3D(x, x, t, m)
    s.load(E)
    s.attempt(1e-6)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(attempts):
        s.attempt(1e-6)
    wall = time.perf_counter() - t0
    out = {"workload": "real-field 3+1D, argon, 800 nm 30 fs 3e14 W/cm^2, %dx%dx%d" % (nx, nx, nt),
           "attempts": attempts, "ms_per_attempt": round(1e3 * wall / attempts, 3),
           "attempts_per_s": round(attempts / wall, 1)}
    if cpu:
        from oracle import realfield as orf
        o = orf.OracleRF(x, x, t, orf.RFMedium("Argon", 800e-9))
        o.E_scale = float(np.max(np.abs(E)))
        En = E.astype(np.float64) / o.E_scale
        t0 = time.perf_counter()
        o.step(En, 1e-6)
        cs = time.perf_counter() - t0
        out["cpu_baseline"] = {"value_s_per_attempt": round(cs, 3), "cores": 1, "kind": "port",
                               "sample": "one attempt of the same step with the FP64 oracle"}
        out["speedup_vs_cpu"] = round(cs / (wall / attempts), 1)
    return out


def bench_adjoint(sigma8=1e-121, L=0.2):
    """1D adjoint: gradient of ||S_out - S_target||^2 with respect to the input spectral phase
    for the config-0 pulse (N_t = 4096) with K = 8 plasma of cross section sigma8 per atm over
    L: one recorded forward solve and one backward pass, timed separately.  The defaults are the
    plasma case of the 1D adjoint GPU test."""
    import torch
    from project_amd import workloads as wl
    from project_amd.adjoint import (FrozenPropagator1D, UPPEPropagate1D, spectral_loss,
                                     super_gaussian_target)
    t, a0 = wl.pulse1d(4096)
    omega = 2 * np.pi * np.fft.fftfreq(4096, t[1] - t[0])
    base = np.fft.fft(a0)
    dev = torch.device("cuda")
    S_t = torch.from_numpy(super_gaussian_target(omega, base)).to(dev)
    prop = FrozenPropagator1D(wl.argon_solver_1d(1.0, sigma8_per_atm=sigma8), t, L, tol=1e-6)
    D0, U0, g0 = prop.tables()
    A0w = torch.from_numpy(base).to(dev).to(torch.complex64)
    prop.run(A0w, D0, U0, g0)          # adaptive run, freezes the step sequence
    fw, bw = [], []
    for _ in range(3):
        phi = torch.zeros(4096, dtype=torch.float64, device=dev, requires_grad=True)
        A = (torch.from_numpy(base).to(dev) * torch.exp(1j * phi)).to(torch.complex64)
        D = torch.from_numpy(D0).to(dev)
        U = torch.from_numpy(U0).to(dev)
        G = torch.tensor(g0, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = UPPEPropagate1D.apply(A, D, U, G, prop)
        loss = spectral_loss(out.to(torch.complex128), S_t)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fw.append(t1 - t0)
        bw.append(t2 - t1)
    f, b = min(fw), min(bw)
    return {"workload": "adjoint 1D, config-0 pulse (Nt 4096), K=8 plasma sigma8 %g, L %g m: "
                        "dL/d(phase[4096]), %d frozen steps, checkpoints every 16"
                        % (sigma8, L, prop.M),
            "forward_s": round(f, 5), "backward_s": round(b, 5),
            "adjoint_over_forward": round(b / f, 3)}


def bench_adjoint3d(nx=256, nt=2048, L=0.3):
    """3+1D adjoint: gradient of the on-axis spectral energy in 600-700 nm with respect to the
    input spatial phase (nx^2) and spectral phase (nt), argon (config 1 physics), 0.8 mJ,
    over L with sqrt(M) checkpoints.  Frozen forward and backward timed separately."""
    import torch
    from project_amd import adjoint3d as a3, media_physics as mp, workloads as wl
    from project_amd.solver_3d import ERK43IP_UPPE_3D_Optimized
    x, y, t = wl.grid3d(nx, nt)
    A_base = wl.beam3d(x, y, t)
    solver = ERK43IP_UPPE_3D_Optimized(x, y, t, mp.Argon(800e-9, 1e5 / 101325,
                                                         tunnel_guard=0.0))
    t0 = time.perf_counter()
    prop = a3.FrozenPropagator3D(solver, A_base, L, tol=1e-5)
    t_rec = time.perf_counter() - t0
    mask = a3.band_mask(solver.omega, solver.omega0, 600e-9, 700e-9)
    dev = torch.device("cuda")
    Ab = torch.from_numpy(A_base.astype(np.complex64)).to(dev)
    fw, bw = [], []
    for _ in range(2):
        phi = torch.zeros((nx, nx), dtype=torch.float32, device=dev, requires_grad=True)
        psi = torch.zeros(nt, dtype=torch.float32, device=dev, requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        B0 = a3.to_state(a3.phase_field(Ab, phi, psi), prop.A_scale)
        Bf = a3.UPPEPropagate3D.apply(B0, prop)
        loss = a3.onaxis_band_energy(Bf, prop.A_scale, mask, nx // 2, nx // 2)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fw.append(t1 - t0)
        bw.append(t2 - t1)
    f, b = min(fw), min(bw)
    gmax = float(max(phi.grad.abs().max().item(), psi.grad.abs().max().item()))
    return {"workload": "adjoint 3+1D %dx%dx%d argon 0.8 mJ, L %.2f m, tol 1e-5: d(on-axis "
                        "600-700 nm energy)/d(spatial phase, spectral phase); %d frozen "
                        "steps, checkpoints every %d" % (nx, nx, nt, L, prop.M, prop.C),
            "record_s": round(t_rec, 4), "forward_s": round(f, 4), "backward_s": round(b, 4),
            "adjoint_over_forward": round(b / f, 3), "loss": float(loss.item()),
            "grad_max_abs": gmax,
            "memory_GB": round(torch.cuda.max_memory_allocated() / 1e9, 2)}


def time_attempts(solver, A0, steps, warmup):
    """`steps` attempts of a loaded 3D solver from a fixed start (load, begin, `warmup`
    attempts through the production driver): first as the production driver launches them (no
    events inside the attempts, HIP events at the two ends), then again from the same start with
    an event between consecutive launch groups (per-kernel times, ms summed over the attempts)."""
    import torch
    from project_amd import _lib
    lib = _lib.load()
    s = _lib.stream_handle()
    stw = _lib.Uppe3DStatus()

    def start():
        solver.load(A0)
        _lib.check(lib.uppe3d_begin(solver._plan, 0.5, 1e-5, 1e-2, 1e-9, s), "uppe3d_begin")
        # warm-up through propagate's driver (uppe3d_run, one status check at its end), which
        # also sizes the plasma launch grid from what it saw
        _lib.check(lib.uppe3d_run(solver._plan, warmup, warmup, s, ctypes.byref(stw)),
                   "uppe3d_run")

    start()
    torch.cuda.synchronize()
    st0 = solver.status()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record()
    for _ in range(steps):
        _lib.check(lib.uppe3d_attempt(solver._plan, s), "uppe3d_attempt")
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    st1 = solver.status()
    if st1.done:
        raise RuntimeError("propagation finished inside the timed region; lower --steps")
    start()
    kms = (ctypes.c_float * 11)()
    tot = ctypes.c_float(0.0)
    _lib.check(lib.uppe3d_attempt_timed(solver._plan, steps, s, kms, ctypes.byref(tot)),
               "uppe3d_attempt_timed")
    torch.cuda.synchronize()
    return {"st0": st0, "st1": st1, "ms_plain": ev0.elapsed_time(ev1), "kms": list(kms),
            "ms_event": float(tot.value), "wall": wall, "path": solver.attempt_path()}


def bench_plasma3d(steps=50, warmup=5, nx=128, nt=1024):
    """The headline attempt with the plasma term active on the full grid: argon 1 bar (Kerr,
    HOKE, LUT ionisation, FP64 plasma recurrence), 4 mJ 40 fs on a 0.3 mm waist in a 2 mm
    window (peak ~6e17 W/m^2, the narrow-waist scan case of the 3D parity tests), tol 1e-5:
    z-steps/s, the t-pass line groups that took the plasma instantiation, per-kernel times."""
    from project_amd import media_physics as mp, workloads as wl
    from project_amd.solver_3d import ERK43IP_UPPE_3D_Optimized
    x, y, t = wl.grid3d(nx, nt, half_width=1e-3)
    A0 = wl.beam3d(x, y, t, energy=4e-3, w0=0.3e-3)
    med = mp.Argon(800e-9, pressure_atm=1e5 / 101325, tunnel_guard=0.0)
    solver = ERK43IP_UPPE_3D_Optimized(x, y, t, med)
    r = time_attempts(solver, A0, steps, warmup)
    st0, st1 = r["st0"], r["st1"]
    att = st1.attempts - st0.attempts
    slots, _, grp = attempt_tables(r["path"], nx, nt)
    # line groups a t-pass launch deals out (plasma deferral unit): 5 t-pass launches per
    # attempt of 16-line groups, or, restructured, T1 (16 lines) and two fused passes (8)
    if r["path"] == "restructured":
        groups = att * (nx * nx // 16 + 2 * nx * nx // (8 if nt >= 1024 else 16))
    else:
        groups = 5 * att * nx * nx // (16 if nt <= 1024 else 8)
    kms = r["kms"]
    per = {g: round(sum(kms[slots.index(k)] for k in sl) / (len(sl) * steps), 5)
           for g, sl in grp.items()}
    return {"workload": "3+1D argon 1 bar, Kerr+HOKE+LUT plasma (tunnel_guard 0), 4 mJ 40 fs "
                        "800 nm, w0 0.3 mm, window 2 mm, %dx%dx%d, tol 1e-5" % (nx, nx, nt),
            "steps": steps, "z_steps_per_s": round(1000.0 * steps / r["ms_plain"], 2),
            "ms_per_step": round(r["ms_plain"] / steps, 4),
            "z_start": st0.z, "z_end": st1.z, "attempts": att,
            "plasma_groups_frac": round((st1.plasma_groups - st0.plasma_groups) / groups, 4),
            "attempt_path": r["path"], "kernel_avg_ms": per}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=128)
    ap.add_argument("--nt", type=int, default=1024)
    ap.add_argument("--cpu-attempts", type=int, default=3,
                    help="oracle attempts timed for the CPU baseline (0 skips it)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the 1D, batched and adjoint legs")
    ap.add_argument("--only", choices=["plasma_3d", "mpc_3d", "realfield_3d", "cavity_1d",
                                       "batched_1d", "config0_1d", "adjoint_1d", "adjoint_3d"],
                    help="run one side leg alone and print its JSON (not the headline line)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before the run, copy a MIB-MiB float32 buffer 3 times (a plain copy "
                         "kernel of known traffic, to scale HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the 128x128x1024 field is not "
              "sharded; see the sharding design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        import torch
        src = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.only:   # one side leg alone (for profiling it); prints that leg's JSON
        legs = {"mpc_3d": bench_mpc, "realfield_3d": lambda: bench_realfield(cpu=False),
                "cavity_1d": lambda: bench_cavity(cpu=False),
                "batched_1d": lambda: bench_batched(cpu=False),
                "config0_1d": lambda: bench_config0(cpu=False), "adjoint_1d": bench_adjoint,
                "adjoint_3d": bench_adjoint3d, "plasma_3d": bench_plasma3d}
        print(json.dumps({args.only: legs[args.only]()}), flush=True)
        return

    import torch
    from project_amd import _lib, media_physics as mp, workloads as wl
    from project_amd.solver_3d import ERK43IP_UPPE_3D_Optimized

    lib = _lib.load()
    nx, nt = args.nx, args.nt
    x, y, t = wl.grid3d(nx, nt)
    A0 = wl.beam3d(x, y, t)
    med = mp.Argon(800e-9, pressure_atm=1e5 / 101325, tunnel_guard=0.0)
    solver = ERK43IP_UPPE_3D_Optimized(x, y, t, med)
    r = time_attempts(solver, A0, args.steps, args.warmup)
    st0, st1, ms_plain, kms, ms_event, wall = (r[k] for k in ("st0", "st1", "ms_plain", "kms",
                                                              "ms_event", "wall"))
    ms_per_step = ms_plain / args.steps
    steps_per_s = 1000.0 / ms_per_step

    field_bytes = 8.0 * nx * nx * nt
    slot_names, passes, groups = attempt_tables(r["path"], nx, nt)
    att_passes = sum(passes.values())
    kernels = {}
    for g, slots in groups.items():
        ms = sum(kms[slot_names.index(k)] for k in slots)
        launches = len(slots) * args.steps
        byts = passes[slots[0]] * field_bytes
        avg_ms = ms / launches
        kernels[g] = {"avg_ms": avg_ms, "share": ms / ms_event,
                      "bytes_per_launch": byts,
                      "GBps": (byts / (avg_ms * 1e-3) / 1e9) if byts else None}
    dom = max((g for g in kernels if kernels[g]["bytes_per_launch"] > 0),
              key=lambda g: kernels[g]["share"])
    d = kernels[dom]
    traffic = measured_traffic(nx, nt, dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(d["GBps"], 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(d["GBps"] / HBM_PEAK_GBS, 4),
                "traffic": traffic}
    if traffic is not None:
        roofline["traffic_source"] = ("profiles/uppe3d_%dx%dx%d_hbm_traffic.md (rocprofv3 "
                                      "FETCH_SIZE + WRITE_SIZE, copy-calibrated)" % (nx, nx, nt))
    step_gbs_37 = 37 * field_bytes * steps_per_s / 1e9
    step_gbs_alg = att_passes * field_bytes * steps_per_s / 1e9

    cpu = None
    if args.cpu_attempts > 0:
        workers = min(os.cpu_count() or 1, 16)
        v, n, secs = cpu_baseline(nx, nt, args.cpu_attempts, workers)
        v1, n1, secs1 = cpu_baseline(nx, nt, 1, 1)
        cpu = {"value": round(v, 5), "unit": "z-steps/s", "cores": workers, "kind": "port",
               "sample": "%d ERK4(3)-IP attempt(s) of the same workload from z=0 with the FP64 "
                         "oracle (NumPy/SciPy pocketfft, scipy workers=%d), %.1f s"
                         % (n, workers, secs),
               "cpu_model": cpu_model(),
               "one_core": {"value": round(v1, 5), "unit": "z-steps/s", "cores": 1,
                            "sample": "%d attempt, scipy workers=1, %.1f s" % (n1, secs1)}}

    out = {
        "metric": "3+1D UPPE z-steps/s (128^2x1024) + HBM GB/s",
        "value": round(steps_per_s, 2),
        "unit": "z-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (complex64; FP64 plasma recurrence and error norm)",
        "data": "synthetic (seeded Gaussian beam + 1e-4 spatial noise)",
        "config": {"workload": "3+1D envelope UPPE ERK4(3)-IP, argon 1 bar (Kerr+HOKE+LUT "
                               "plasma, tunnel exponent without the 1e-20 guard: "
                               "tunnel_guard=0), 0.8 mJ 40 fs 800 nm, w0 1 mm, tol 1e-5, L 0.5 m",
                   "grid": [nx, nx, nt], "z_start": st0.z, "z_end": st1.z,
                   "attempts": st1.attempts - st0.attempts,
                   "accepted": st1.accepted - st0.accepted,
                   "wall_s": round(wall, 4)},
        "roofline": roofline,
        "attempt_path": r["path"],
        "step_hbm_GBps": {"model_37F": round(step_gbs_37, 1),
                          "algorithmic": round(step_gbs_alg, 1),
                          "algorithmic_field_passes": att_passes,
                          "frac_algorithmic": round(step_gbs_alg / HBM_PEAK_GBS, 4)},
        "kernels": {g: {k: (round(v, 5) if isinstance(v, float) else v) for k, v in kd.items()}
                    for g, kd in kernels.items()},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(steps_per_s / cpu["value"], 1)
    if not args.no_extra:
        out["config0_1d"] = bench_config0(cpu=args.cpu_attempts > 0)
        out["batched_1d"] = bench_batched(cpu=args.cpu_attempts > 0)
        out["cavity_1d"] = bench_cavity(cpu=args.cpu_attempts > 0)
        for key, leg in (("plasma_3d", bench_plasma3d), ("mpc_3d", bench_mpc),
                         ("realfield_3d", lambda: bench_realfield(cpu=args.cpu_attempts > 0))):
            try:
                out[key] = leg()
            except Exception as e:   # a failing side leg is reported, not fatal to the line
                out[key] = {"error": "%s: %s" % (type(e).__name__, e)}
        # the 1D adjoint is timed on the plasma case of its GPU test; the config-3 case (1 m,
        # sigma8 1e-130) is kept off the GPU (tests/test_gpu_adjoint.py)
        out["adjoint_1d"] = bench_adjoint()
        out["adjoint_1d"]["config3"] = ("not timed: the 1 m, sigma8 1e-130 case is kept off "
                                        "the GPU after a fault with an experimental, since "
                                        "removed, 1D kernel build")
        try:
            out["adjoint_3d"] = bench_adjoint3d()
        except NotImplementedError as e:
            out["adjoint_3d"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
