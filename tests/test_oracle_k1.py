"""Pins the plasma-density recurrence (media_physics.py:11-66 restated in oracle/k1.py and
oracle/plasma_k1.c) by closed forms of the recurrence and by the two restatements agreeing."""
import numpy as np

from oracle import k1


def test_c_and_numpy_restatements_agree():
    """Same operations in the same order; they differ only by libm vs NumPy exp rounding,
    amplified where the depletion factor (rho_nt - rho)/rho_nt cancels."""
    rng = np.random.default_rng(0)
    w = (10 ** rng.uniform(-5, 16, (7, 300))).astype(np.float32)
    wa = (10 ** rng.uniform(-3, 14, (7, 300))).astype(np.float32)
    for args in [(w, np.zeros_like(w), 1e9, 0.4e-15, 2.45e25), (w, wa, 1e9, 0.4e-15, 2.45e25),
                 (w * 1e3, wa, 6.7e12, 1e-15, 2.1e28)]:
        r1, g1 = k1.k1(*args)
        r2, g2 = k1.k1_numpy(*args)
        np.testing.assert_allclose(r1, r2, rtol=1e-8, atol=0)
        np.testing.assert_allclose(g1, g2, rtol=1e-8, atol=0)


def test_constant_rate_without_recombination():
    """rate_rec = 0: rho_{t+1} = rho_t + w (rho_nt - rho_t) dt  =>  rho_t = rho_nt (1 - (1-w dt)^t)."""
    n, dt, rnt, wv = 500, 1e-15, 2.5e25, 3e11
    w = np.full(n, wv, np.float32)
    rho, gen = k1.k1(w, np.zeros_like(w), 0.0, dt, rnt)
    dtd, rntd, wd = float(np.float32(dt)), float(np.float32(rnt)), float(np.float32(wv))
    exact = rntd * (1 - (1 - wd * dtd) ** np.arange(n))
    np.testing.assert_allclose(rho, exact, rtol=1e-11)
    np.testing.assert_allclose(gen, wd * (rntd - rho), rtol=1e-11)


def test_recombination_only_decays_nothing_from_zero_and_clamps():
    n = 64
    rho, gen = k1.k1(np.zeros(n, np.float32), np.zeros(n, np.float32), 1e12, 1e-15, 1e25)
    assert np.all(rho == 0) and np.all(gen == 0)
    # a rate far above 1/dt saturates the density at rho_nt after one step (clamp)
    w = np.full(n, 1e20, np.float32)
    rho, _ = k1.k1(w, np.zeros_like(w), 1e9, 1e-15, 1e25)
    rnt = float(np.float32(1e25))
    assert rho[0] == 0 and rho[1] == rnt
    # saturated: depletion is 0, recombination pulls rho below rho_nt by 1 - exp(-r dt),
    # and the next step's source clamps it back
    assert np.all(rho[1:] >= rnt * (1 - 2e-6)) and np.all(rho[1:] <= rnt)


def test_exponential_branch_matches_exact_integrator():
    """Large |gdt| takes the exp branch: one step from 0 gives S (e^{G dt} - 1)/G."""
    w = np.array([1e10, 0], np.float32)
    wa = np.array([0, 0], np.float32)
    rr, dt, rnt = 1e12, 1e-11, 1e25          # gdt = -10 (clamp edge) -> exp branch
    rho, _ = k1.k1(w, wa, rr, dt, rnt)
    G = -float(np.float32(rr))
    dtd = float(np.float32(dt))
    S = float(np.float32(1e10)) * float(np.float32(rnt))
    np.testing.assert_allclose(rho[1], min(S * ((np.exp(G * dtd) - 1) / G), float(np.float32(rnt))),
                               rtol=1e-14)
