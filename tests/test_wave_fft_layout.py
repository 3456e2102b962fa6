"""The register-resident t-line FFT of the 3D t-pass (four-step NT = 64 x R, radix-2 stages
across lanes, bit-reversed lane positions) restated in NumPy index by index: it must equal the
DFT for every supported line length, in both directions."""
import numpy as np
import pytest


def brev6(l):
    return int("{:06b}".format(l)[::-1], 2)


def lanes(x, tw, nt, dif, conj):
    x = x.copy()
    ln = np.arange(64)
    ms = [32, 16, 8, 4, 2, 1] if dif else [1, 2, 4, 8, 16, 32]
    for m in ms:
        w = tw[(ln & (m - 1)) * (nt // (2 * m))]
        w = np.conj(w) if conj else w
        hi = ((ln & m) != 0)[:, None]
        if dif:
            p = x[ln ^ m]
            x = np.where(hi, (p - x) * w[:, None], x + p)
        else:
            tt = np.where(hi, x * w[:, None], x)
            p = tt[ln ^ m]
            x = np.where(hi, p - tt, tt + p)
    return x


@pytest.mark.parametrize("R", [2, 4, 8, 16, 32])
def test_wave_fft_index_algebra(R):
    nt = 64 * R
    tw = np.exp(-2j * np.pi * np.arange(nt) / nt)
    rl = np.array([brev6(l) for l in range(64)])
    k = np.arange(R)
    rng = np.random.default_rng(R)
    X = rng.standard_normal(nt) + 1j * rng.standard_normal(nt)
    v = X[rl[:, None] + 64 * k[None, :]]                 # lane lam holds X[l + 64 r]
    v = np.fft.ifft(v, axis=1) * R                       # R-point DFT (inverse sign)
    v = v * np.conj(tw[rl[:, None] * k[None, :]])
    v = lanes(v, tw, nt, dif=False, conj=True) / nt
    np.testing.assert_allclose(v.reshape(-1), np.fft.ifft(X), atol=1e-13)   # t = k + R lam
    y = rng.standard_normal(nt) + 1j * rng.standard_normal(nt)
    v = y.reshape(64, R)
    v = lanes(v, tw, nt, dif=True, conj=False)
    v = v * tw[rl[:, None] * k[None, :]]
    v = np.fft.fft(v, axis=1)
    Y = np.zeros(nt, complex)
    Y[rl[:, None] + 64 * k[None, :]] = v
    np.testing.assert_allclose(Y, np.fft.fft(y), atol=1e-10)


# -- NT = 1024 line transform through a wave-private LDS region (16 x 16 x 4) --------------
P1, S1, REGION = 68, 260, 16 * 68


def w1024_tables():
    k = np.arange(16)[:, None]
    lam = np.arange(64)[None, :]
    twA = np.exp(-2j * np.pi * (lam * k) / 1024)          # twA[k, lane]
    twB = np.exp(-2j * np.pi * ((lam & 3) * k) / 64)      # twB[k, lane]
    return twA, twB


def w1024_inverse(v, twA, twB):
    """v[lane, r] = X[lane + 64 r]  ->  v[lane, 4 j + b] = x[t1 + 16 (4 ug + j) + 256 b]."""
    lane = np.arange(64)
    v = np.fft.ifft(v, axis=1) * 16
    v = v * np.conj(twA.T)
    buf = np.full(REGION, np.nan + 0j)
    addr = np.arange(16)[None, :] * P1 + lane[:, None]       # [lane, k]
    assert len(np.unique(addr)) == addr.size and addr.max() < REGION
    buf[addr] = v
    t1p, la = lane >> 2, lane & 3
    addr = t1p[:, None] * P1 + la[:, None] + 4 * np.arange(16)[None, :]
    v = buf[addr]
    v = np.fft.ifft(v, axis=1) * 16
    v = v * np.conj(twB.T)
    buf = np.full(REGION, np.nan + 0j)
    addr = la[:, None] * S1 + np.arange(16)[None, :] * 16 + t1p[:, None]
    assert len(np.unique(addr)) == addr.size and addr.max() < REGION
    buf[addr] = v
    t1, ug = lane & 15, lane >> 4
    out = np.zeros((64, 16), complex)
    for j in range(4):
        w = buf[np.arange(4)[None, :] * S1 + ((4 * ug + j) * 16 + t1)[:, None]]
        w = np.fft.ifft(w, axis=1) * 4
        out[:, 4 * j:4 * j + 4] = w / 1024
    assert np.isfinite(out).all()
    return out


def w1024_forward(v, twA, twB):
    lane = np.arange(64)
    t1, ug = lane & 15, lane >> 4
    buf = np.full(REGION, np.nan + 0j)
    for j in range(4):
        w = np.fft.fft(v[:, 4 * j:4 * j + 4], axis=1)
        buf[np.arange(4)[None, :] * S1 + ((4 * ug + j) * 16 + t1)[:, None]] = w
    t1p, la = lane >> 2, lane & 3
    v = buf[la[:, None] * S1 + np.arange(16)[None, :] * 16 + t1p[:, None]]
    v = np.fft.fft(v * twB.T, axis=1)
    buf = np.full(REGION, np.nan + 0j)
    buf[t1p[:, None] * P1 + la[:, None] + 4 * np.arange(16)[None, :]] = v
    v = buf[np.arange(16)[None, :] * P1 + lane[:, None]]
    v = np.fft.fft(v * twA.T, axis=1)
    assert np.isfinite(v).all()
    return v


def w1024_time_index():
    lane = np.arange(64)
    t1, ug = lane & 15, lane >> 4
    j, b = np.arange(16) // 4, np.arange(16) % 4
    return t1[:, None] + 16 * (4 * ug[:, None] + j[None, :]) + 256 * b[None, :]


def test_w1024_index_algebra():
    twA, twB = w1024_tables()
    rng = np.random.default_rng(7)
    X = rng.standard_normal(1024) + 1j * rng.standard_normal(1024)
    v = X[np.arange(64)[:, None] + 64 * np.arange(16)[None, :]]
    x = w1024_inverse(v, twA, twB)
    tix = w1024_time_index()
    assert sorted(tix.reshape(-1)) == list(range(1024))
    np.testing.assert_allclose(x, np.fft.ifft(X)[tix], atol=1e-14)
    y = rng.standard_normal(1024) + 1j * rng.standard_normal(1024)
    V = w1024_forward(y[tix], twA, twB)
    np.testing.assert_allclose(V, np.fft.fft(y)[np.arange(64)[:, None] + 64 * np.arange(16)[None, :]],
                               atol=1e-10)
