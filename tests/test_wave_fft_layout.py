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
