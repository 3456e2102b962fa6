"""The workgroup LDS FFT of the 1D kernels (LdsFFT in uppe1d.hip: Stockham stages of radix 8,
then 4 or 2, ping-ponging between two buffers) restated in NumPy index by index, with the
intermediate stages stored at fpad(n) = n + n / 32: it must equal the DFT for every supported
length and workgroup size, stay inside the padded buffers, and its 64-bit LDS accesses must
be at most two-way bank-conflicted for 16 lanes per cycle (eight-way without the padding)."""
import numpy as np
import pytest


def fpad(n):
    return n + (n >> 5)


def stages(N):
    out, NS = [], 1
    while NS < N:
        rem = N // NS
        R = 8 if rem >= 8 else (4 if rem == 4 else 2)
        out.append((NS, R))
        NS *= R
    return out


def degree(addr, group=16):
    addr = np.asarray(addr)
    return max(np.bincount(addr[g:g + group] % group, minlength=group).max()
               for g in range(0, len(addr) - group + 1, group))


def lds_fft(x, T, sign, padded=True):
    N = len(x)
    size = N + N // 32
    src = np.full(size, np.nan + 0j)
    dst = np.full(size, np.nan + 0j)
    src[:N] = x
    tw = np.exp(-2j * np.pi * np.arange(N) / N)
    st = stages(N)
    worst = 1
    for si, (NS, R) in enumerate(st):
        pin = fpad if (padded and si > 0) else (lambda n: n)
        pout = fpad if (padded and si < len(st) - 1) else (lambda n: n)
        NB, STEP = N // R, N // (NS * R)
        for j0 in range(0, NB, T):
            j = np.arange(j0, min(j0 + T, NB))
            idx = [pin(j + r * NB) for r in range(R)]
            assert max(i.max() for i in idx) < size
            v = np.stack([src[i] for i in idx], 1)
            k = j & (NS - 1)
            if NS > 1:
                w = tw[(k * STEP)[:, None] * np.arange(R)[None, :]]
                v = v * (w if sign > 0 else np.conj(w))
            v = np.fft.fft(v, axis=1) if sign > 0 else np.fft.ifft(v, axis=1) * R
            b = (j - k) * R + k
            for r in range(R):
                o = pout(b + r * NS)
                assert o.max() < size
                dst[o] = v[:, r]
                for w0 in range(0, len(j) - 63, 64):     # one wavefront's store
                    worst = max(worst, degree(o[w0:w0 + 64]))
            for i in idx:
                for w0 in range(0, len(j) - 63, 64):     # one wavefront's load
                    worst = max(worst, degree(i[w0:w0 + 64]))
        src, dst = dst, np.full(size, np.nan + 0j)
    return src[:N], worst


@pytest.mark.parametrize("N,T", [(256, 256), (1024, 256), (2048, 256), (2048, 1024),
                                 (4096, 512), (8192, 1024), (8192, 256)])
def test_padded_lds_fft(N, T):
    rng = np.random.default_rng(N + T)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    y, worst = lds_fft(x, T, 1)
    np.testing.assert_allclose(y, np.fft.fft(x), atol=1e-9 * np.sqrt(N))
    assert worst <= 2
    y, _ = lds_fft(x, T, -1)
    np.testing.assert_allclose(y / N, np.fft.ifft(x), atol=1e-12)
    if N // 8 >= 64:   # a full wavefront of radix-8 butterflies in the first stage
        _, worst_unpadded = lds_fft(x, T, 1, padded=False)
        assert worst_unpadded == 8
