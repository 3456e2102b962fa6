"""The mixed-radix t-line transform of the 3D kernels for t axes that are not a power of two
(gtpass / gload / gstore / gdiag in uppe3d.hip) restated in NumPy index by index: the radix
plan (make_gplan), the Stockham stages with their twiddle index k r N / (Ns R), the in-place
rule (every stage but the last reads all inputs before writing; the last writes exactly the
positions it read), the symmetric odd-radix DFT and the per-lane time chunks.  Each must give
the DFT for the supported lengths, in both directions; a dropped twiddle, a wrong output
base or a flipped sign in the odd-radix DFT fails it."""
import numpy as np
import pytest

BIG = (17, 19, 23, 29, 31)


def make_gplan(nt):
    """Radices ascending; powers of two in 16s (remainder 8, 4, 2); None if unsupported."""
    if not 128 <= nt <= 2048:
        return None
    m, twos, r, big = nt, 0, [], 0
    while m % 2 == 0:
        m //= 2
        twos += 1
    for p in range(3, 32, 2):
        while m % p == 0 and m > 1:
            m //= p
            r.append(p)
            big += p >= 17
    if m != 1 or big > 1:
        return None
    while twos >= 4:
        r.append(16)
        twos -= 4
    if twos:
        r.append(1 << twos)
    r.sort()
    return r if len(r) <= 12 else None


def dft_odd(v, R, sign):
    """The kernel's symmetric DFT: pair sums / differences, cos and sin of 2 pi q / R."""
    H = (R - 1) // 2
    s = v.copy()
    for m in range(1, H + 1):
        a, b = v[m], v[R - m]
        s[m], s[R - m] = a + b, a - b
    o = np.empty_like(v)
    o[0] = s[0] + sum(s[m] for m in range(1, H + 1))
    for k in range(1, H + 1):
        A = s[0] + sum(np.cos(2 * np.pi * ((m * k) % R) / R) * s[m] for m in range(1, H + 1))
        B = sum(np.sin(2 * np.pi * ((m * k) % R) / R) * s[R - m] for m in range(1, H + 1))
        o[k] = A - 1j * B if sign > 0 else A + 1j * B
        o[R - k] = A + 1j * B if sign > 0 else A - 1j * B
    return o


def dft_r(v, R, sign):
    if R % 2:
        return dft_odd(v, R, sign)
    return np.fft.fft(v) if sign > 0 else np.fft.ifft(v) * R


def transform(x, rads, sign):
    """Stockham stages on one line in place, one 64-lane wave; sign +1 forward, -1 inverse."""
    n = len(x)
    line = x.astype(complex).copy()
    tw = np.exp(-2j * np.pi * np.arange(n) / n)
    ns = 1
    for si, R in enumerate(rads):
        last = si == len(rads) - 1
        nb = n // R
        assert not last or ns * R == n
        reads, writes = [], {}
        if last:
            for j in range(nb):   # butterflies one at a time (lanes j, j + 64, ...)
                v = line[j + np.arange(R) * nb].copy()
                base = butterfly(v, j, ns, n, R, sign, tw)
                pos = base + np.arange(R) * ns
                assert set(pos) == set(j + np.arange(R) * nb)   # in place
                line[pos] = dft_out(v, j, ns, n, R, sign, tw)
        else:
            vals = [line[j + np.arange(R) * nb].copy() for j in range(nb)]   # all reads first
            for j in range(nb):
                base = butterfly(vals[j], j, ns, n, R, sign, tw)
                line[base + np.arange(R) * ns] = dft_out(vals[j], j, ns, n, R, sign, tw)
        ns *= R
    return line


def butterfly(v, j, ns, n, R, sign, tw):
    k = j % ns
    return (j // ns) * ns * R + k


def dft_out(v, j, ns, n, R, sign, tw):
    k = j % ns
    v = v.copy()
    if ns > 1:
        step = k * (n // (ns * R))
        w = tw[step * np.arange(R)]
        v = v * (w if sign > 0 else np.conj(w))
    return dft_r(v, R, sign)


@pytest.mark.parametrize("nt,expect", [(2046, [2, 3, 11, 31]), (1536, [2, 3, 16, 16]),
                                       (1001, [7, 11, 13]), (250, [2, 5, 5, 5]),
                                       (2000, [5, 5, 5, 16]), (729, [3, 3, 3, 3, 3, 3]),
                                       (1798, None), (2039, None), (4094, None), (126, None),
                                       (2048, [8, 16, 16])])
def test_radix_plan(nt, expect):
    r = make_gplan(nt)
    assert r == expect
    if r is not None:
        assert int(np.prod(r)) == nt
        assert all(p not in BIG for p in r[:-1])   # a large radix only in the last stage


@pytest.mark.parametrize("R", [3, 5, 7, 11, 13, 17, 19, 23, 29, 31])
def test_odd_radix_dft(R):
    rng = np.random.default_rng(R)
    v = rng.standard_normal(R) + 1j * rng.standard_normal(R)
    np.testing.assert_allclose(dft_odd(v, R, 1), np.fft.fft(v), atol=1e-12)
    np.testing.assert_allclose(dft_odd(v, R, -1), np.fft.ifft(v) * R, atol=1e-12)


@pytest.mark.parametrize("nt", [2046, 1536, 1001, 250, 2000, 729, 130, 1122])
def test_mixed_radix_line_transform(nt):
    rads = make_gplan(nt)
    assert rads is not None
    rng = np.random.default_rng(nt)
    x = rng.standard_normal(nt) + 1j * rng.standard_normal(nt)
    np.testing.assert_allclose(transform(x, rads, 1), np.fft.fft(x), atol=1e-9 * np.sqrt(nt))
    np.testing.assert_allclose(transform(x, rads, -1) / nt, np.fft.ifft(x), atol=1e-12)


@pytest.mark.parametrize("nt", [2046, 1001, 130])
def test_time_chunks_cover_the_line_in_order(nt):
    """Lane l holds t = l c + k, k < cnt_l, c = ceil(nt / 64): every sample once, and time
    order is lane-major (what the plasma scan across lanes relies on)."""
    c = (nt + 63) // 64
    seen = []
    for lane in range(64):
        t0 = lane * c
        cnt = max(0, min(c, nt - t0))
        seen += [t0 + k for k in range(cnt)]
    assert seen == list(range(nt))
    assert c <= 32
