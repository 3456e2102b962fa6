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


def test_w2048_two_w1024_and_radix2():
    """W2048 (uppe3d.hip): v[lane, r] = X[lane + 64 r], r < 32.  Inverse: e = X[k] + X[k +
    1024], o = (X[k] - X[k + 1024]) conj(W2048^k), two W1024 inverses, / 2: v[:, :16] hold the
    even samples and v[:, 16:] the odd ones, each in the W1024 time layout.  Forward: the
    transposed sequence, X[k] = E + W2048^k O, X[k + 1024] = E - W2048^k O."""
    twA, twB = w1024_tables()
    lane = np.arange(64)[:, None]
    r = np.arange(16)[None, :]
    twC = np.exp(-2j * np.pi * (lane + 64 * r) / 2048)     # twC[r * 64 + lane] as [lane, r]
    rng = np.random.default_rng(13)
    X = rng.standard_normal(2048) + 1j * rng.standard_normal(2048)
    v = X[np.arange(64)[:, None] + 64 * np.arange(32)[None, :]]
    a, b = v[:, :16], v[:, 16:]
    e = w1024_inverse(a + b, twA, twB) / 2
    o = w1024_inverse((a - b) * np.conj(twC), twA, twB) / 2
    tix = w1024_time_index()
    x = np.fft.ifft(X)
    np.testing.assert_allclose(e, x[2 * tix], atol=1e-14)
    np.testing.assert_allclose(o, x[2 * tix + 1], atol=1e-14)
    y = rng.standard_normal(2048) + 1j * rng.standard_normal(2048)
    E = w1024_forward(y[2 * tix], twA, twB)
    O = w1024_forward(y[2 * tix + 1], twA, twB) * twC
    V = np.concatenate([E + O, E - O], axis=1)
    np.testing.assert_allclose(V, np.fft.fft(y)[np.arange(64)[:, None] + 64 * np.arange(32)[None, :]],
                               atol=1e-10)


def test_r256_row_algebra():
    """r256 (uppe3d.hip): four 256-point rows per wavefront, lane = 16 s + j, value r of row s
    at y = j + 16 r.  Forward: DFT16 over r, x W256^(j k1), write wb[k1 * 65 + lane], read
    wb[k1 * 65 + 16 s + j'] (lane 16 s + k1), DFT16 over j' -> v[k2] = X[k1 + 16 k2].  Inverse:
    DFT16^-1 over k2, write wb[j * 65 + 16 s + k1], read wb[k1 * 65 + 16 s + k] (lane 16 s + j),
    x conj(W256^(j k)), DFT16^-1 -> 256 x[j + 16 r]."""
    PK256, WR = 65, 16 * 65
    lane = np.arange(64)
    s16, j = lane & 48, lane & 15
    tw = np.exp(-2j * np.pi * (j[:, None] * np.arange(16)[None, :]) / 256)   # [lane, k]
    rng = np.random.default_rng(17)
    rows = rng.standard_normal((4, 256)) + 1j * rng.standard_normal((4, 256))
    v = rows[(lane >> 4)[:, None], j[:, None] + 16 * np.arange(16)[None, :]]
    v = np.fft.fft(v, axis=1) * tw
    wb = np.full(WR, np.nan + 0j)
    addr = np.arange(16)[None, :] * PK256 + lane[:, None]
    assert len(np.unique(addr)) == addr.size and addr.max() < WR
    wb[addr] = v
    v = np.fft.fft(wb[j[:, None] * PK256 + s16[:, None] + np.arange(16)[None, :]], axis=1)
    K = np.fft.fft(rows, axis=1)
    np.testing.assert_allclose(v, K[(lane >> 4)[:, None], j[:, None] + 16 * np.arange(16)[None, :]],
                               atol=1e-10)
    u = np.fft.ifft(v, axis=1) * 16
    wb = np.full(WR, np.nan + 0j)
    addr = np.arange(16)[None, :] * PK256 + s16[:, None] + j[:, None]
    assert len(np.unique(addr)) == addr.size and addr.max() < WR
    wb[addr] = u
    w = wb[j[:, None] * PK256 + s16[:, None] + np.arange(16)[None, :]]
    w = np.fft.ifft(w * np.conj(tw), axis=1) * 16
    np.testing.assert_allclose(w / 256, rows[(lane >> 4)[:, None], j[:, None] + 16 * np.arange(16)[None, :]],
                               atol=1e-12)


# -- 128 x 128 plane transform: 16 x 8 register DFTs per 8 lanes, wave and plane transposes --
# LDS layouts of the device code: wave region (k, l) at 64 k + (k >> 1) + l, plane (row, col) at
# 133 row + 3 (row & 1) + col
WREG, PP = 1032, 133
PLANE = PP * 127 + 3 + 128


def widx(k, l):
    return 64 * k + (k >> 1) + l


def pidx(row, col):
    return PP * row + 3 * (row & 1) + col


def tw128():
    k = np.arange(16)[:, None]
    lane = np.arange(64)[None, :]
    return np.exp(-2j * np.pi * ((lane & 7) * k) / 128)    # [k, lane]


def fwd128(v, tw):
    """v[lane, r] = a_(lane>>3)[(lane & 7) + 8 r] -> v[lane, 8 h + k2] = A_(lane>>3)[k1 + 16 k2],
    k1 = 2 (lane & 7) + h."""
    lane = np.arange(64)
    v = np.fft.fft(v, axis=1) * tw.T
    wb = np.full(WREG, np.nan + 0j)
    addr = widx(np.arange(16)[None, :], lane[:, None])
    assert len(np.unique(addr)) == addr.size and addr.max() < WREG
    wb[addr] = v
    sq, jj = lane >> 3, lane & 7
    out = np.zeros((64, 16), complex)
    for h in range(2):
        u = wb[widx((2 * jj + h)[:, None], 8 * sq[:, None] + np.arange(8)[None, :])]
        out[:, 8 * h:8 * h + 8] = np.fft.fft(u, axis=1)
    assert np.isfinite(out).all()
    return out


def inv128(v, tw):
    lane = np.arange(64)
    sq, jj = lane >> 3, lane & 7
    wb = np.full(WREG, np.nan + 0j)
    for h in range(2):
        u = np.fft.ifft(v[:, 8 * h:8 * h + 8], axis=1) * 8
        wb[widx((2 * jj + h)[:, None], 8 * sq[:, None] + np.arange(8)[None, :])] = u
    v = wb[widx(np.arange(16)[None, :], lane[:, None])]
    v = np.fft.ifft(v * np.conj(tw.T), axis=1) * 16
    assert np.isfinite(v).all()
    return v


def test_x128_plane_algebra():
    tw = tw128()
    rng = np.random.default_rng(11)
    A = rng.standard_normal((128, 128)) + 1j * rng.standard_normal((128, 128))
    lane = np.arange(64)
    sq, jj = lane >> 3, lane & 7
    P = np.full(PLANE, np.nan + 0j)
    regs = []
    for wv in range(16):
        row = 8 * wv + sq
        v = A[row[:, None], jj[:, None] + 8 * np.arange(16)[None, :]]
        v = fwd128(v, tw)
        for h in range(2):
            for k2 in range(8):
                addr = pidx(2 * jj + h + 16 * k2, 8 * wv + sq)
                P[addr] = v[:, 8 * h + k2]
    assert np.isfinite(P[[pidx(ky, x) for ky in range(128) for x in range(128)]]).all()
    K = np.zeros((128, 128), complex)          # K[kx, ky]
    for wv in range(16):
        v = P[pidx((8 * wv + sq)[:, None], jj[:, None] + 8 * np.arange(16)[None, :])]
        v = fwd128(v, tw)
        regs.append(v)
        ky = 8 * wv + sq
        for n in range(16):
            kx = 2 * jj + (n >> 3) + 16 * (n & 7)
            K[kx, ky] = v[:, n]
    np.testing.assert_allclose(K, np.fft.fft2(A), atol=1e-9)
    # inverse back to rows
    P = np.full(PLANE, np.nan + 0j)
    for wv in range(16):
        v = inv128(regs[wv], tw)
        P[pidx((8 * wv + sq)[:, None], jj[:, None] + 8 * np.arange(16)[None, :])] = v
    out = np.zeros((128, 128), complex)
    for wv in range(16):
        v = np.zeros((64, 16), complex)
        for h in range(2):
            for k2 in range(8):
                v[:, 8 * h + k2] = P[pidx(2 * jj + h + 16 * k2, 8 * wv + sq)]
        v = inv128(v, tw) / 16384
        out[(8 * wv + sq)[:, None], jj[:, None] + 8 * np.arange(16)[None, :]] = v
    np.testing.assert_allclose(out, A, atol=1e-12)
    assert max(16 * WREG, PLANE) * 8 <= 160 * 1024


def bank_degree(addr, group):
    """Largest number of lanes of one group of `group` consecutive lanes whose 64-bit accesses
    fall on the same pair of 4-byte banks (2 group banks)."""
    addr = np.asarray(addr)
    return max(np.bincount(addr[g:g + group] % group, minlength=group).max()
               for g in range(0, 64, group))


def test_x128_transposes_conflict_free():
    """Every wavefront access of the four transposes is conflict-free with 16 lanes of 64-bit
    accesses per cycle and at most two-way with 32 (the previous layout, pitch 65 and 129, was
    two-way and up to four-way); offsets from the per-lane base are the same for every lane."""
    lane = np.arange(64)
    sq, jj = lane >> 3, lane & 7
    pats = []
    pats += [widx(k, lane) for k in range(16)]
    pats += [widx(2 * jj + h, 8 * sq + j) for h in range(2) for j in range(8)]
    pats += [pidx(2 * jj + h + 16 * k2, 8 * wv + sq) for h in range(2) for k2 in range(8)
             for wv in range(16)]
    pats += [pidx(8 * wv + sq, jj + 8 * r) for r in range(16) for wv in range(16)]
    assert max(bank_degree(a, 16) for a in pats) == 1
    assert max(bank_degree(a, 32) for a in pats) <= 2
    # constant offsets: the device adds (h + 16 k2) PP + 3 h to one per-lane base
    for h in range(2):
        for k2 in range(8):
            np.testing.assert_array_equal(pidx(2 * jj + h + 16 * k2, 8 * sq),
                                          pidx(2 * jj, 8 * sq) + (h + 16 * k2) * PP + 3 * h)


def cis_f32(x):
    """Float32 restatement of the device exp(i x) (FMA Cody-Waite reduction by pi/2, Taylor
    polynomials, quadrant rotation); FMAs are emulated in float64 then rounded."""
    f = np.float32
    x = np.asarray(x, np.float32)
    k = np.rint(x * f(0.636619772367581343)).astype(np.float32)
    r = (-k.astype(np.float64) * np.float64(f(1.57079637050628662)) + x).astype(np.float32)
    r = (-k.astype(np.float64) * np.float64(f(-4.37113900018624283e-8)) + r).astype(np.float32)
    r2 = r * r

    def fma(a, b, c):
        return (np.float64(a) * np.float64(b) + np.float64(c)).astype(np.float32)
    sp = fma(r2, f(1 / 362880), f(-1 / 5040))
    sp = fma(r2, sp, f(1 / 120))
    sp = fma(r2, sp, f(-1 / 6))
    sn = fma(r * r2, sp, r)
    cp = fma(r2, f(-1 / 3628800), f(1 / 40320))
    cp = fma(r2, cp, f(-1 / 720))
    cp = fma(r2, cp, f(1 / 24))
    cp = fma(r2, cp, f(-0.5))
    cs = fma(r2, cp, f(1))
    q = k.astype(np.int64) & 3
    c = np.where(q & 1, sn, cs)
    s = np.where(q & 1, cs, sn)
    return np.where((q == 1) | (q == 2), -c, c) + 1j * np.where(q >= 2, -s, s)


def test_cis_f32_accuracy():
    rng = np.random.default_rng(5)
    for scale in (1e-3, 1.0, 30.0, 3e3, 1e5):
        x = (rng.uniform(-1, 1, 20000) * scale).astype(np.float32)
        got = cis_f32(x)
        ref = np.exp(1j * x.astype(np.float64))
        err = np.abs(got - ref)
        # a few float32 ulps of the result plus the reduction error k * 2e-15
        assert err.max() < 4e-7 + 2e-15 * scale, (scale, err.max())
