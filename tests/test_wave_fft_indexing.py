"""CPU check of the index arithmetic of the one-wavefront-per-ring FFT
(gshtrans_amd/csrc/gsh_fft_wave.hip, wave_fft and the real-pair split in
fft_anal_w), re-traced lane by lane in numpy:

* the four-step decomposition N = R * 8 * 8 with the same register, lane,
  twiddle and bit-reversal bookkeeping reproduces numpy.fft (both signs);
* every LDS address of the two transposes stays inside the 1024-entry
  per-wave buffer, is written by exactly one lane, is read only after it was
  written in the same round, and every 16-lane phase of a 16-byte access
  touches 16 distinct 16-byte bank groups (no bank conflicts on 64 banks);
* the real-pair unpacking X_a = (Z[k] + conj Z[N-k]) / 2,
  X_b = (Z[k] - conj Z[N-k]) / 2i with the partner lane / register used by
  the kernel reproduces the FFTs of the two real rings.
"""
import numpy as np
import pytest


def bitrev(x, bits):
    r = 0
    for i in range(bits):
        r |= ((x >> i) & 1) << (bits - 1 - i)
    return r


def dft_reg(v, sign):
    """Radix-2 DIF over a list of per-lane registers, as dft_reg<R, SIGN>."""
    R = len(v)
    h = R // 2
    while h >= 1:
        for base in range(0, R, 2 * h):
            for j in range(h):
                a, b = v[base + j], v[base + j + h]
                v[base + j] = a + b
                v[base + j + h] = (a - b) * np.exp(1j * sign * np.pi * (16 * j // h) / 16)
        h //= 2


def check_phases(addr):
    for ph in range(4):
        s = addr[16 * ph:16 * ph + 16] % 16
        assert len(set(s.tolist())) == 16


def wave_fft(x, R, sign):
    N = 64 * R
    lane = np.arange(64)
    G, P, H = R // 8, 64 // R, (2 if R >= 32 else 1)
    LR = int(np.log2(R))
    v = [x[64 * n1 + lane].astype(complex) for n1 in range(R)]
    dft_reg(v, sign)
    twa = np.exp(-2j * np.pi * np.outer(np.arange(R), np.arange(64)) / N)
    for k1 in range(1, R):
        w = twa[k1] if sign < 0 else np.conj(twa[k1])
        v[bitrev(k1, LR)] = v[bitrev(k1, LR)] * w
    lds = np.full(1024, np.nan + 0j)
    b, cp = lane & 7, lane >> 3
    u = [[None] * 8 for _ in range(G)]
    for h in range(H):
        lds[:] = np.nan
        for kl in range(R // H):
            k1 = h * (R // H) + kl
            addr = kl * 64 + (lane ^ (8 * (k1 & 1)))
            assert addr.max() < 1024 and len(set(addr.tolist())) == 64
            check_phases(addr)
            lds[addr] = v[bitrev(k1, LR)]
        for gl in range(G // H):
            for a in range(8):
                addr = (8 * gl + cp) * 64 + ((b + 8 * a) ^ (8 * (cp & 1)))
                check_phases(addr)
                val = lds[addr]
                assert not np.isnan(val).any()
                u[h * (G // H) + gl][a] = val.copy()
    for g in range(G):
        dft_reg(u[g], sign)
    tw = np.exp(-2j * np.pi * np.arange(N) / N)
    for c in range(1, 8):
        w = tw[b * c * R] if sign < 0 else np.conj(tw[b * c * R])
        for g in range(G):
            u[g][bitrev(c, 3)] = u[g][bitrev(c, 3)] * w
    k1o, cm = lane % R, lane // R
    o = [[None] * 8 for _ in range(G)]
    for h in range(H):
        lds[:] = np.nan
        for g in range(G):
            for cl in range(8 // H):
                c = h * (8 // H) + cl
                k1 = 8 * g + cp
                sw = ((k1 >> 1) ^ ((c & 1) << 2)) & 7
                addr = (cl * R + k1) * 8 + (b ^ sw)
                assert addr.max() < 1024 and len(set(addr.tolist())) == 64
                check_phases(addr)
                lds[addr] = u[g][bitrev(c, 3)]
        for cql in range(G // H):
            cq = h * (G // H) + cql
            c = cm + P * cq
            cl = c - h * (8 // H)
            sw = ((k1o >> 1) ^ ((c & 1) << 2)) & 7
            for bb in range(8):
                addr = (cl * R + k1o) * 8 + (bb ^ sw)
                check_phases(addr)
                val = lds[addr]
                assert not np.isnan(val).any()
                o[cq][bb] = val.copy()
    for cq in range(G):
        dft_reg(o[cq], sign)
    out = np.zeros(N, complex)
    for cq in range(G):
        for d in range(8):
            out[lane + 64 * (cq + G * d)] = o[cq][bitrev(d, 3)]
    return out


@pytest.mark.parametrize("R", [8, 16, 32])
@pytest.mark.parametrize("sign", [-1, 1])
def test_wave_fft_matches_numpy(R, sign):
    N = 64 * R
    rng = np.random.default_rng(R)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    y = wave_fft(x, R, sign)
    ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * N
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("R", [8, 16, 32])
def test_real_pair_split(R):
    N = 64 * R
    rng = np.random.default_rng(100 + R)
    xa, xb = rng.standard_normal(N), rng.standard_normal(N)
    Z = wave_fft(xa + 1j * xb, R, -1)
    Fa, Fb = np.fft.fft(xa), np.fft.fft(xb)
    lane = np.arange(64)
    for t in range(R // 2 + 1):
        q = Z[((64 - lane) & 63) + 64 * (R - 1 - t)].copy()   # ds_bpermute source
        q[0] = Z[64 * ((R - t) % R)]                          # lane 0: own register
        k = lane + 64 * t
        m = k <= N // 2
        xa_k = (Z[k] + np.conj(q)) / 2
        xb_k = (Z[k] - np.conj(q)) / 2j
        assert np.max(np.abs(xa_k[m] - Fa[k[m]])) < 1e-12 * np.sqrt(N)
        assert np.max(np.abs(xb_k[m] - Fb[k[m]])) < 1e-12 * np.sqrt(N)
