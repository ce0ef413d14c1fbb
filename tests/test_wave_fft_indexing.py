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
  the kernel reproduces the FFTs of the two real rings;
* the exchange of order bins between the 4 or 8 rings of a workgroup
  through LDS (fft_synth_w gather, fft_anal_w scatter, with the ring counts
  of synth_waves / anal_waves): every LDS entry of a round is written
  exactly once and stays inside the workgroup buffer, every 16-lane phase
  is free of bank conflicts, each ring receives exactly the bins the
  per-ring formulas give (complex rings, real pairs, lmax below and at N/2,
  rings without a field), and every wave instruction of the cooperative
  global access covers whole 128-byte lines of F / FA (the workgroup's half
  of each line for 4 complex rings).
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



# ---------------------------------------------------------------------------
# Bin exchange between the rings of a workgroup (fft_synth_w / fft_anal_w).
# ---------------------------------------------------------------------------
def lds_per_wave(R):
    return (R // 2 if R >= 32 else R) * 64


def synth_waves(R):
    """Rings per synthesis workgroup (synth_waves in gsh_fft_wave.hip)."""
    return 4 if R >= 32 else 8


def anal_waves(R, real):
    """Rings per analysis workgroup (anal_waves in gsh_fft_wave.hip)."""
    return 4 if R >= 32 and real else 8


def xch(c, b, W):
    return b ^ ((16 // W) * c)


def check_b128_phases(addr16):
    """addr16: the 64 lane addresses, in 16-byte units, of one b128 access."""
    for ph in range(4):
        s = np.asarray(addr16[16 * ph:16 * ph + 16]) % 16
        assert len(set(s.tolist())) == 16


def check_whole_lines(lanes, line, piece, npieces=8):
    """Global accesses of one wave instruction (or of the xa / xb pair of a
    real pair): every 128-byte line touched by lanes `lanes` is covered by its
    first `npieces` 16-byte pieces (all 8, or the 4 of a 4-ring workgroup of
    complex fields, whose paired workgroup moves the other 4 through the same
    L2)."""
    got = {}
    for t in np.flatnonzero(lanes):
        got.setdefault(int(line[t]), set()).add(int(piece[t]))
    for ln, pcs in got.items():
        assert pcs == set(range(npieces)), (ln, sorted(pcs))


CASES = [(R, real, part, nf) for R in (8, 16, 32) for real in (False, True)
         for part in (False, True) for nf in ("full", "short")]


def case_params(R, real, part, nf, W=8):
    N = 64 * R
    lmax = N // 2 - (37 if part else 0)
    nfield = (2 * W if real else W) if nf == "full" else (2 * W - 5 if real else W - 3)
    m_first = 0 if real else -lmax
    nslot = lmax + 1 if real else 2 * lmax + 1
    return N, lmax, nfield, m_first, nslot


def coop_threads(real, nfield, W):
    """Cooperative share of thread tid: ring c, bins 64 i + bq; first column
    2 cf of the ring in its workgroup's columns."""
    tid = np.arange(64 * W)
    c, bq = tid % W, tid // W
    cf = c * (2 if real else 1)
    return c, bq, cf, cf < nfield, real & (cf + 1 < nfield)


@pytest.mark.parametrize("R,real,part,nf", CASES)
def test_synthesis_bin_gather(R, real, part, nf):
    W = synth_waves(R)
    N, lmax, nfield, m_first, nslot = case_params(R, real, part, nf, W)
    LPW = lds_per_wave(R)
    H, CPT, BH = N // LPW, LPW // 64, LPW // 2
    CPH = BH // 64
    rng = np.random.default_rng(R + 7 * real + 3 * part)
    ncg = 2 if real else 1
    F = rng.standard_normal((ncg, nslot, 16))   # one colatitude: F[cg][slot][16]
    c, bq, cf, cv, cv1 = coop_threads(real, nfield, W)
    col = 2 * cf
    lane = np.arange(64)

    def ld2(colv, slot, mask):
        out = np.zeros(len(slot), complex)
        cg, cc, sl = colv[mask] // 16, colv[mask] % 16, slot[mask]
        out[mask] = F[cg, sl, cc] + 1j * F[cg, sl, cc + 1]
        return out

    def whole_lines(mask, colv, slot):
        """mask / colv / slot over 512 threads, or 1024 for the xa, xb pair."""
        if nf != "full":
            return
        T = 64 * W
        n = len(mask) // T
        for wv in range(W):
            sel = np.zeros(len(mask), bool)
            for h in range(n):
                sel[T * h + 64 * wv:T * h + 64 * wv + 64] = True
            check_whole_lines(mask & sel, (colv // 16) * nslot + slot, (colv % 16) // 2,
                              min(8, W * (2 if real else 1)))

    v = np.full((W, R, 64), np.nan + 0j)
    if not real:
        st = {}
        for r in range(H):
            for i in range(CPT):
                k = r * LPW + 64 * i + bq
                lo = k <= lmax
                m = cv & (lo | (k >= N - lmax))
                sl = np.where(m, np.where(lo, k - m_first, k - N - m_first), 0)
                whole_lines(m, col, sl)
                st[r, i] = ld2(col, sl, m)
        RH, IH = (N // 2) // LPW, ((N // 2) % LPW) // 64
        both = cv & (bq == 0) & (N // 2 <= lmax)
        st[RH, IH] = st[RH, IH] + ld2(col, np.where(both, -N // 2 - m_first, 0), both)
        for r in range(H):
            lds = np.full(W * LPW, np.nan + 0j)
            written = np.zeros(W * LPW, int)
            for i in range(CPT):
                addr = c * LPW + xch(c, 64 * i + bq, W)
                assert addr.min() >= 0 and addr.max() < W * LPW
                for wv in range(W):
                    check_b128_phases(addr[64 * wv:64 * wv + 64])
                lds[addr] = st[r, i]
                np.add.at(written, addr, 1)
            assert (written == 1).all()
            for w in range(W):
                for j in range(CPT):
                    a = w * LPW + xch(w, 64 * j + lane, W)
                    check_b128_phases(a)
                    v[w, r * CPT + j] = lds[a]
    else:
        sa, sb = {}, {}
        for r in range(H):
            for i in range(CPH):
                kk = r * BH + 64 * i + bq
                ma, mb = cv & (kk <= lmax), cv1 & (kk <= lmax)
                ks = np.where(ma, kk, 0)
                whole_lines(np.concatenate([ma, mb]), np.concatenate([col, col + 2]),
                            np.concatenate([ks, ks]))
                sa[r, i], sb[r, i] = ld2(col, ks, ma), ld2(col + 2, ks, mb)
        for r in range(H):
            lds = np.full(W * LPW, np.nan + 0j)
            written = np.zeros(W * LPW, int)
            for i in range(CPH):
                addr = c * LPW + xch(c, 64 * i + bq, W)
                assert addr.min() >= 0 and addr.max() + BH < W * LPW
                for wv in range(W):
                    check_b128_phases(addr[64 * wv:64 * wv + 64])
                    check_b128_phases(addr[64 * wv:64 * wv + 64] + BH)
                lds[addr], lds[addr + BH] = sa[r, i], sb[r, i]
                np.add.at(written, addr, 1)
                np.add.at(written, addr + BH, 1)
            assert (written == 1).all()
            for w in range(W):
                for j in range(CPH):
                    a = w * LPW + xch(w, 64 * j + lane, W)
                    check_b128_phases(a)
                    xa, xb = lds[a].copy(), lds[a + BH].copy()
                    if r == 0 and j == 0:
                        xa[0], xb[0] = xa[0].real, xb[0].real
                    v[w, r * CPH + j] = xa + 1j * xb
                for n1 in range(R - (r + 1) * CPH, min(R - r * CPH, R - 1) + 1):
                    kk = N - 64 * n1 - lane
                    m = (kk >= r * BH) & (kk < (r + 1) * BH)
                    a = w * LPW + xch(w, np.where(m, kk - r * BH, 0), W)
                    assert (a >= w * LPW).all() and (a < w * LPW + BH).all()
                    xa, xb = lds[a], lds[a + BH]
                    # first touch of the register unless round r - 1 set it
                    keep = v[w, n1] if n1 == R - r * CPH else np.zeros(64, complex)
                    v[w, n1] = np.where(m, np.conj(xa) + 1j * np.conj(xb), keep)
        for w in range(W):   # slot N/2: lane 0 of register R/2, loaded by the ring
            f0 = 2 * w
            na = nb = 0
            if f0 < nfield and N // 2 <= lmax:
                cg, cc = (2 * f0) // 16, (2 * f0) % 16
                na = F[cg, N // 2, cc]
                nb = F[cg, N // 2, cc + 2] if f0 + 1 < nfield else 0
            v[w, R // 2, 0] = na + 1j * nb
    # Reference: the bins of ring w from the per-ring formulas.
    for w in range(W):
        f0 = w * (2 if real else 1)
        if f0 >= nfield:
            continue
        for n1 in range(R):
            k = 64 * n1 + lane
            if not real:
                ref = np.zeros(64, complex)
                for m, sl in ((k <= lmax, k - m_first), (k >= N - lmax, k - N - m_first)):
                    ref[m] += F[0, sl[m], 2 * f0] + 1j * F[0, sl[m], 2 * f0 + 1]
            else:
                kk = np.where(k <= N // 2, k, N - k)
                m = kk <= lmax
                ks = np.minimum(kk, nslot - 1)
                cg, cc = (2 * f0) // 16, (2 * f0) % 16
                xa = np.where(m, F[cg, ks, cc] + 1j * F[cg, ks, cc + 1], 0)
                xb = (np.where(m, F[cg, ks, cc + 2] + 1j * F[cg, ks, cc + 3], 0)
                      if f0 + 1 < nfield else np.zeros(64, complex))
                edge = (kk == 0) | (kk == N // 2)
                xa = np.where(edge, xa.real, np.where(k > N // 2, np.conj(xa), xa))
                xb = np.where(edge, xb.real, np.where(k > N // 2, np.conj(xb), xb))
                ref = xa + 1j * xb
            assert np.array_equal(v[w, n1], ref), (w, n1)


@pytest.mark.parametrize("R,real,part,nf", CASES)
def test_analysis_bin_scatter(R, real, part, nf):
    W = anal_waves(R, real)
    N, lmax, nfield, m_first, nslot = case_params(R, real, part, nf, W)
    LPW = lds_per_wave(R)
    H, CPT, BH = N // LPW, LPW // 64, LPW // 2
    CPH = BH // 64
    rng = np.random.default_rng(50 + R + 7 * real + 3 * part)
    c, bq, cf, cv, cv1 = coop_threads(real, nfield, W)
    lane = np.arange(64)
    ncg = 2 if real else 1
    FA = {}   # (slot, column) -> value

    def store(mask, slot, colv, val):
        for t in np.flatnonzero(mask):
            key = (int(slot[t]), int(colv[t]))
            assert key not in FA, key
            FA[key] = val[t]

    def whole_lines(masks_slots_cols):
        if nf != "full":
            return
        for wv in range(W):
            line, piece, sel = [], [], []
            for mask, slot, colv in masks_slots_cols:
                sl = slice(64 * wv, 64 * wv + 64)
                sel.append(mask[sl])
                line.append((colv[sl] // 16) * nslot + slot[sl])
                piece.append((colv[sl] % 16) // 2)
            check_whole_lines(np.concatenate(sel), np.concatenate(line), np.concatenate(piece),
                              min(8, W * (2 if real else 1)))

    if not real:
        X = rng.standard_normal((W, N)) + 1j * rng.standard_normal((W, N))   # weighted bins
        for r in range(H):
            lds = np.full(W * LPW, np.nan + 0j)
            for w in range(W):
                for j in range(CPT):
                    a = w * LPW + xch(w, 64 * j + lane, W)
                    check_b128_phases(a)
                    assert np.isnan(lds[a]).all()
                    lds[a] = X[w, 64 * (r * CPT + j) + lane]
            for i in range(CPT):
                b = 64 * i + bq
                k = r * LPW + b
                a = c * LPW + xch(c, b, W)
                for wv in range(W):
                    check_b128_phases(a[64 * wv:64 * wv + 64])
                val = lds[a]
                assert not np.isnan(val).any()
                m1, m2 = cv & (k <= lmax), cv & (k >= N - lmax)
                s1, s2 = np.where(m1, k - m_first, 0), np.where(m2, k - N - m_first, 0)
                store(m1, s1, 2 * cf, val)
                store(m2, s2, 2 * cf, val)
                whole_lines([(m1, s1, 2 * cf)])
                whole_lines([(m2, s2, 2 * cf)])
        expect = {}
        for w in range(min(W, nfield)):
            for k in range(N):
                if k <= lmax:
                    expect[(k - m_first, 2 * w)] = X[w, k]
                if k >= N - lmax:
                    expect[(k - N - m_first, 2 * w)] = X[w, k]
    else:
        XA = rng.standard_normal((W, N // 2)) + 1j * rng.standard_normal((W, N // 2))
        XB = rng.standard_normal((W, N // 2)) + 1j * rng.standard_normal((W, N // 2))
        for r in range(H):
            lds = np.full(W * LPW, np.nan + 0j)
            for w in range(W):
                for j in range(CPH):
                    kl = 64 * (r * CPH + j) + lane
                    a = w * LPW + xch(w, 64 * j + lane, W)
                    check_b128_phases(a)
                    check_b128_phases(a + BH)
                    lds[a], lds[a + BH] = XA[w, kl], XB[w, kl]
            for i in range(CPH):
                b = 64 * i + bq
                k = r * BH + b
                a = c * LPW + xch(c, b, W)
                assert a.max() + BH < W * LPW
                for wv in range(W):
                    check_b128_phases(a[64 * wv:64 * wv + 64])
                m = cv & (k <= lmax)
                mb = m & cv1
                store(m, k, 2 * cf, lds[a])
                store(mb, k, 2 * cf + 2, lds[a + BH])
                whole_lines([(m, k, 2 * cf), (mb, k, 2 * cf + 2)])
        expect = {}
        for w in range(W):
            f0 = 2 * w
            for k in range(min(lmax + 1, N // 2)):
                if f0 < nfield:
                    expect[(k, 2 * f0)] = XA[w, k]
                if f0 + 1 < nfield:
                    expect[(k, 2 * f0 + 2)] = XB[w, k]
    assert FA.keys() == expect.keys()
    for key, val in expect.items():
        assert FA[key] == val, key


@pytest.mark.parametrize("R", [16, 32])
def test_real_pair_radix2_split(R):
    """fft_anal_w for real pairs at R >= 16: lane loads the even / odd samples
    x[2m], x[2m + 1] (m = 64 a + lane) of both rings, runs two half-length
    wavefront FFTs (twiddles of the length-N tables at stride 2, which are the
    half-length ones), and combines X[k] = E[k] + W_N^k O[k],
    X[k + N/2] = E[k] - W_N^k O[k] into v[t] = X[lane + 64 t]."""
    N = 64 * R
    rng = np.random.default_rng(300 + R)
    xa, xb = rng.standard_normal(N), rng.standard_normal(N)
    z = xa + 1j * xb
    lane = np.arange(64)
    R2 = R // 2
    # Loads: pair a of the lane holds samples 128 a + 2 lane (+1) of each ring.
    ein = np.zeros(64 * R2, complex)
    oin = np.zeros(64 * R2, complex)
    for a in range(R2):
        idx = 128 * a + 2 * lane
        ein[64 * a + lane] = xa[idx] + 1j * xb[idx]
        oin[64 * a + lane] = xa[idx + 1] + 1j * xb[idx + 1]
    E = wave_fft(ein, R2, -1)
    O = wave_fft(oin, R2, -1)
    tw = np.exp(-2j * np.pi * np.arange(N) / N)
    v = np.zeros((R, 64), complex)
    for t in range(R2):
        k = lane + 64 * t
        wo = O[k] * tw[k]
        v[t] = E[k] + wo
        v[t + R2] = E[k] - wo
    X = np.fft.fft(z)
    for t in range(R):
        assert np.max(np.abs(v[t] - X[lane + 64 * t])) < 1e-12 * np.sqrt(N)
    # Twiddle-table indices at stride 2 stay inside the length-N tables.
    for k1 in range(1, R2):
        assert 2 * k1 < R
    assert 7 * 7 * R2 * 2 < N


# ---------------------------------------------------------------------------
# Workgroup order of the wavefront ring FFTs (xcd_pair_map, xcd_pair_grid).
# ---------------------------------------------------------------------------
def xcd_pair_grid(nfg, ntheta):
    units = (nfg >> 1) * ((ntheta + 3) >> 2)
    return (units + 7) // 8 * 64


def xcd_pair_map(b, nfg, ntheta):
    x, sq = b & 7, b >> 3
    u = ((sq >> 3) << 3) | x
    np_ = nfg >> 1
    quad = u // np_
    fg = 2 * (u - quad * np_) + (sq & 1)
    theta = 4 * quad + ((sq >> 1) & 3)
    return (fg, theta) if theta < ntheta else None


@pytest.mark.parametrize("nfg,ntheta", [(2, 1), (2, 3), (4, 5), (8, 257), (16, 1025),
                                        (6, 65), (10, 130), (16, 513)])
def test_xcd_pair_map(nfg, ntheta):
    """Every (field group, colatitude) runs exactly once; the two field groups
    of a pair and the four colatitudes of a quad run on one XCD (b % 8),
    within 8 consecutive workgroups of that XCD."""
    where = {}
    for b in range(xcd_pair_grid(nfg, ntheta)):
        r = xcd_pair_map(b, nfg, ntheta)
        if r is None:
            continue
        fg, t = r
        assert 0 <= fg < nfg and 0 <= t < ntheta
        assert r not in where
        where[r] = (b & 7, (b >> 3) >> 3)   # XCD, unit slot on that XCD
    assert len(where) == nfg * ntheta
    for (fg, t), loc in where.items():
        for other in ((fg ^ 1, t), (fg, t - t % 4), (fg ^ 1, t - t % 4)):
            if other in where:
                assert where[other] == loc
