"""FP64 parity of the HIP transform path against the CPU oracle (which
restates GSHTrans/src/Wigner.h and GaussLegendreGrid.h step for step).

Bar: relative L2 <= 1e-12 against the oracle on the same seeded inputs, and
the reference's own round-trip tolerance |c - A(S(c))| <= 5e4 eps
(tests/CheckCoeff2Coeff.h:80-83).  Every call goes through the C ABI in
libgshtrans_hip.so.
"""
import sys

import numpy as np
import pytest

import gshtrans_amd as G
import oracle

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps
SEED0 = 20261202


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available(), "GPU tests need a device"
    return _t


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.detach().cpu().numpy()


_ORACLES = {}


def oracle_grid(L, nmax, mrange="All", nrange="All"):
    key = (L, nmax, mrange, nrange)
    if key not in _ORACLES:
        _ORACLES[key] = oracle.OracleGrid(L, nmax, mrange, nrange, table=False)
    return _ORACLES[key]


def check_pair(torch, L, nmax, lmax, n, real, decay, seed, grid=None, tol=1e-12):
    say(f"[parity] L={L} lmax={lmax} n={n} real={real} seed={seed}")
    g = grid or G.GaussLegendreGrid(L, nmax)
    og = oracle_grid(L, nmax)
    c = G.random_coefficients(lmax, n, real=real, seed=seed, decay=decay, grid_lmax=L)
    f_ref = og.inverse(lmax, n, c, real)
    f_gpu = host(g.inverse(lmax, n, dev(torch, c), real=real))
    torch.cuda.synchronize()
    e_syn = rel_l2(f_gpu, f_ref)
    c_ref = og.forward(lmax, n, f_ref, real)
    c_gpu = host(g.forward(lmax, n, dev(torch, f_ref), real=real))
    e_ana = rel_l2(c_gpu, c_ref)
    rt = host(g.forward(lmax, n, g.inverse(lmax, n, dev(torch, c), real=real), real=real))
    e_rt = float(np.max(np.abs(rt - c)))
    say(f"    synthesis {e_syn:.2e}  analysis {e_ana:.2e}  round trip {e_rt:.2e}")
    assert e_syn <= tol
    assert e_ana <= tol
    assert e_rt <= 5e4 * EPS
    return g


def test_config1_plain_L64(torch):
    """lMax=64, nMax=2, one complex n=0 field, c ~ CN(0,1)/(l+1)."""
    check_pair(torch, 64, 2, 64, 0, False, 1.0, SEED0 + 0)


@pytest.mark.parametrize("n", [-2, -1, 0, 1, 2])
def test_config2_plain_L512(torch, n):
    """lMax=512, n in -2..2, complex, decay (l+1)^-1.5, 513 x 1024 grid."""
    g = _grid(512, 2)
    check_pair(torch, 512, 2, 512, n, False, 1.5, SEED0 + 1 + n, grid=g)


_GRIDS = {}


def _grid(L, nmax, mrange="All", nrange="All"):
    key = (L, nmax, mrange, nrange)
    if key not in _GRIDS:
        _GRIDS[key] = G.GaussLegendreGrid(L, nmax, mrange, nrange)
    return _GRIDS[key]


@pytest.mark.parametrize("L,lmax,n", [(64, 64, 0), (64, 64, 2), (64, 40, -1), (16, 16, 0),
                                      (128, 100, 1), (512, 512, 0)])
def test_real_fields(torch, L, lmax, n):
    check_pair(torch, L, 2, lmax, n, True, 1.0, SEED0 + 7, grid=_grid(L, 2))


@pytest.mark.parametrize("L,lmax,n", [(64, 40, 1), (32, 5, -2), (8, 8, 2), (2, 2, 1),
                                      (1, 1, 0), (4, 2, 2), (256, 255, 0)])
def test_truncation_and_small_grids(torch, L, lmax, n):
    nmax = min(2, L)
    check_pair(torch, L, nmax, lmax, n, False, 1.0, SEED0 + 11, grid=_grid(L, nmax))


@pytest.mark.parametrize("L,lmax,n,real", [(24, 24, 0, False), (15, 15, 1, True),
                                           (100, 80, -1, False), (33, 33, 2, False),
                                           (48, 48, 0, True), (300, 300, 2, False)])
def test_bluestein_ring_lengths(torch, L, lmax, n, real):
    """Ring lengths 2L that are not powers of two (Bluestein path)."""
    nmax = min(2, L)
    check_pair(torch, L, nmax, lmax, n, real, 1.0, SEED0 + 13, grid=_grid(L, nmax))


def test_nonnegative_grid_real(torch):
    g = G.GaussLegendreGrid(32, 2, mrange="NonNegative")
    og = oracle.OracleGrid(32, 2, mrange="NonNegative", table=False)
    c = G.random_coefficients(32, 1, real=True, seed=3)
    f = host(g.inverse(32, 1, dev(torch, c), real=True))
    assert rel_l2(f, og.inverse(32, 1, c, True)) < 1e-12
    with pytest.raises(G.GSHError) as e:
        g.inverse(32, 1, dev(torch, G.random_coefficients(32, 1)), real=False)
    assert e.value.status == 4


def test_batch_matches_single(torch):
    """Five n=1 fields in one call equal five single calls; strides honoured."""
    L, n = 128, 1
    g = _grid(L, 2)
    cs = G.random_coefficients(L, n, seed=21, batch=5)
    fb = host(g.inverse(L, n, dev(torch, cs)))
    for b in range(5):
        fs = host(g.inverse(L, n, dev(torch, cs[b])))
        assert np.array_equal(fb[b], fs)
    cb = host(g.forward(L, n, dev(torch, fb)))
    for b in range(5):
        assert np.array_equal(cb[b], host(g.forward(L, n, dev(torch, fb[b]))))


def test_strided_batch_via_capi(torch):
    import ctypes
    L, n = 32, 0
    g = _grid(L, 2)
    C = G.coefficient_size(L, n)
    cs = G.random_coefficients(L, n, seed=4, batch=3)
    padded = np.zeros((3, C + 7), dtype=np.complex128)
    padded[:, :C] = cs
    fsz = g.FieldSize()
    out = torch.zeros((3, fsz + 5), dtype=torch.complex128, device="cuda")
    din = dev(torch, padded)
    st = G.library().gsh_inverse(g.handle, L, n, G.COMPLEX, 0, 0, 3,
                                 ctypes.c_void_p(din.data_ptr()), C + 7,
                                 ctypes.c_void_p(out.data_ptr()), fsz + 5, None)
    torch.cuda.synchronize()
    assert st == 0
    o = host(out)
    for b in range(3):
        ref = host(g.inverse(L, n, dev(torch, cs[b]))).ravel()
        assert np.array_equal(o[b, :fsz], ref)
        assert np.all(o[b, fsz:] == 0)


def test_forward_accumulates_like_reference(torch):
    """ForwardTransformation adds into `out` (GaussLegendreGrid.h:192), and
    zeroes (L, L) for complex fields at lMax == L (:209-215)."""
    L, n = 32, 0
    g = _grid(L, 2)
    og = oracle_grid(L, 2)
    f = og.inverse(L, n, G.random_coefficients(L, n, seed=8))
    start = G.random_coefficients(L, n, seed=9, grid_lmax=L + 1)
    out = start.copy()
    g.ForwardTransformation(L, n, f, out)
    ref = start + og.forward(L, n, f)
    ref[-1] = 0
    assert rel_l2(out, ref) < 1e-12 and out[-1] == 0
    outd = dev(torch, start.copy())
    g.ForwardTransformation(L, n, dev(torch, f), outd)
    assert rel_l2(host(outd), ref) < 1e-12


def test_host_range_calls_match_device(torch):
    L, n = 48, -1
    g = _grid(L, 2)
    c = G.random_coefficients(L, n, seed=12)
    f = np.zeros((L + 1, 2 * L), dtype=np.complex128)
    g.InverseTransformation(L, n, c, f)
    assert np.array_equal(f, host(g.inverse(L, n, dev(torch, c))))
    fr = np.zeros((L + 1, 2 * L))
    cr = G.random_coefficients(L, n, real=True, seed=13)
    g.InverseTransformation(L, n, cr, fr)
    assert np.array_equal(fr, host(g.inverse(L, n, dev(torch, cr), real=True)))


def test_lmax_zero_quirk(torch):
    """lMax == 0 touches element 0 only (GaussLegendreGrid.h:137-140, 251-260)."""
    g = _grid(8, 2)
    c = np.array([1.5 - 0.5j])
    f = np.full((9, 16), 7.0 + 1j)
    g.InverseTransformation(0, 0, c, f)
    assert f[0, 0] == c[0] * 2 / (1 / np.sqrt(np.pi))
    assert np.all(f.ravel()[1:] == 7.0 + 1j)
    out = np.array([5.0 + 5j])
    h = np.full((9, 16), 3.0 + 1j)
    g.ForwardTransformation(0, 0, h, out)
    assert out[0] == h[0, 0] * (1 / np.sqrt(np.pi)) / 2


def test_argument_errors(torch):
    g = G.GaussLegendreGrid(8, 2, nrange="NonNegative")
    with pytest.raises(G.GSHError) as e:
        g.inverse(8, -1, dev(torch, G.random_coefficients(8, -1)))
    assert e.value.status == 2
    with pytest.raises(G.GSHError) as e:
        g.inverse(9, 0, dev(torch, G.random_coefficients(9, 0)))
    assert e.value.status == 3
    with pytest.raises(G.GSHError) as e:
        g.inverse(1, 2, dev(torch, np.zeros(4, dtype=np.complex128)))
    assert e.value.status == 2


def test_L1024_parity_one_field(torch):
    """Full bench size: one complex n=2 field at lMax=1024 against the oracle."""
    check_pair(torch, 1024, 2, 1024, 2, False, 1.0, SEED0 + 2, grid=_grid(1024, 2))


@pytest.mark.parametrize("n,real", [(0, True), (2, False)])
def test_config3_fp64_batch64(torch, n, real):
    """The headline workload in FP64: a batch of 64 real n=0 fields or 64
    complex n=2 fields at lMax=1024 (128 columns per launch), (l+1)^-1
    decay.  Fields 0 and 63 against the oracle's synthesis and analysis
    (<= 1e-12), and the whole batch's round trip (5e4 eps)."""
    L, B = 1024, 64
    g = _grid(L, 2)
    og = oracle_grid(L, 2)
    say(f"[config3 fp64] L={L} n={n} real={real} batch={B}")
    cs = G.random_coefficients(L, n, real=real, seed=SEED0 + 2, decay=1.0, batch=B)
    f = g.inverse(L, n, dev(torch, cs), real=real)
    c2 = host(g.forward(L, n, f, real=real))
    fh = host(f)
    for b in (0, B - 1):
        f_ref = og.inverse(L, n, cs[b], real)
        e_syn = rel_l2(fh[b], f_ref)
        c_ref = og.forward(L, n, f_ref, real)
        c_gpu = host(g.forward(L, n, dev(torch, f_ref), real=real))
        e_ana = rel_l2(c_gpu, c_ref)
        say(f"    field {b}: synthesis {e_syn:.2e}  analysis {e_ana:.2e}")
        assert e_syn <= 1e-12 and e_ana <= 1e-12, b
    e_rt = float(np.max(np.abs(c2 - cs)))
    say(f"    batch round trip max {e_rt:.2e}")
    assert e_rt <= 5e4 * EPS


def test_L1024_batch_round_trip_and_linearity(torch):
    """Size-independent properties on a 16-field batch at lMax=1024."""
    L = 1024
    g = _grid(L, 2)
    for (n, real) in [(2, False), (0, True)]:
        say(f"[roundtrip] L={L} n={n} real={real} batch=16")
        cs = G.random_coefficients(L, n, real=real, seed=SEED0 + 30 + n, batch=16)
        dc = dev(torch, cs)
        f = g.inverse(L, n, dc, real=real)
        c2 = host(g.forward(L, n, f, real=real))
        assert float(np.max(np.abs(c2 - cs))) < 5e4 * EPS
        # linearity: S(a c1 + c2) = a S(c1) + S(c2)
        f2 = g.inverse(L, n, dc[:2] * 0.5 + dc[2:4], real=real)
        lin = 0.5 * f[:2] + f[2:4]
        assert rel_l2(host(f2), host(lin)) < 1e-13


@pytest.mark.parametrize("L,lmax,n,real,batch", [
    (256, 256, 0, False, 9), (256, 256, 1, True, 9), (256, 130, -2, False, 3),
    (256, 70, 2, True, 2), (512, 512, 1, True, 3), (512, 400, 0, False, 17),
    (1024, 1024, 0, True, 3)])
def test_wavefront_ring_fft_batches(torch, L, lmax, n, real, batch):
    """Ring lengths 512, 1024 and 2048 take the one-wavefront-per-ring FFT;
    batches that leave a workgroup's field group (8 complex or 16 real
    fields) partly empty, odd real batches (the last real pair has one
    field), and truncation lmax < L, each field against the oracle."""
    say(f"[wavefft] L={L} lmax={lmax} n={n} real={real} batch={batch}")
    g = _grid(L, 2)
    og = oracle_grid(L, 2)
    cs = G.random_coefficients(lmax, n, real=real, seed=SEED0 + 40 + L + lmax, batch=batch,
                               grid_lmax=L)
    f = host(g.inverse(lmax, n, dev(torch, cs), real=real))
    c2 = host(g.forward(lmax, n, dev(torch, f), real=real))
    check = sorted({0, batch - 1} if L == 1024 else set(range(batch)))
    for b in check:
        f_ref = og.inverse(lmax, n, cs[b], real)
        assert rel_l2(f[b], f_ref) <= 1e-12, b
        c_ref = og.forward(lmax, n, f[b], real)
        assert rel_l2(c2[b], c_ref) <= 1e-12, b
    assert float(np.max(np.abs(c2 - cs))) < 5e4 * EPS


@pytest.mark.parametrize("L,n,real", [(1100, 1, False), (1100, 0, True)])
def test_long_bluestein_rings(torch, L, n, real):
    """lMax past 1024 with a ring length that is not a power of two (2200
    points): the LDS ring FFT's 8192-point Bluestein convolution, against the
    oracle (<= 1e-12) and through the round trip."""
    say(f"[long rings] L={L} n={n} real={real}")
    g = G.GaussLegendreGrid(L, 1)
    og = oracle.OracleGrid(L, 1, table=False)
    cs = G.random_coefficients(L, n, real=real, seed=SEED0 + 70, batch=2)
    f = host(g.inverse(L, n, dev(torch, cs), real=real))
    c2 = host(g.forward(L, n, dev(torch, f), real=real))
    f_ref = og.inverse(L, n, cs[1], real)
    e_syn = rel_l2(f[1], f_ref)
    e_ana = rel_l2(c2[1], og.forward(L, n, f[1], real))
    e_rt = float(np.max(np.abs(c2 - cs)))
    say(f"    synthesis {e_syn:.2e} analysis {e_ana:.2e} round trip {e_rt:.2e}")
    assert e_syn <= 1e-12 and e_ana <= 1e-12 and e_rt < 5e4 * EPS
    g.close()
    og.close()


def test_L2048_grid(torch):
    """The largest grid (lMax = 2048, 4096-point rings in the LDS FFT, a
    colatitude count one past a multiple of 256): the FP64 round trip of
    complex n = 2 and real n = 0 fields, the split-BF16 synthesis and
    analysis within 1e-7 of FP64, and the refusal of lMax = 2049."""
    L = 2048
    g = G.GaussLegendreGrid(L, 2)
    for n, real in ((2, False), (0, True)):
        say(f"[L2048] n={n} real={real}")
        cs = G.random_coefficients(L, n, real=real, seed=SEED0 + 80 + n, batch=2)
        dc = dev(torch, cs)
        f = g.inverse(L, n, dc, real=real)
        c2 = g.forward(L, n, f, real=real)
        e_rt = float(np.max(np.abs(host(c2) - cs)))
        fm = g.inverse(L, n, dc, real=real, precision="bf16x6")
        cm = g.forward(L, n, f, real=real, precision="bf16x6")
        e_syn = rel_l2(host(fm), host(f))
        e_ana = rel_l2(host(cm), host(c2))
        say(f"    round trip {e_rt:.2e}; bf16x6 synthesis {e_syn:.2e} analysis {e_ana:.2e}")
        assert e_rt < 5e4 * EPS and e_syn <= 1e-7 and e_ana <= 1e-7
    g.close()
    with pytest.raises(G.GSHError) as e:
        G.GaussLegendreGrid(L + 1, 0)
    assert e.value.status == 5
