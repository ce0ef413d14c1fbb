"""Mixed-precision Legendre contraction (split-BF16 MFMA, FP64 recurrence,
seeds, weights and FFT): relative L2 error of each field against the FP64
GPU transform, which is itself pinned to the oracle at <= 1e-12 by
test_gpu_parity.py.  Bars: bf16x6 <= 1e-7 (the precision target); bf16x3
<= 3e-5 (two-way split, ~16 significant bits per operand; measured up to
9.2e-6 at lMax=1024, batch 64) and fp32 <= 3e-6 (FP32 MFMA on FP32-rounded
operands with FP32 accumulation over up to L + 1 terms; measured up to
7.6e-7 at lMax=1024, batch 64): these two modes are reported, not held to
the 1e-7 target, and their bars sit about 3-4x above the measured maxima so
that a lost product term or chunk shows.

The lMax=1024 batch-64 cases are the headline workload: 64 real n=0 fields
and 64 complex n=2 fields, i.e. 128 columns, one whole 8-tile super-group
of the split-BF16 kernels."""
import sys

import numpy as np
import pytest

import gshtrans_amd as G

pytestmark = pytest.mark.gpu
SEED0 = 20261202
BARS = {"bf16x6": 1e-7, "bf16x3": 3e-5, "fp32": 3e-6}


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    return _t


_GRIDS = {}


def grid(L):
    if L not in _GRIDS:
        _GRIDS[L] = G.GaussLegendreGrid(L, min(2, L))
    return _GRIDS[L]


def per_field_rel(torch, a, b):
    d = (a - b).reshape(a.shape[0], -1)
    return (torch.linalg.vector_norm(d, dim=1) /
            torch.linalg.vector_norm(b.reshape(b.shape[0], -1), dim=1)).cpu().numpy()


@pytest.mark.parametrize("prec", ["bf16x6", "bf16x3", "fp32"])
@pytest.mark.parametrize("L,n,real,batch", [(64, 0, False, 3), (64, 2, True, 5),
                                            (512, 2, False, 8), (300, -1, False, 2),
                                            (1024, 0, True, 16), (1024, 2, False, 16),
                                            (100, 1, True, 70), (1024, 0, True, 64),
                                            (1024, 2, False, 64), (63, 0, True, 4),
                                            (255, 0, False, 3), (512, 0, False, 20),
                                            (40, 0, True, 150), (64, 1, False, 5),
                                            (129, -2, False, 3), (127, 2, False, 3)])
def test_mixed_synthesis_error(torch, prec, L, n, real, batch):
    say(f"[mixed synthesis] {prec} L={L} n={n} real={real} batch={batch}")
    g = grid(L)
    c = torch.from_numpy(G.random_coefficients(L, n, real=real, seed=SEED0 + L + n,
                                               batch=batch)).cuda()
    f64 = g.inverse(L, n, c, real=real)
    fmx = g.inverse(L, n, c, real=real, precision=prec)
    e = per_field_rel(torch, fmx, f64)
    say(f"    max {e.max():.2e} (field {int(e.argmax())}) median {np.median(e):.2e}")
    assert e.max() <= BARS[prec]
    assert e.min() > 0 or prec == "fp64"


@pytest.mark.parametrize("prec", ["bf16x6", "bf16x3"])
@pytest.mark.parametrize("L,n,real,batch", [(256, 1, False, 3), (257, 0, True, 5),
                                            (512, 2, False, 4), (1024, 0, True, 6)])
def test_mixed_synthesis_last_rows(torch, prec, L, n, real, batch):
    """With ntheta = L + 1 one or two rows past a multiple of 256 (257, 258,
    513, 1025), the split-BF16 synthesis takes those last colatitudes out of
    its MFMA items: the coefficient split adds their partial sums per block
    from the FP64 coefficients and the FP64 recurrence values, and a last
    kernel sums the blocks.  Upper index 0 is synthesized over the northern
    colatitudes (their mirrors from the same sums) and its tail row is the
    equator.  Complex fields of other upper indices pair orders m and -m,
    order -m at the mirrored rows: their tail rows are the last ones for
    m >= 0 and the first ones for m < 0, checked on fields of one order sign
    each.  The whole field meets the mode's bar, and those rows equal the
    FP64 transform to 1e-12 in either mode when the sums reach the ring FFT
    in FP64 (the LDS ring FFT, here L = 257), and to one FP32 rounding of the
    bins (2^-23) when they reach the wavefront ring FFT in FP32 (L = 256,
    512, 1024); a wrong degree, order, column or block shows far above
    either."""
    say(f"[mixed synthesis, last rows] {prec} L={L} n={n} real={real} batch={batch}")
    g = grid(L)
    c = torch.from_numpy(G.random_coefficients(L, n, real=real, seed=SEED0 + 7 * L + n,
                                               batch=batch)).cuda()
    f64 = g.inverse(L, n, c, real=real)
    fmx = g.inverse(L, n, c, real=real, precision=prec)
    e = per_field_rel(torch, fmx, f64)
    say(f"    field: max {e.max():.2e}")
    assert e.max() <= BARS[prec]
    nt = L + 1
    assert 1 <= nt % 256 <= 2
    row_bar = 2.0 ** -23 if 2 * L in (512, 1024, 2048) else 1e-12

    def check(fm, fr, tail, mfma_row):
        for t in tail:
            e = per_field_rel(torch, fm[:, t:t + 1], fr[:, t:t + 1])
            say(f"    row {t}: max {e.max():.2e} (bar {row_bar:.1e})")
            assert e.max() <= row_bar
        # the row next to them comes from the MFMA items
        if tail:
            t = mfma_row
            assert per_field_rel(torch, fm[:, t:t + 1], fr[:, t:t + 1]).max() > 1e-12

    if n == 0:
        # synthesized over the northern colatitudes; the tail row is the
        # equator of an odd ntheta
        tail = [nt // 2] if nt % 2 else []
        check(fmx, f64, tail, nt // 2 - 1)
    elif real:
        tail = list(range(nt // 256 * 256, nt))
        check(fmx, f64, tail, tail[0] - 1)
    else:
        # orders m, -m paired: order -m's rows mirrored
        k = nt % 128
        l = np.arange(abs(n), L + 1)
        order = np.concatenate([np.arange(-j, j + 1) for j in l])
        for sign, tail, mfma_row in ((1, list(range(nt - k, nt)), nt - k - 1),
                                     (-1, list(range(k)), k)):
            keep = torch.from_numpy((order >= 0) if sign > 0 else (order < 0)).cuda()
            cs = c * keep
            say(f"    orders {'m >= 0' if sign > 0 else 'm < 0'} only")
            check(g.inverse(L, n, cs, real=real, precision=prec), g.inverse(L, n, cs, real=real),
                  tail, mfma_row)


@pytest.mark.parametrize("prec", ["bf16x6", "bf16x3", "fp32"])
@pytest.mark.parametrize("L,n,real,batch", [(64, 0, False, 3), (64, 2, True, 5),
                                            (512, 2, False, 8), (300, -1, False, 2),
                                            (1024, 0, True, 16), (1024, 2, False, 16),
                                            (100, 1, True, 70), (24, 1, False, 4),
                                            (1024, 0, True, 64), (1024, 2, False, 64),
                                            (63, 0, True, 4), (255, 0, False, 3),
                                            (512, 0, False, 20), (64, 1, False, 5),
                                            (129, -2, False, 3), (127, 2, False, 3)])
def test_mixed_analysis_error(torch, prec, L, n, real, batch):
    say(f"[mixed analysis] {prec} L={L} n={n} real={real} batch={batch}")
    g = grid(L)
    c = torch.from_numpy(G.random_coefficients(L, n, real=real, seed=SEED0 + 7 * L + n,
                                               batch=batch)).cuda()
    f = g.inverse(L, n, c, real=real)
    a64 = g.forward(L, n, f, real=real)
    amx = g.forward(L, n, f, real=real, precision=prec)
    e = per_field_rel(torch, amx, a64)
    rt = per_field_rel(torch, g.forward(L, n, g.inverse(L, n, c, real=real, precision=prec),
                                        real=real, precision=prec), c)
    say(f"    analysis max {e.max():.2e} (field {int(e.argmax())}) median {np.median(e):.2e}; "
        f"round trip max {rt.max():.2e}")
    assert e.max() <= BARS[prec]
    assert rt.max() <= 2 * BARS[prec]


@pytest.mark.parametrize("prec", ["bf16x6", "bf16x3"])
@pytest.mark.parametrize("L,n,real", [(64, 0, True), (64, 0, False), (63, 0, True),
                                      (1024, 0, True), (64, 1, False), (63, -2, False),
                                      (1024, 2, False)])
def test_mixed_analysis_general_field(torch, prec, L, n, real):
    """Upper index 0 is analysed over the northern colatitudes against the sum
    and difference of each row and its mirror; complex fields of other upper
    indices analyse order -m against the mirrored rows.  A field of
    independent random samples (not a synthesized, band-limited one) makes
    the northern and southern rows unrelated: the split-BF16 analysis must
    still meet the mode's bar against FP64, for odd (equator row) and even
    row counts."""
    say(f"[mixed analysis, random field] {prec} L={L} n={n} real={real}")
    g = grid(L)
    gen = torch.Generator(device="cuda").manual_seed(SEED0 + L + real)
    shape = (5, L + 1, 2 * L)
    if real:
        f = torch.randn(shape, dtype=torch.float64, device="cuda", generator=gen)
    else:
        f = torch.complex(torch.randn(shape, dtype=torch.float64, device="cuda", generator=gen),
                          torch.randn(shape, dtype=torch.float64, device="cuda", generator=gen))
    a64 = g.forward(L, n, f, real=real)
    amx = g.forward(L, n, f, real=real, precision=prec)
    e = per_field_rel(torch, amx, a64)
    say(f"    analysis max {e.max():.2e}")
    assert e.max() <= BARS[prec]


@pytest.mark.parametrize("n", [0, 1])
def test_mixed_accumulate_and_zero_last(torch, n):
    """Accumulating analysis adds into the output; a complex field at the
    grid degree zeroes coefficient (L, L) only (order -L at degree L, which
    the paired-order kernel writes from order L's sums, stays)."""
    L = 32
    g = grid(L)
    c = torch.from_numpy(G.random_coefficients(L, n, seed=3, batch=2)).cuda()
    f = g.inverse(L, n, c)
    base = torch.from_numpy(G.random_coefficients(L, n, seed=4, batch=2, grid_lmax=L + 1)).cuda()
    out = base.clone()
    g.forward(L, n, f, coeffs=out, precision="bf16x6", accumulate=True)
    ref = base + g.forward(L, n, f)
    ref[:, -1] = 0
    assert float(per_field_rel(torch, out, ref).max()) < 1e-7
    assert bool((out[:, -1] == 0).all())


@pytest.mark.parametrize("prec", ["bf16x6", "bf16x3"])
def test_mixed_synthesis_on_fresh_grid(torch, prec):
    """The split-BF16 synthesis starts its recurrences at checkpoints that the
    first call on a grid must build: synthesis before any analysis."""
    say(f"[mixed synthesis, fresh grid] {prec} L=96 n=1")
    g = G.GaussLegendreGrid(96, 1)
    c = torch.from_numpy(G.random_coefficients(96, 1, real=False, seed=SEED0 + 7,
                                               batch=3)).cuda()
    fmx = g.inverse(96, 1, c, real=False, precision=prec)
    f64 = g.inverse(96, 1, c, real=False)
    e = per_field_rel(torch, fmx, f64)
    say(f"    max {e.max():.2e}")
    assert e.max() <= BARS[prec]
