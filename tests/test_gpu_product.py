"""Analysis of pointwise field products (gsh_forward_product): the product
field of CanonicalComponentFieldBinary.h:203-253 (upper index N0 + N1,
CanonicalComponentFieldOverloads.h:207-233) analysed with the product formed
inside the ring FFT's loads.  Checked against the unfused path (the product
field formed in memory, then the analysis) and against the oracle's
synthesis -> product -> analysis on the CPU, for complex and real fields, the
LDS ring FFT (L = 64) and the wavefront one (L = 256, 512), a batch, and the
split-BF16 mode."""
import sys

import numpy as np
import pytest

import gshtrans_amd as G
import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    return _t


def rel(torch, a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


@pytest.mark.parametrize("L,n0,n1,real,batch", [(64, 1, -1, False, 2), (64, 0, 0, True, 3),
                                                (256, 1, 1, False, 2), (256, 0, 0, True, 5),
                                                (512, 0, 2, False, 1), (512, 0, 0, True, 2)])
def test_fused_product_matches_unfused(torch, L, n0, n1, real, batch):
    say(f"[product] L={L} n0={n0} n1={n1} real={real} batch={batch}")
    g = G.GaussLegendreGrid(L, 2)
    u = torch.from_numpy(G.random_coefficients(L, n0, real=real, seed=3, batch=batch)).cuda()
    v = torch.from_numpy(G.random_coefficients(L, n1, real=real, seed=4, batch=batch)).cuda()
    fa = g.inverse(L, n0, u, real=real)
    fb = g.inverse(L, n1, v, real=real)
    fused = g.forward_product(L, n0 + n1, fa, fb, real=real)
    unfused = g.forward(L, n0 + n1, fa * fb, real=real)
    e = rel(torch, fused, unfused)
    say(f"    fused vs unfused {e:.2e}")
    assert e < 1e-13
    assert torch.equal(g.product(L, u, n0, v, n1, real=real), fused)
    mixed = g.product(L, u, n0, v, n1, real=real, precision="bf16x6")
    em = rel(torch, mixed, fused)
    say(f"    bf16x6 product vs FP64 {em:.2e}")
    # Two bf16x6 syntheses and one bf16x6 analysis, each within 1e-7 of FP64.
    assert em < 2e-7


@pytest.mark.parametrize("real", [False, True])
def test_product_matches_oracle(torch, real):
    L, n0, n1 = 32, (0 if real else 1), (0 if real else 1)
    og = oracle.OracleGrid(L, 2, table=False)
    g = G.GaussLegendreGrid(L, 2)
    u = G.random_coefficients(L, n0, real=real, seed=11)
    v = G.random_coefficients(L, n1, real=real, seed=12)
    ref = og.forward(L, n0 + n1, og.inverse(L, n0, u, real) * og.inverse(L, n1, v, real), real)
    got = g.product(L, torch.from_numpy(u).cuda(), n0, torch.from_numpy(v).cuda(), n1,
                    real=real).cpu().numpy()
    e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    say(f"[product vs oracle] real={real}: {e:.2e}")
    assert e < 1e-12


def test_product_argument_errors(torch):
    g = G.GaussLegendreGrid(16, 1)
    f = torch.zeros((1, 17, 32), dtype=torch.complex128, device="cuda")
    with pytest.raises(G.GSHError):
        g.forward_product(16, 2, f, f)          # n0 + n1 outside the grid's upper indices
    with pytest.raises(G.GSHError):
        g.forward_product(0, 0, f, f)           # the lMax = 0 quirk has no product form
