"""Adjoint of the synthesis (gsh_adjoint_inverse) and the differentiable
synthesis behind torch.autograd (gshtrans_amd/autograd.py).

Pinned by:
* the defining identity Re <S c, f> = Re <c, S^H f> on random c and f
  (complex and real fields, truncated and full lmax, every FFT path);
* its relation to the analysis that the oracle pins:
  A(f) = S^H(W f), W = w_theta * 2 pi / N per ring (complex fields);
* central finite differences of the least-squares loss
  ||S_n(c) - d||^2 at 100 seeded coefficients (lMax=512, n=1, 16 complex
  fields, d = S(c_true) + N(0, 0.01) noise, i.e. sigma = 0.1 per part).
"""
import sys

import numpy as np
import pytest

import gshtrans_amd as G

pytestmark = pytest.mark.gpu
SEED0 = 20261202


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available(), "GPU tests need a device"
    return _t


_GRIDS = {}


def grid(L, nmax=2):
    if (L, nmax) not in _GRIDS:
        _GRIDS[(L, nmax)] = G.GaussLegendreGrid(L, nmax)
    return _GRIDS[(L, nmax)]


def rdot(a, b):
    return float((a.conj() * b).real.sum())


@pytest.mark.parametrize("L,lmax,n,real", [
    (32, 32, 0, False), (32, 32, 1, True), (32, 20, -2, False), (24, 24, 1, True),
    (256, 256, 2, False), (256, 256, 0, True), (512, 300, -1, True), (1024, 1024, 2, False)])
def test_adjoint_identity(torch, L, lmax, n, real):
    say(f"[adjoint] identity L={L} lmax={lmax} n={n} real={real}")
    g = grid(L)
    B = 3
    c = torch.from_numpy(G.random_coefficients(lmax, n, real=real, seed=SEED0 + L + lmax,
                                               batch=B, grid_lmax=L)).cuda()
    rng = np.random.default_rng(L + 7 * lmax)
    shape = (B, L + 1, 2 * L)
    f = rng.standard_normal(shape) if real else \
        rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    fd = torch.from_numpy(f).cuda()
    Sc = g.inverse(lmax, n, c, real=real)
    Shf = g.adjoint(lmax, n, fd, real=real)
    lhs = rdot(Sc, fd)
    rhs = rdot(c, Shf)
    scale = float(torch.linalg.vector_norm(Sc) * torch.linalg.vector_norm(fd))
    say(f"    <Sc,f>={lhs:.6e} <c,S^H f>={rhs:.6e} rel={abs(lhs - rhs) / scale:.2e}")
    assert abs(lhs - rhs) <= 1e-13 * scale


@pytest.mark.parametrize("L,lmax,n", [(64, 40, 1), (256, 200, -2), (512, 511, 2)])
def test_adjoint_of_weighted_field_is_analysis(torch, L, lmax, n):
    """A(f) = S^H(W f) for complex fields (lmax < L: no (L, L) zeroing)."""
    g = grid(L)
    c = torch.from_numpy(G.random_coefficients(lmax, n, seed=SEED0 + 3, batch=2,
                                               grid_lmax=L)).cuda()
    f = g.inverse(lmax, n, c)
    w = torch.from_numpy(g.CoLatitudeWeights() * (2 * np.pi / (2 * L))).cuda()
    a = g.forward(lmax, n, f)
    b = g.adjoint(lmax, n, f * w[None, :, None])
    err = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(a))
    assert err <= 1e-13, err


def test_adjoint_lmax0(torch):
    g = grid(8)
    for real in (False, True):
        c = torch.tensor([[1.5 - 0.5j]], dtype=torch.complex128, device="cuda")
        f = torch.zeros((1, 9, 16), dtype=torch.float64 if real else torch.complex128,
                        device="cuda")
        f[0, 0, 0] = 0.75 if real else 0.75 + 0.25j
        lhs = rdot(g.inverse(0, 0, c, real=real), f)
        rhs = rdot(c, g.adjoint(0, 0, f, real=real))
        assert abs(lhs - rhs) <= 1e-15 * abs(lhs)


def test_mixed_precision_adjoint(torch):
    L, n = 1024, 2
    g = grid(L)
    rng = np.random.default_rng(5)
    f = torch.from_numpy(rng.standard_normal((4, L + 1, 2 * L)) +
                         1j * rng.standard_normal((4, L + 1, 2 * L))).cuda()
    a64 = g.adjoint(L, n, f)
    a6 = g.adjoint(L, n, f, precision="bf16x6")
    err = float(torch.linalg.vector_norm(a6 - a64) / torch.linalg.vector_norm(a64))
    say(f"[adjoint] bf16x6 vs fp64 rel L2 {err:.2e}")
    assert err <= 1e-7


def test_least_squares_gradient_config5(torch):
    """lMax=512, n=1, 16 complex fields: autograd gradient of
    ||S(c) - d||^2 through the adjoint kernel against central finite
    differences at 100 seeded coefficients, and against the analysis of the
    weighted residual."""
    from gshtrans_amd.autograd import synthesize
    L, n, B = 512, 1, 16
    g = grid(L)
    c_true = torch.from_numpy(G.random_coefficients(L, n, seed=SEED0 + 5, batch=B)).cuda()
    d = g.inverse(L, n, c_true)
    gen = torch.Generator(device="cuda").manual_seed(SEED0 + 5)
    noise = torch.randn(d.shape, dtype=torch.complex128, device="cuda", generator=gen)
    # torch's complex normal has variance 1 split evenly; scale to sigma 0.1 per part
    d = d + noise * (0.1 * np.sqrt(2.0))
    c0 = torch.from_numpy(G.random_coefficients(L, n, seed=SEED0 + 6, batch=B)).cuda() * 0.5
    c = c0.clone().requires_grad_(True)
    loss = (synthesize(g, L, n, c) - d).abs().pow(2).sum()
    loss.backward()
    grad = c.grad.detach()
    # against the explicit adjoint of the residual: dL/dc = 2 S^H(S c - d)
    r = g.inverse(L, n, c0) - d
    ref = 2 * g.adjoint(L, n, r)
    assert float(torch.linalg.vector_norm(grad - ref) / torch.linalg.vector_norm(ref)) < 1e-14
    # finite differences on 100 coefficients (field b, index k), real and imaginary parts
    rng = np.random.default_rng(SEED0 + 5)
    C = c0.shape[1]
    picks = [(int(rng.integers(B)), int(rng.integers(C))) for _ in range(100)]
    h = 1e-3
    worst = 0.0
    for b, k in picks:
        base = c0[b].clone()
        for part, unit in ((0, 1.0), (1, 1j)):
            vals = []
            for sgn in (1, -1):
                cc = base.clone()
                cc[k] += sgn * h * unit
                vals.append(float((g.inverse(L, n, cc) - d[b]).abs().pow(2).sum()))
            fd = (vals[0] - vals[1]) / (2 * h)
            an = float(grad[b, k].real if part == 0 else grad[b, k].imag)
            worst = max(worst, abs(fd - an) / max(abs(an), 1.0))
    say(f"[adjoint] finite differences: worst rel err {worst:.2e}")
    assert worst < 1e-6
