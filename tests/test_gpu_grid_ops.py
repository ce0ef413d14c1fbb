"""Grid helpers on the device (GridBase.h:60-73, 115-172;
ScalarFieldOverloads.h:288-316): Weights, ProjectFunction, Integrate, the L2
inner product and the seeded coefficient generator.  Pinned by closed forms:
the sphere's area 4 pi; Parseval for band-limited fields, <S u, S v> =
sum conj(u) v, which the Gauss-Legendre grid integrates exactly (the D^l_{mn}
of one n are orthonormal in the Ortho normalisation); the device generator
against the host stream of gshtrans_amd.random_coefficients."""
import sys

import numpy as np
import pytest

import gshtrans_amd as G

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    return _t


@pytest.mark.parametrize("L", [16, 255, 1024])
def test_weights_project_integrate(torch, L):
    g = G.GaussLegendreGrid(L, 2)
    w = g.weights()
    assert abs(float(w.sum()) - 4 * np.pi) < 1e-12
    one = g.project(lambda th, ph: torch.ones_like(th))
    assert abs(complex(g.integrate(one)) - 4 * np.pi) < 1e-12
    # cos(theta) and sin(theta) cos(phi) integrate to zero; cos^2 to 4 pi / 3.
    c = g.project(lambda th, ph: torch.cos(th))
    s = g.project(lambda th, ph: torch.sin(th) * torch.cos(ph))
    c2 = g.project(lambda th, ph: torch.cos(th) ** 2)
    assert abs(complex(g.integrate(c))) < 1e-13 and abs(complex(g.integrate(s))) < 1e-13
    assert abs(complex(g.integrate(c2)) - 4 * np.pi / 3) < 1e-12
    # The same weighted sums in torch.
    f = torch.randn(3, L + 1, 2 * L, dtype=torch.complex128, device="cuda")
    ref = (f * w).sum(dim=(1, 2))
    assert torch.allclose(g.integrate(f), ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("L,n,real", [(64, 0, False), (64, 2, False), (200, -1, False),
                                      (64, 0, True), (512, 0, True)])
def test_inner_product_parseval(torch, L, n, real):
    g = G.GaussLegendreGrid(L, 2)
    u = torch.from_numpy(G.random_coefficients(L, n, real=real, seed=5, batch=2)).cuda()
    v = torch.from_numpy(G.random_coefficients(L, n, real=real, seed=6, batch=2)).cuda()
    fu, fv = g.inverse(L, n, u, real=real), g.inverse(L, n, v, real=real)
    got = g.integrate(fu, fv)
    if real:
        # Real field: c_{l,-m} = (-1)^m conj(c_lm) (NonNegative storage), so
        # orders 0 < m < N/2 count twice; the c2r synthesis takes the Nyquist
        # bin m = N/2 = L once (Re c cos(L phi), sum_j cos^2 = N), so it counts once.
        ms = np.concatenate([np.arange(0, l + 1) for l in range(abs(n), L + 1)])
        wgt = torch.from_numpy(np.where((ms == 0) | (ms == L), 1.0, 2.0)).cuda()
        ref = (wgt * (u.conj() * v).real).sum(dim=1).to(torch.complex128)
    else:
        ref = (u.conj() * v).sum(dim=1)
    e = float((got - ref).abs().max() / ref.abs().max())
    say(f"[parseval] L={L} n={n} real={real}: {e:.2e}")
    assert e < 1e-12


@pytest.mark.parametrize("L,lmax,n,real,decay", [(32, 32, 0, False, 1.0), (32, 20, 2, False, 0.0),
                                                 (64, 64, 0, True, 1.5), (64, 40, 1, True, 1.0),
                                                 (1024, 1024, 2, False, 1.0)])
def test_device_random_coefficients(torch, L, lmax, n, real, decay):
    g = G.GaussLegendreGrid(L, 2)
    host = G.random_coefficients(lmax, n, real=real, seed=17, decay=decay, grid_lmax=L, batch=3)
    dev = g.random_coefficients(lmax, n, real=real, seed=17, decay=decay, batch=3).cpu().numpy()
    assert dev.shape == host.shape
    assert np.array_equal(dev == 0, host == 0) and np.array_equal(dev.imag == 0, host.imag == 0)
    e = float(np.abs(dev - host).max() / np.abs(host).max())
    say(f"[random] L={L} lmax={lmax} n={n} real={real}: {e:.2e}")
    assert e < 1e-14
