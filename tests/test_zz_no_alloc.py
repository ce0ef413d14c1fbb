"""The allocation-free integration entry point (gsh_integrate_noalloc) against
gsh_integrate, through the C ABI: same kernels, so the results must be
bit-identical, on two calls in a row (the second smaller, over the scratch
contents the first left behind), plus the refusal of a call larger than the
static scratch."""
import ctypes

import numpy as np
import pytest

import gshtrans_amd as G

pytestmark = pytest.mark.gpu

SCRATCH_RING_SUMS = 262144  # (field, colatitude) pairs the static scratch holds


def _fields(L, batch, real, seed):
    import torch
    rng = np.random.default_rng(seed)
    shape = (batch, L + 1, 2 * L)
    a = rng.standard_normal(shape)
    if not real:
        a = a + 1j * rng.standard_normal(shape)
    return torch.from_numpy(a).to("cuda")


def _integrate(fn, grid, real, fa, fb):
    import torch
    out = torch.full((fa.shape[0],), float("nan"), dtype=torch.complex128, device="cuda")
    st = fn(grid.handle, G.REAL if real else G.COMPLEX, fa.shape[0], fa.data_ptr(),
            None if fb is None else fb.data_ptr(), grid.FieldSize(), out.data_ptr(), None)
    torch.cuda.synchronize()
    return st, out.cpu().numpy()


@pytest.mark.parametrize("real", [True, False])
@pytest.mark.parametrize("inner", [False, True])
def test_noalloc_equals_integrate(real, inner):
    lib = G.library()
    L = 96
    grid = G.GaussLegendreGrid(L, 0)
    # Two calls in a row, the second with fewer fields and a smaller grid's
    # worth of entries left over from the first in the scratch.
    for batch, seed in ((12, 1), (5, 2)):
        print(f"noalloc: L={L} batch={batch} real={real} inner={inner}", flush=True)
        fa = _fields(L, batch, real, seed)
        fb = _fields(L, batch, real, seed + 100) if inner else None
        st0, ref = _integrate(lib.gsh_integrate, grid, real, fa, fb)
        st1, got = _integrate(lib.gsh_integrate_noalloc, grid, real, fa, fb)
        assert st0 == G.GSH_OK and st1 == G.GSH_OK
        assert np.all(np.isfinite(ref))
        assert np.array_equal(got, ref)
    grid.close()


def test_noalloc_second_grid_and_area():
    # A second grid on the same device shares the scratch; the integral of 1
    # is the sphere's area.
    import torch
    lib = G.library()
    for L in (64, 33):
        grid = G.GaussLegendreGrid(L, 0)
        ones = torch.ones((3, L + 1, 2 * L), dtype=torch.float64, device="cuda")
        st, got = _integrate(lib.gsh_integrate_noalloc, grid, True, ones, None)
        assert st == G.GSH_OK
        assert np.max(np.abs(got - 4 * np.pi)) < 1e-12
        grid.close()


def test_noalloc_refuses_calls_past_the_scratch():
    import torch
    lib = G.library()
    L = 1024
    grid = G.GaussLegendreGrid(L, 0)
    most = SCRATCH_RING_SUMS // (L + 1)
    # Only the size is examined before the refusal: no kernel is launched, so
    # one field's buffer serves as the pointer.
    f = torch.zeros((1, L + 1, 2 * L), dtype=torch.float64, device="cuda")
    out = torch.zeros(most + 1, dtype=torch.complex128, device="cuda")
    st = lib.gsh_integrate_noalloc(grid.handle, G.REAL, most + 1, f.data_ptr(), None,
                                   grid.FieldSize(), out.data_ptr(), None)
    assert st == G.GSH_ERR_SCRATCH_LIMIT
    assert "static scratch" in lib.gsh_status_string(st).decode()
    # The largest batch that fits runs and agrees with gsh_integrate.
    fa = torch.ones((most, L + 1, 2 * L), dtype=torch.float64, device="cuda")
    st0, ref = _integrate(lib.gsh_integrate, grid, True, fa, None)
    st1, got = _integrate(lib.gsh_integrate_noalloc, grid, True, fa, None)
    assert st0 == G.GSH_OK and st1 == G.GSH_OK and np.array_equal(got, ref)
    grid.close()
