"""Analysis of pointwise field expressions (gsh_forward_expr): the expression
fields the reference builds with CanonicalComponentFieldOverloads.h:20-266
(sum, difference, product, quotient: CanonicalComponentFieldBinary.h;
negation, conj, real, imag, scalar multiples and shifts:
CanonicalComponentFieldUnary.h), analysed with the expression formed inside
the ring FFT's loads.  Every operation, for complex and real fields, is held
to the oracle's synthesis -> expression (numpy) -> analysis on the CPU at
L = 32 (LDS ring FFT), 256 (wavefront FFT, R = 8) and 512 (wavefront FFT,
real rings in even/odd pairs), within 1e-12 relative L2."""
import sys

import numpy as np
import pytest

import gshtrans_amd as G
import oracle

pytestmark = pytest.mark.gpu

OPS = ["product", "sum", "difference", "quotient", "negate", "conj", "real", "imag", "scale",
       "shift"]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    return _t


def upper_indices(op, real):
    """(n0 of a, n1 of b, n of the expression) allowed by the reference's types."""
    if op == "product":
        return (1, 1, 2) if not real else (0, -1, -1)
    if op == "quotient":
        return (1, 0, 1) if not real else (0, 0, 0)
    if op == "shift":
        return 0, 0, 0
    return (-1, -1, -1) if not real else (1, 1, 1)


def expr_numpy(op, a, b, s):
    return {
        "product": lambda: a * b,
        "sum": lambda: a + b,
        "difference": lambda: a - b,
        "quotient": lambda: a / b,
        "negate": lambda: -a,
        "conj": lambda: np.conj(a),
        "real": lambda: np.real(a).copy(),
        "imag": lambda: np.imag(a).copy() if np.iscomplexobj(a) else np.zeros_like(a),
        "scale": lambda: s * a,
        "shift": lambda: a + s,
    }[op]()


_GRIDS = {}


def grids(L):
    if L not in _GRIDS:
        _GRIDS[L] = (G.GaussLegendreGrid(L, 2), oracle.OracleGrid(L, 2, table=False))
    return _GRIDS[L]


@pytest.mark.parametrize("L", [32, 256, 512])
@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("op", OPS)
def test_expression_matches_oracle(torch, L, real, op):
    g, og = grids(L)
    n0, n1, n = upper_indices(op, real)
    say(f"[expr] L={L} real={real} op={op} n0={n0} n1={n1} n={n}")
    u = G.random_coefficients(L, n0, real=real, seed=21, decay=1.0)
    v = G.random_coefficients(L, n1, real=real, seed=22, decay=1.0)
    if op == "quotient":
        v = v * 0.1
        v[0] += 12.0  # l = 0, m = 0: the divisor stays away from zero
    s = 1.5 if real else (0.3 - 1.7j)
    fa_h = og.inverse(L, n0, u, real)
    fb_h = og.inverse(L, n1, v, real)
    e_h = expr_numpy(op, fa_h, fb_h, s)
    out_real = real or op in ("real", "imag")
    ref = og.forward(L, n, e_h, out_real)
    fa = g.inverse(L, n0, torch.from_numpy(u).cuda(), real=real)
    fb = g.inverse(L, n1, torch.from_numpy(v).cuda(), real=real)
    got = g.forward_expr(L, n, op, fa, fb if op in OPS[:4] else None,
                         scalar=s if op in ("scale", "shift") else None, real=real)
    got = got.cpu().numpy()
    den = np.linalg.norm(ref)
    if op == "imag" and real:
        assert np.max(np.abs(got)) == 0.0
        return
    e = np.linalg.norm(got - ref) / den
    say(f"    vs oracle {e:.2e}")
    assert e < 1e-12


def test_expression_accumulates_and_splits(torch):
    """accumulate adds into the coefficients (GaussLegendreGrid.h:192); a batch
    equals its fields one by one; the split-BF16 analysis of a sum is within
    the mode's bar of the FP64 one."""
    L = 512
    g, _ = grids(L)
    u = torch.from_numpy(G.random_coefficients(L, 2, seed=5, batch=3)).cuda()
    v = torch.from_numpy(G.random_coefficients(L, 2, seed=6, batch=3)).cuda()
    fa, fb = g.inverse(L, 2, u), g.inverse(L, 2, v)
    base = torch.from_numpy(G.random_coefficients(L, 2, seed=7, batch=3)).cuda()
    acc = base.clone()
    g.forward_expr(L, 2, "difference", fa, fb, coeffs=acc, accumulate=True)
    plain = g.forward_expr(L, 2, "difference", fa, fb)
    assert torch.allclose(acc, base + plain, rtol=0, atol=1e-14)
    one = g.forward_expr(L, 2, "difference", fa[1], fb[1])
    assert torch.equal(one, plain[1])
    mixed = g.forward_expr(L, 2, "sum", fa, fb, precision="bf16x6")
    fp64 = g.forward_expr(L, 2, "sum", fa, fb)
    e = float(torch.linalg.vector_norm(mixed - fp64) / torch.linalg.vector_norm(fp64))
    say(f"[expr] bf16x6 sum vs FP64 {e:.2e}")
    assert e < 1e-7


def test_expression_argument_errors(torch):
    g = G.GaussLegendreGrid(16, 1)
    f = torch.zeros((1, 17, 32), dtype=torch.complex128, device="cuda")
    r = torch.zeros((1, 17, 32), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        g.forward_expr(16, 0, "sum", f)                     # second operand missing
    with pytest.raises(ValueError):
        g.forward_expr(16, 0, "scale", r, scalar=1j, real=True)  # complex scalar, real field
    with pytest.raises(G.GSHError):
        g.forward_expr(16, 2, "conj", f)                    # n outside the grid's range
    lib = G.library()
    c = torch.zeros((1, G.coefficient_size(16, 0)), dtype=torch.complex128, device="cuda")
    st = lib.gsh_forward_expr(g.handle, 16, 0, G.COMPLEX, 11, 0, 1, f.data_ptr(), None,
                              g.FieldSize(), 0.0, 0.0, c.data_ptr(), c.shape[1], 0, None)
    assert st == 1  # GSH_ERR_INVALID_ARGUMENT: no such expression
    gr = G.GaussLegendreGrid(16, 1, mrange="NonNegative")
    with pytest.raises(G.GSHError):
        gr.forward_expr(16, 0, "real", f)                   # complex fields need MRange = All
