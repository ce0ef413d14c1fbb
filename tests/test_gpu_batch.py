"""Batch plans (gsh_plan_*): many fields of uneven truncation degree and
upper index through grouped, autotuned launches and HIP-graph replays.

Every field of a plan must equal the plain single-field call (gsh_inverse /
gsh_forward) on the same input, which the FP64 parity tests pin to the
oracle; a few fields are also checked against the oracle directly.  FP32
storage is checked against the FP64 plan within FP32 rounding.  Graph
replays are interleaved with eager work and each replay is checked against
the eager result.
"""
import sys

import numpy as np
import pytest

import gshtrans_amd as G
import oracle

pytestmark = pytest.mark.gpu
SEED0 = 20261202


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available(), "GPU tests need a device"
    return _t


def rel(a, b):
    import torch
    return float(torch.linalg.vector_norm((a - b).reshape(-1)) /
                 max(float(torch.linalg.vector_norm(b.reshape(-1))), 1e-300))


def make_coeffs(torch, lmax, n, real, cmax, seed, dtype=None):
    out = np.zeros((len(lmax), cmax), dtype=np.complex128)
    for i, (l, m) in enumerate(zip(lmax, n)):
        c = G.random_coefficients(int(l), int(m), real=real, seed=seed + i)
        out[i, :c.size] = c
    t = torch.from_numpy(out).cuda()
    return t if dtype is None else t.to(dtype)


@pytest.mark.parametrize("L,real", [(128, False), (128, True), (48, False), (256, True)])
def test_plan_matches_single_calls(torch, L, real):
    rng = np.random.default_rng(L + real)
    count = 21
    lmax = np.clip(np.round(np.exp(rng.uniform(np.log(4), np.log(L), count))), 2, L).astype(int)
    lmax[0] = L
    n = rng.integers(-2, 3, count)
    if real:
        n = np.abs(n)
    n = np.where(np.abs(n) > lmax, 0, n)
    g = G.GaussLegendreGrid(L, 2)
    plan = G.BatchPlan(g, lmax, n, real=real)
    say(f"[plan] L={L} real={real} groups={plan.info()}")
    c = make_coeffs(torch, lmax, n, real, plan.cmax, SEED0 + L)
    f = plan.inverse(c)
    c2 = plan.forward(f)
    say(f"[plan] tuned {plan.info()}")
    for i in range(count):
        C = plan.sizes[i]
        fs = g.inverse(int(lmax[i]), int(n[i]), c[i, :C].contiguous(), real=real)
        assert rel(f[i], fs) <= 1e-13, i
        cs = g.forward(int(lmax[i]), int(n[i]), fs, real=real)
        assert rel(c2[i, :C], cs) <= 1e-13, i
        assert float((c2[i, C:]).abs().max()) == 0.0 if C < plan.cmax else True
    og = oracle.OracleGrid(L, 2, table=False)
    for i in (0, count - 1):
        C = plan.sizes[i]
        ref = og.inverse(int(lmax[i]), int(n[i]), c[i, :C].cpu().numpy(), real)
        assert rel(f[i].cpu(), torch.from_numpy(ref)) <= 1e-12
    og.close()


def test_plan_fp32_storage_and_graph_replays(torch):
    L = 128
    rng = np.random.default_rng(3)
    count = 12
    lmax = rng.integers(8, L + 1, count)
    n = rng.integers(-2, 3, count)
    g = G.GaussLegendreGrid(L, 2)
    p64 = G.BatchPlan(g, lmax, n)
    p32 = G.BatchPlan(g, lmax, n, storage="fp32")
    c = make_coeffs(torch, lmax, n, False, p64.cmax, SEED0 + 9)
    f64 = p64.inverse(c)
    c32 = c.to(torch.complex64)
    f32 = p32.inverse(c32)
    assert f32.dtype == torch.complex64
    assert rel(f32.to(torch.complex128), f64) < 1e-6
    a32 = p32.forward(f32)
    assert a32.dtype == torch.complex64
    assert rel(a32.to(torch.complex128), c) < 1e-5
    # Graph: capture on first use, replay with eager work and syncs between.
    fg = torch.empty_like(f32)
    eager = p32.inverse(c32).clone()
    ag = torch.zeros_like(c32)
    eager_a = p32.forward(eager).clone()
    for rep in range(4):
        fg.zero_()
        p32.inverse(c32, field=fg, graph=True)
        torch.cuda.synchronize()
        assert torch.equal(fg, eager), rep
        ag.zero_()
        p32.forward(fg, coeffs=ag, graph=True)
        torch.cuda.synchronize()
        assert torch.equal(ag, eager_a), rep
        _ = g.inverse(int(lmax[0]), int(n[0]), c[0, :p64.sizes[0]].contiguous())  # eager work
        torch.cuda.synchronize()


def test_config4_uneven_batch_round_trip(torch):
    """256 fields, lMax log-uniform in [32, 1024], n uniform in {-2..2},
    complex FP32 storage, coefficients (l+1)^-1: synthesis then analysis
    through one plan (graph replay), round trip within FP32 rounding, and
    five fields against the plain FP64 single-field calls."""
    L, count = 1024, 256
    rng = np.random.default_rng(SEED0 + 4)
    lmax = np.round(np.exp(rng.uniform(np.log(32), np.log(1024), count))).astype(int)
    n = rng.integers(-2, 3, count)
    g = G.GaussLegendreGrid(L, 2)
    plan = G.BatchPlan(g, lmax, n, storage="fp32")
    c = make_coeffs(torch, lmax, n, False, plan.cmax, SEED0 + 400, dtype=torch.complex64)
    f = plan.inverse(c, graph=True)
    f = plan.inverse(c, field=f, graph=True)
    c2 = plan.forward(f, graph=True)
    c2 = plan.forward(f, coeffs=c2, graph=True)
    torch.cuda.synchronize()
    say(f"[config4] groups {plan.info()}")
    err = max(rel(c2[i, :plan.sizes[i]], c[i, :plan.sizes[i]]) for i in range(count))
    say(f"[config4] worst round-trip rel L2 {err:.2e}")
    assert err < 2e-6
    for i in rng.choice(count, 5, replace=False):
        C = plan.sizes[i]
        ref = g.inverse(int(lmax[i]), int(n[i]), c[i, :C].to(torch.complex128).contiguous())
        assert rel(f[i].to(torch.complex128), ref) < 1e-6, i


def test_plan_graph_recaptured_after_buffer_growth(torch):
    """A captured plan graph embeds the grid's work buffer and the plan's
    FP32 staging buffer.  A larger eager call on the same grid reallocates
    the work buffer, and a larger coefficient stride on an FP32 plan
    reallocates the staging buffer; the plan must then capture again rather
    than replay a graph holding freed memory.  Every replay equals the eager
    result of the same call."""
    L = 128
    g = G.GaussLegendreGrid(L, 2)
    lmax = np.array([40, 33, 20, 12])
    n = np.array([0, 1, -1, 2])
    for storage in ("fp64", "fp32"):
        say(f"[plan graph, buffer growth] storage={storage}")
        plan = G.BatchPlan(g, lmax, n, storage=storage)
        cdt = plan.coeff_dtype()
        c = make_coeffs(torch, lmax, n, False, plan.cmax, SEED0 + 77, dtype=cdt)
        eager_f = plan.inverse(c).clone()
        eager_c = plan.forward(eager_f).clone()
        fg = torch.empty_like(eager_f)
        cg = torch.zeros_like(c)
        for rep in range(3):
            fg.zero_()
            plan.inverse(c, field=fg, graph=True)
            cg.zero_()
            plan.forward(fg, coeffs=cg, graph=True)
            torch.cuda.synchronize()
            assert torch.equal(fg, eager_f), (storage, rep)
            assert torch.equal(cg, eager_c), (storage, rep)
            # Eager work between replays that grows the grid's work buffer
            # (8, 24, then 40 complex fields at the grid's full degree).
            big = torch.from_numpy(G.random_coefficients(L, 0, seed=SEED0 + rep,
                                                         batch=8 + 16 * rep)).cuda()
            _ = g.forward(L, 0, g.inverse(L, 0, big))
            torch.cuda.synchronize()
        if storage == "fp32":
            # A wider coefficient stride reallocates the staging buffer; the
            # graph captured for the first stride is stale afterwards.
            wide = torch.zeros((plan.count, plan.cmax + 300), dtype=cdt, device="cuda")
            wide[:, :plan.cmax] = c
            fw = torch.empty_like(eager_f)
            for rep in range(2):
                fw.zero_()
                plan.inverse(wide, field=fw, graph=True)
                torch.cuda.synchronize()
                assert torch.equal(fw, eager_f), rep
                fg.zero_()
                plan.inverse(c, field=fg, graph=True)
                torch.cuda.synchronize()
                assert torch.equal(fg, eager_f), rep
        plan.close()


def test_plan_outlives_grid(torch):
    """Destroying a grid detaches its plans: later plan calls raise, and
    closing the plan afterwards is safe."""
    g = G.GaussLegendreGrid(32, 1)
    plan = G.BatchPlan(g, [32, 16], [0, 1])
    c = make_coeffs(torch, [32, 16], [0, 1], False, plan.cmax, SEED0 + 5)
    plan.inverse(c, graph=True)
    torch.cuda.synchronize()
    g.close()
    with pytest.raises(G.GSHError):
        plan.inverse(c)
    plan.close()


def test_tuned_choices_are_cached_and_persisted(torch, tmp_path, monkeypatch):
    """Tuned choices are kept per launch shape: with GSH_TUNE_CACHE naming a
    file, the first plan of a shape appends its choices there, a second plan
    of the same shape takes them without timing (the file does not grow) and
    reports the same choices, and its results equal the first plan's."""
    cache = tmp_path / "tune.txt"
    cache.write_text("not a tuning line\n")  # unreadable lines are skipped
    monkeypatch.setenv("GSH_TUNE_CACHE", str(cache))
    L = 96
    lmax = np.array([96, 80, 71, 40, 33, 12, 90, 64, 55] * 2)
    n = np.array([0, 1, -1, 2, 0, 0, 1, -2, 2] * 2)
    g = G.GaussLegendreGrid(L, 2)
    p1 = G.BatchPlan(g, lmax, n)
    c = make_coeffs(torch, lmax, n, False, p1.cmax, SEED0 + 500)
    f1 = p1.inverse(c)
    a1 = p1.forward(f1)
    torch.cuda.synchronize()
    lines = [x for x in cache.read_text().splitlines() if x.startswith("gsh2:")]
    say(f"[tune cache] {len(lines)} shapes written")
    assert len(lines) == 2 * len(p1.info())  # every group, both directions
    p2 = G.BatchPlan(g, lmax, n)
    f2 = p2.inverse(c)
    a2 = p2.forward(f2)
    torch.cuda.synchronize()
    assert cache.read_text().count("gsh2:") == len(lines)
    assert p2.info() == p1.info()
    assert torch.equal(f1, f2) and torch.equal(a1, a2)
    p1.close()
    p2.close()


def test_split_bf16_analysis_pass_choice(torch, tmp_path, monkeypatch):
    """The split-BF16 analysis runs in 64- or 128-degree passes, whichever the
    tuner timed faster per group.  For upper index 0 both pass lengths
    accumulate the same products in the same order, so the result does not
    depend on the choice; complex fields of other upper indices run orders m
    and -m together in 64-degree passes (the -m sums over mirrored
    colatitudes, chunked from the other pole), which agree with the
    128-degree kernel to the split-BF16 error.  A second cache file that pins
    every shape to the other pass length (read when GSH_TUNE_CACHE names a new
    file; its lines replace the process's choices) forces it: the n = 0
    field stays bit-identical, the others within 1e-7 of each other."""
    L = 512
    lmax = np.array([512, 400, 512, 300])
    n = np.array([0, 1, -1, 2])
    g = G.GaussLegendreGrid(L, 2)
    fp = G.BatchPlan(g, lmax, n)
    c = make_coeffs(torch, lmax, n, False, fp.cmax, SEED0 + 600)
    f = fp.inverse(c)
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("GSH_TUNE_CACHE", str(cache))
    p1 = G.BatchPlan(g, lmax, n, precision="bf16x6")
    a1 = p1.forward(f)
    torch.cuda.synchronize()
    passes = [d["analysis_pass"] for d in p1.info()]
    say(f"[pass] tuned analysis passes {passes}")
    assert all(p in (64, 128) for p in passes)
    lines = [x.split() for x in cache.read_text().splitlines() if x.startswith("gsh2:")]
    flipped = []
    for key, t, ff, pl in lines:
        if int(pl) in (64, 128):
            pl = "128" if pl == "64" else "64"
        flipped.append(" ".join([key, t, ff, pl]))
    cache2 = tmp_path / "tune2.txt"
    cache2.write_text("\n".join(flipped) + "\n")
    monkeypatch.setenv("GSH_TUNE_CACHE", str(cache2))
    p2 = G.BatchPlan(g, lmax, n, precision="bf16x6")
    a2 = p2.forward(f)
    torch.cuda.synchronize()
    passes2 = [d["analysis_pass"] for d in p2.info()]
    say(f"[pass] forced analysis passes {passes2}")
    assert all(a + b == 192 for a, b in zip(passes, passes2))
    assert torch.equal(a1[0], a2[0])  # n = 0
    d = float(torch.linalg.vector_norm(a1[1:] - a2[1:]) / torch.linalg.vector_norm(a1[1:]))
    say(f"[pass] n != 0: 64- vs 128-degree passes {d:.2e}")
    assert 0 < d < 1e-7
    fp64 = fp.forward(f)
    for a in (a1, a2):
        e = float(torch.linalg.vector_norm(a - fp64) / torch.linalg.vector_norm(fp64))
        say(f"[pass] bf16x6 vs FP64 {e:.2e}")
        assert e < 1e-7
    for p in (fp, p1, p2):
        p.close()
