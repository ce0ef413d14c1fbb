#!/usr/bin/env python3
"""GSH transform throughput on one MI355X.

Workload (the precision configuration the headline metric is quoted on):
GaussLegendreGrid lMax=1024, nMax=2 (1025 x 2048 grid); a batch of 64
real-valued n=0 fields and a batch of 64 complex n=2 fields, coefficients
normal with (l+1)^-1 decay (seeded, synthetic).  One step = synthesis then
analysis of every field = 256 transforms.  Inputs are resident in HBM before
timing starts.

Printed: one JSON line.  `value` is transforms/s of the headline mode, the
bf16x6 split-BF16 Legendre contraction (<= 1e-7 relative L2 against FP64);
the FP64 figure, the speedup, the bf16x3 and FP32-MFMA modes, and the
per-field relative L2 error of each mode against FP64 are reported beside it.
`roofline` is the dominant kernel's algorithmic rate (HIP events around each
launch on its own stream) against the MI355X peak of the unit it runs on (for
split-BF16 the MFMA flops of all products against the bf16 peak), with its
per-launch HBM traffic from the committed counter summary under profiles/;
`cpu_baseline` times the CPU oracle (a restatement of the reference's
serial transform with its precomputed Wigner table) on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GSH transforms/sec at lMax=1024, B=64: mixed vs FP64 GPU; rel. L2 err"
L = 1024
NMAX = 2
BATCH = 64
SEED = 20261202 + 2
# Published MI355X dense peaks of the MFMA type each mode runs on, and the
# MFMA products each mode spends per logical product (split-BF16 planes).
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0, "bf16x6": 2500.0}
PRODUCTS = {"fp64": 1, "fp32": 1, "bf16x3": 3, "bf16x6": 6}
HBM_GBPS = 8000.0
# Per-dispatch HBM bytes of each kernel, from FETCH_SIZE / WRITE_SIZE counter
# runs scaled by a plain copy's reading (rocprofv3 summary, text).
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "transform_kernels_lmax1024_b64_counters.md")
# Kernel instantiations of each work family per mode (real and complex
# batch), as regular expressions over the summary's kernel names.
FAMILY_KERNELS = {
    # split-BF16 Legendre kernels: the n=0 batch runs the equatorially
    # symmetric variants (_ws, _s), the n=2 batch orders m and -m paired (_pm)
    "legendre_analysis": {"fp64": r"leg_anal_f64<4, false>", "fp32": r"leg_anal_f64<4, true>",
                          "bf16x3": r"leg_anal_bf16_(ws<2, \d>|pm<2>)",
                          "bf16x6": r"leg_anal_bf16_(ws<3, \d>|pm<3>)"},
    "legendre_synthesis": {"fp64": r"leg_synth_f64<2, \d, false>",
                           "fp32": r"leg_synth_f64<2, \d, true>",
                           "bf16x3": r"leg_synth_bf16_(s<2, \d>|pm<2>)",
                           "bf16x6": r"leg_synth_bf16_(s<3, \d>|pm<3>)"},
    # FA in FP64 (fp64, fp32) or FP32 (split-BF16 modes): last template argument
    "fft_analysis": {"fp64": r"fft_anal_w<32, \d, 0, 0>", "fp32": r"fft_anal_w<32, \d, 0, 0>",
                     "bf16x3": r"fft_anal_w<32, \d, 0, 1>", "bf16x6": r"fft_anal_w<32, \d, 0, 1>"},
    # F in FP64 (fp64, fp32) or FP32 (split-BF16 modes): last template argument
    "fft_synthesis": {"fp64": r"fft_synth_w<32, \d, 0, 0>", "fp32": r"fft_synth_w<32, \d, 0, 0>",
                      "bf16x3": r"fft_synth_w<32, \d, 0, 1>", "bf16x6": r"fft_synth_w<32, \d, 0, 1>"},
    "coef_split": {"bf16x3": r"coef_split<2, \d, \d>", "bf16x6": r"coef_split<3, \d, \d>"},
}
HEADLINE = "bf16x6"            # <= 1e-7 relative L2 against FP64
REPORTED = ["bf16x3", "fp32"]  # reported beside it (throughput and errors)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_counts():
    """Per-field algorithmic work: Legendre flops 4*ntheta*C (complex) or
    4*ntheta*R (real); ring-FFT HBM bytes (field read/write + F bins).

    Also the Legendre flops the kernels issue: the n=0 kernels fold each
    colatitude with its mirror (d^l_{m0}(pi - theta) = (-1)^(l+m) d^l_{m0}(theta)
    on the equatorially symmetric Gauss-Legendre nodes), so they multiply over
    the ceil(ntheta/2) northern colatitudes only."""
    import gshtrans_amd as G
    nt, N = L + 1, 2 * L
    R0 = G.coefficient_size(L, 0, real=True)
    C2 = G.coefficient_size(L, 2, real=False)
    leg = {"real": 4.0 * nt * R0, "complex": 4.0 * nt * C2}
    issued = {"real": 4.0 * ((nt + 1) // 2) * R0, "complex": 4.0 * nt * C2}
    fft_bytes = {"real": nt * N * 8 + nt * (L + 1) * 16,
                 "complex": nt * N * 16 + nt * (2 * L + 1) * 16}
    return leg, fft_bytes, issued


def run_gpu(args):
    import torch
    import gshtrans_amd as G

    dev = torch.device("cuda:0")
    grid = G.GaussLegendreGrid(L, NMAX)
    grid.prepare(0)
    grid.prepare(2)
    log("[bench] generating inputs")
    cr = torch.from_numpy(G.random_coefficients(L, 0, real=True, seed=SEED, batch=BATCH)).to(dev)
    cc = torch.from_numpy(G.random_coefficients(L, 2, real=False, seed=SEED + 1,
                                                batch=BATCH)).to(dev)
    fr = torch.empty((BATCH, L + 1, 2 * L), dtype=torch.float64, device=dev)
    fc = torch.empty((BATCH, L + 1, 2 * L), dtype=torch.complex128, device=dev)
    outs = {}

    def step(prec, o):
        grid.inverse(L, 0, cr, field=o["fr"], real=True, precision=prec)
        grid.forward(L, 0, o["fr"], coeffs=o["cr"], real=True, precision=prec)
        grid.inverse(L, 2, cc, field=o["fc"], real=False, precision=prec)
        grid.forward(L, 2, o["fc"], coeffs=o["cc"], real=False, precision=prec)

    def buffers(prec):
        if prec == "fp64":
            return {"fr": fr, "fc": fc, "cr": torch.empty_like(cr), "cc": torch.empty_like(cc)}
        return {"fr": torch.empty_like(fr), "fc": torch.empty_like(fc),
                "cr": torch.empty_like(cr), "cc": torch.empty_like(cc)}

    # FP64, the headline mixed mode, then the other reduced-precision modes
    # (each reported with its own throughput and errors).
    headline = args.precision or HEADLINE
    extra = REPORTED if args.modes is None else [m for m in args.modes.split(",") if m]
    modes = ["fp64"] + [m for m in [headline] + extra if m != "fp64"]
    modes = list(dict.fromkeys(modes))
    for m in modes:
        outs[m] = buffers(m)
    mixed = headline if headline != "fp64" else None
    results = {}
    for m in modes:
        o = outs[m]
        for _ in range(args.warmup):
            step(m, o)
        torch.cuda.synchronize()
        if m == headline:
            grid.timing_enable(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(m, o)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        kt = grid.timing_read() if m == headline else None
        grid.timing_enable(False)
        tps = args.steps * 4 * BATCH / dt
        results[m] = {"transforms_per_s": tps, "ms_per_step": 1e3 * dt / args.steps,
                      "kernels": kt}
        log(f"[bench] {m}: {tps:.1f} transforms/s, {1e3 * dt / args.steps:.2f} ms/step")

    # Two streams (unless --no-two-streams): the real and complex batches are
    # independent jobs; the same step with each batch on its own HIP stream
    # (default priority) and grid handle (its own work buffers), so one
    # batch's HBM-bound ring FFTs overlap the other's Legendre kernels.
    # Reported beside the one-stream figures, which are value and carry the
    # per-kernel timings (profiles run with --no-two-streams, so that their
    # kernel durations are one-stream ones too).
    if args.two_streams:
        gc = G.GaussLegendreGrid(L, NMAX)
        gc.prepare(2)
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()

    def step2(prec, o):
        cur = torch.cuda.current_stream()
        s1.wait_stream(cur)
        s2.wait_stream(cur)
        grid.inverse(L, 0, cr, field=o["fr"], real=True, precision=prec, stream=s1)
        grid.forward(L, 0, o["fr"], coeffs=o["cr"], real=True, precision=prec, stream=s1)
        gc.inverse(L, 2, cc, field=o["fc"], real=False, precision=prec, stream=s2)
        gc.forward(L, 2, o["fc"], coeffs=o["cc"], real=False, precision=prec, stream=s2)
        cur.wait_stream(s1)
        cur.wait_stream(s2)

    two = {}
    for m in (dict.fromkeys(["fp64", headline]) if args.two_streams else []):
        o = outs[m]
        for _ in range(args.warmup):
            step2(m, o)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step2(m, o)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        two[m] = {"transforms_per_s": args.steps * 4 * BATCH / dt,
                  "ms_per_step": 1e3 * dt / args.steps}
        log(f"[bench] two streams, {m}: {two[m]['transforms_per_s']:.1f} transforms/s")
    if two and headline in two and "fp64" in two:
        two["speedup_vs_fp64"] = two[headline]["transforms_per_s"] / two["fp64"]["transforms_per_s"]
    results["two_streams"] = two or None

    # Accuracy of each mode against FP64 per field (relative L2).
    def per_field(a, b):
        d = (a - b).reshape(a.shape[0], -1)
        return (torch.linalg.vector_norm(d, dim=1) /
                torch.linalg.vector_norm(b.reshape(b.shape[0], -1), dim=1)).cpu().numpy()

    errors = {}
    ref = outs["fp64"]
    for m in modes:
        if m == "fp64":
            continue
        o = outs[m]
        # (the two-stream steps wrote the same values into these buffers)
        # Analysis on the same (FP64-synthesised) fields as the FP64 analysis.
        ar = torch.empty_like(cr)
        ac = torch.empty_like(cc)
        grid.forward(L, 0, ref["fr"], coeffs=ar, real=True, precision=m)
        grid.forward(L, 2, ref["fc"], coeffs=ac, real=False, precision=m)
        torch.cuda.synchronize()
        e = {"synthesis_real": per_field(o["fr"], ref["fr"]),
             "synthesis_complex": per_field(o["fc"], ref["fc"]),
             "analysis_real": per_field(ar, ref["cr"]),
             "analysis_complex": per_field(ac, ref["cc"])}
        allv = sorted(float(x) for v in e.values() for x in v)
        errors[m] = {k + "_max": float(v.max()) for k, v in e.items()}
        errors[m]["max"] = allv[-1]
        errors[m]["median"] = allv[len(allv) // 2]
        errors[m]["round_trip_max"] = float(max(per_field(o["cr"], cr).max(),
                                                per_field(o["cc"], cc).max()))
    # FP64 round trip as a size-independent self-check.
    rt_r = per_field(outs["fp64"]["cr"], cr)
    rt_c = per_field(outs["fp64"]["cc"], cc)
    errors["fp64_round_trip_max"] = float(max(rt_r.max(), rt_c.max()))
    return results, headline, mixed, errors


def roofline(results, headline):
    leg, fftb, issued = algorithmic_counts()
    kt = results[headline]["kernels"]
    # per-step work: real batch + complex batch, one launch each per kernel
    work = {
        "legendre_synthesis": ("mfma", BATCH * (leg["real"] + leg["complex"])),
        "legendre_analysis": ("mfma", BATCH * (leg["real"] + leg["complex"])),
        "fft_synthesis": ("hbm", BATCH * (fftb["real"] + fftb["complex"])),
        "fft_analysis": ("hbm", BATCH * (fftb["real"] + fftb["complex"])),
    }
    name = max(work, key=lambda k: kt[k][0])
    ms, launches = kt[name]
    bound, per_step = work[name]
    steps = launches / 2.0
    total = per_step * steps
    sec = ms / 1e3
    issued_frac = None
    if bound == "mfma":
        # MFMA flops for the algorithmic products: each logical product costs
        # PRODUCTS[mode] MFMA products.
        achieved = total * PRODUCTS[headline] / sec / 1e12
        peak = PEAK_TFLOPS[headline]
        unit = "TFLOP/s"
        # the same over the products the kernels issue (n=0: northern half)
        done = BATCH * (issued["real"] + issued["complex"]) * steps
        issued_frac = done * PRODUCTS[headline] / sec / 1e12 / peak
    else:
        achieved = total / sec / 1e9
        peak = HBM_GBPS
        unit = "GB/s"
    per_kernel = {k: {"avg_launch_ms": v[0] / max(v[1], 1), "launches": v[1]}
                  for k, v in kt.items()}
    return {"kernel": name, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "issued_frac": issued_frac,
            "traffic": traffic_per_launch(name, headline),
            "avg_launch_ms": ms / max(launches, 1), "per_kernel": per_kernel}


def traffic_per_launch(family, mode):
    """HBM bytes per launch of a work family (mean over its real and complex
    instantiations) from the committed counter summary, or None."""
    from gshtrans_amd.profile_summary import load_table
    t = load_table(TRAFFIC_FILE)
    if not t:
        return None
    pat = FAMILY_KERNELS.get(family)
    pat = pat.get(mode) if isinstance(pat, dict) else pat
    if pat is None:
        return None
    rows = [v for k, v in t["kernels"].items() if re.fullmatch(pat, k)]
    if not rows:
        return None
    return sum(r["fetch_bytes"] + r["write_bytes"] for r in rows) / len(rows)


def host_info():
    """The host the CPU baseline ran on: logical CPUs, the CPUs this process
    may use, and the CPU model (from /proc/cpuinfo)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "model": model}


def plan_vs_loop(args):
    """The batching workload: 256 complex fields, lMax drawn log-uniform from
    [32, 1024], n uniform in -2..2, coefficients normal with (l+1)^-1 decay
    (seeded), FP32 storage, FP64 arithmetic: one batch plan (every group's
    synthesis, then every group's analysis, replayed from a HIP graph)
    against a loop of the plain per-field calls (gsh_inverse / gsh_forward,
    FP64 storage) over the same fields.  512 transforms per step."""
    import numpy as np
    import torch
    import gshtrans_amd as G
    rng = np.random.default_rng(SEED + 3)
    count = 256
    lmax = np.rint(np.exp(rng.uniform(np.log(32), np.log(1024), count))).astype(int)
    lmax = np.clip(lmax, 32, 1024)
    nn = rng.integers(-2, 3, count)
    grid = G.GaussLegendreGrid(L, NMAX)
    for n in range(-NMAX, NMAX + 1):
        grid.prepare(n)
    sizes = [G.coefficient_size(int(a), int(b)) for a, b in zip(lmax, nn)]
    cmax = max(sizes)
    c64 = torch.zeros((count, cmax), dtype=torch.complex128)
    for i in range(count):
        c64[i, :sizes[i]] = torch.from_numpy(G.random_coefficients(int(lmax[i]), int(nn[i]),
                                                                   seed=SEED + 10 + i,
                                                                   grid_lmax=L))
    c64 = c64.cuda()
    plan = G.BatchPlan(grid, lmax, nn, real=False, precision="fp64", storage="fp32")
    c32 = c64.to(torch.complex64)
    f32 = torch.empty((count, L + 1, 2 * L), dtype=torch.complex64, device="cuda")
    o32 = torch.empty_like(c32)

    def plan_step():
        plan.inverse(c32, field=f32, graph=True)
        plan.forward(f32, coeffs=o32, graph=True)

    fbuf = torch.empty((1, L + 1, 2 * L), dtype=torch.complex128, device="cuda")
    obuf = [torch.empty((1, s), dtype=torch.complex128, device="cuda") for s in sizes]
    cin = [c64[i:i + 1, :sizes[i]].contiguous() for i in range(count)]

    def loop_step():
        for i in range(count):
            grid.inverse(int(lmax[i]), int(nn[i]), cin[i], field=fbuf)
            grid.forward(int(lmax[i]), int(nn[i]), fbuf, coeffs=obuf[i])

    out = {}
    for name, fn in (("plan_graph", plan_step), ("loop", loop_step)):
        for _ in range(2):  # tuning, graph capture, warm-up
            fn()
        torch.cuda.synchronize()
        reps = max(2, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        out[name] = {"transforms_per_s": 2 * count / dt, "ms_per_step": 1e3 * dt}
        log(f"[bench] batching workload, {name}: {2 * count / dt:.1f} transforms/s")
    # The plan's FP32-stored results against the FP64 loop.
    err = []
    for i in range(count):
        a = o32[i, :sizes[i]].to(torch.complex128)
        b = obuf[i][0]
        err.append(float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)))
    info = plan.info()
    plan.close()
    return {"workload": "256 complex fields, lMax log-uniform [32, 1024], n in -2..2, FP32 "
                        "storage, FP64 arithmetic; synthesis + analysis of every field "
                        "(512 transforms/step)",
            "plan_graph": out["plan_graph"], "loop_per_field": out["loop"],
            "speedup_plan_vs_loop": out["plan_graph"]["transforms_per_s"] /
            out["loop"]["transforms_per_s"],
            "plan_vs_loop_rel_l2_max": max(err), "groups": len(info)}


def cpu_baseline(args):
    """The CPU oracle restating the reference's path: Wigner table built with
    OpenMP on all host cores (setup, reported separately), then serial
    per-field transforms streaming the table, as GaussLegendreGrid.h:163, 281.
    Sample: one real n=0 field and one complex n=2 field, synthesis and
    analysis each (4 transforms), on NRange=Single grids (8.6 GB table each)."""
    import numpy as np
    import gshtrans_amd as G
    import oracle
    t0 = time.perf_counter()
    g0 = oracle.OracleGrid(L, 0, nrange="Single", table=True)
    g2 = oracle.OracleGrid(L, 2, nrange="Single", table=True)
    setup = time.perf_counter() - t0
    cr = G.random_coefficients(L, 0, real=True, seed=SEED)
    cc = G.random_coefficients(L, 2, real=False, seed=SEED + 1)
    t0 = time.perf_counter()
    f = g0.inverse(L, 0, cr, real=True)
    g0.forward(L, 0, f, real=True)
    f = g2.inverse(L, 2, cc)
    g2.forward(L, 2, f)
    dt = time.perf_counter() - t0
    g0.close()
    g2.close()
    return {"value": 4.0 / dt, "unit": "transforms/s", "cores": 1, "kind": "port",
            "sample": "lMax=1024: 1 real n=0 + 1 complex n=2 field, synthesis + analysis "
                      "(4 transforms, serial as the reference) with precomputed Wigner tables",
            "setup_s": setup, "setup_threads": oracle.lib().gsho_num_threads(),
            "host": host_info()}


def config_rows(args):
    """The CPU-baseline rows of the other workloads, each beside the FP64 GPU
    figure on the same seeded inputs (one JSON line of its own; not the
    headline).  CPU: the oracle with its precomputed Wigner table (built with
    OpenMP on every usable core, reported as setup_s), serial transforms as
    GaussLegendreGrid.h:163, 281; median of the repetitions.
      grid64    lMax=64, n=0, one complex field, (l+1)^-1 decay
      grid512   lMax=512, n=-2..2, one complex field per n, (l+1)^-1.5 decay
      lsq512    lMax=512, n=1, 16 complex fields: the GPU adjoint of the
                synthesis (S^H, the least-squares gradient's transform)
                against the oracle's analysis, its CPU counterpart (the
                same contraction with quadrature weights)."""
    import statistics
    import numpy as np
    import torch
    import gshtrans_amd as G
    import oracle

    def med(fn, reps):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts)

    def gpu_med(fn, reps):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    rows = {}
    for name, Lc, ns, decay in (("grid64", 64, [0], 1.0), ("grid512", 512, [-2, -1, 0, 1, 2], 1.5)):
        t0 = time.perf_counter()
        og = oracle.OracleGrid(Lc, 2, table=True)
        setup = time.perf_counter() - t0
        g = G.GaussLegendreGrid(Lc, 2)
        per_n = {}
        for n in ns:
            c = G.random_coefficients(Lc, n, seed=SEED + 100 + n, decay=decay)
            reps = 5 if Lc <= 64 else 3
            f_cpu = og.inverse(Lc, n, c)
            a_cpu = og.forward(Lc, n, f_cpu)
            ts = med(lambda: og.inverse(Lc, n, c), reps)
            ta = med(lambda: og.forward(Lc, n, f_cpu), reps)
            cd = torch.from_numpy(c).cuda()
            fd = g.inverse(Lc, n, cd)
            ad = g.forward(Lc, n, fd)
            gs = gpu_med(lambda: g.inverse(Lc, n, cd, field=fd), 10)
            ga = gpu_med(lambda: g.forward(Lc, n, fd, coeffs=ad), 10)
            per_n[str(n)] = {
                "cpu_synthesis_s": ts, "cpu_analysis_s": ta,
                "gpu_synthesis_s": gs, "gpu_analysis_s": ga,
                "rel_l2_synthesis_vs_cpu": rel(fd.cpu().numpy(), f_cpu),
                "rel_l2_analysis_vs_cpu": rel(ad.cpu().numpy(), a_cpu),
                "round_trip_rel_l2": rel(ad.cpu().numpy(), c)}
        rows[name] = {"lmax": Lc, "setup_s": setup, "fields": per_n}
        og.close()
        g.close()
        log(f"[bench] {name}: setup {setup:.2f} s")
    Lc, n, B = 512, 1, 16
    og = oracle.OracleGrid(Lc, 1, nrange="Single", table=True)
    g = G.GaussLegendreGrid(Lc, 1)
    c = G.random_coefficients(Lc, n, seed=SEED + 200, batch=B)
    f_cpu = og.inverse(Lc, n, c[0])
    ta = med(lambda: og.forward(Lc, n, f_cpu), 3)
    fd = g.inverse(Lc, n, torch.from_numpy(c).cuda())
    adj = g.adjoint(Lc, n, fd)
    gt = gpu_med(lambda: g.adjoint(Lc, n, fd, coeffs=adj), 10)
    rows["lsq512"] = {"lmax": Lc, "n": n, "batch": B, "cpu_analysis_s_per_field": ta,
                      "gpu_adjoint_s_per_field": gt / B,
                      "gpu_adjoint_batch_s": gt}
    og.close()
    g.close()
    return {"config_rows": rows, "host": host_info(),
            "setup_threads": oracle.lib().gsho_num_threads()}


def copy_probe():
    """A plain copy kernel of known size: 1 GiB read + 1 GiB written per
    dispatch, the calibration for FETCH_SIZE / WRITE_SIZE counter runs."""
    import torch
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_per_dispatch": a.numel() * 8}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default=None,
                    help="headline mode (default bf16x6)")
    ap.add_argument("--modes", default=None,
                    help="comma-separated reduced-precision modes reported beside the "
                         "headline (default bf16x3,fp32)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batching", action="store_true",
                    help="skip the batching workload (plan graph replay against a loop of "
                         "per-field calls)")
    ap.add_argument("--no-two-streams", dest="two_streams", action="store_false",
                    help="skip the extra timing of FP64 and the headline mode with the real "
                         "and complex batches on two HIP streams (reported beside value, "
                         "never as value)")
    ap.add_argument("--config-rows", action="store_true",
                    help="only time the CPU-baseline rows of the other workloads (lMax=64, "
                         "lMax=512 for n=-2..2, the lMax=512 adjoint) beside the FP64 GPU "
                         "figures, print them as one JSON line and exit")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy 5 times (the "
                         "known-size reference for HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.config_rows:
        print(json.dumps(config_rows(args)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(the transform batch shards over GPUs without exchange: one GPU's rate is its share)",
              file=sys.stderr)
        sys.exit(2)
    results, headline, mixed, errors = run_gpu(args)
    rl = roofline(results, headline)
    batching = None if args.no_batching else plan_vs_loop(args)
    cpu = None if args.no_cpu_baseline else cpu_baseline(args)
    head = results[headline]
    fp64 = results["fp64"]
    line = {
        "metric": METRIC,
        "value": head["transforms_per_s"],
        "unit": "transforms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if headline == "fp64" else f"{headline} Legendre products, f64 recurrence/FFT",
        "data": "synthetic (seeded normal coefficients, (l+1)^-1 decay)",
        "config": {"workload": "GaussLegendreGrid lMax=1024 nMax=2: 64 real n=0 + 64 complex "
                               "n=2 fields, synthesis then analysis (256 transforms/step)",
                   "lmax": L, "batch_real": BATCH, "batch_complex": BATCH,
                   "mode": headline},
        "fp64_transforms_per_s": fp64["transforms_per_s"],
        "mixed_transforms_per_s": results[mixed]["transforms_per_s"] if mixed else None,
        "speedup_mixed_vs_fp64": (results[mixed]["transforms_per_s"] /
                                  fp64["transforms_per_s"]) if mixed else None,
        "modes": {m: {"transforms_per_s": r["transforms_per_s"], "ms_per_step": r["ms_per_step"],
                      "speedup_vs_fp64": r["transforms_per_s"] / fp64["transforms_per_s"]}
                  for m, r in results.items() if m != "two_streams"},
        "two_streams": dict(results["two_streams"], note=(
            "the real and complex batches of each step on two HIP streams, each with its "
            "own grid handle; value and the per-kernel figures are the one-stream run"))
        if results["two_streams"] else None,
        "rel_l2": errors,
        "batching": batching,
        "roofline": rl,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
