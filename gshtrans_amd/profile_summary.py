"""Summaries of rocprofv3 output for the transform kernels (markdown to stdout).

    python -m gshtrans_amd.profile_summary stats DIR --command CMD
        kernel-trace statistics (DIR holds *kernel_stats.csv)
    python -m gshtrans_amd.profile_summary counters --stats DIR --fetch DIR
        --write DIR --copy-fetch DIR --copy-write DIR --sq DIR --command CMD
        per-dispatch HBM bytes (FETCH_SIZE / WRITE_SIZE scaled by what the same
        counter reads for a plain 1 GiB copy), MFMA and FP64 VALU rates of each
        kernel, against the MI355X peaks, plus a JSON block that bench.py reads
        for its roofline `traffic`.

Counter conventions used (per dispatch): FETCH_SIZE and WRITE_SIZE in KiB;
MFMA FLOPs = 512 * SQ_INSTS_VALU_MFMA_MOPS_{BF16,F32,F64}; FP64 VALU FLOPs
= 64 * (ADD_F64 + MUL_F64 + 2 * FMA_F64) instructions.
"""
import argparse
import collections
import csv
import glob
import json
import os
import re

HBM_GBPS = 8000.0
PEAK_TFLOPS = {"BF16": 2500.0, "F32": 157.3, "F64": 78.6}
COPY_BYTES = 1 << 30  # bench.py --copy-probe: 1 GiB read + 1 GiB written per dispatch


def short(name):
    n = re.sub(r"void |gsh::\(anonymous namespace\)::|gsh::", "", name)
    n = re.sub(r"\((LegParams|FFTParams)(, long)?\)$", "", n)
    return n


def _one(d, pattern):
    f = glob.glob(os.path.join(d, "**", pattern), recursive=True)
    if not f:
        raise SystemExit(f"no {pattern} under {d}")
    return f[0]


def read_stats(d):
    out = {}
    for r in csv.DictReader(open(_one(d, "*kernel_stats.csv"))):
        out[short(r["Name"])] = {"calls": int(r["Calls"]), "avg_ns": float(r["AverageNs"]),
                                 "total_ns": float(r["TotalDurationNs"]),
                                 "pct": float(r["Percentage"])}
    return out


def read_counters(d):
    """{kernel: {counter: mean per dispatch}} (one row per dispatch and counter)."""
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(_one(d, "*counter_collection.csv"))):
        acc[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in acc.items()}


def copy_factor(d, counter):
    """Bytes per KiB of counter for the plain copy (largest dispatch of the run)."""
    vals = [v[counter] for v in read_counters(d).values() if counter in v]
    return COPY_BYTES / (max(vals) * 1024.0)


def cmd_stats(a):
    st = read_stats(a.dir)
    print(f"Command (from the repository root, on one MI355X):\n\n```\n{a.command}\n```\n")
    print("| kernel | calls | avg ms | total % |\n|---|---:|---:|---:|")
    for k, v in sorted(st.items(), key=lambda kv: -kv[1]["total_ns"])[: a.top]:
        print(f"| `{k}` | {v['calls']} | {v['avg_ns'] / 1e6:.3f} | {v['pct']:.2f} |")


def cmd_counters(a):
    st = read_stats(a.stats)
    fe, wr, sq = read_counters(a.fetch), read_counters(a.write), read_counters(a.sq)
    ff, wf = copy_factor(a.copy_fetch, "FETCH_SIZE"), copy_factor(a.copy_write, "WRITE_SIZE")
    print(f"Commands (from the repository root, on one MI355X; each counter set in a run "
          f"of its own):\n\n```\n{a.command}\n```\n")
    print(f"Calibration: the plain 1 GiB copy (`bench.py --copy-probe`) reads "
          f"FETCH_SIZE x {ff:.3f} and WRITE_SIZE x {wf:.3f} bytes per KiB counted; every "
          f"byte figure below is scaled by these factors.\n")
    print("| kernel | avg ms | HBM read GB | HBM write GB | HBM GB/s | of 8 TB/s | MFMA TFLOP/s "
          "| of peak | FP64 VALU TFLOP/s | of 78.6 |\n|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|")
    table = {}
    for k, v in sorted(st.items(), key=lambda kv: -kv[1]["total_ns"]):
        if k not in fe or k not in wr or not k.startswith(("leg_", "fft_", "coef_split",
                                                           "bin_split", "rec_check")):
            continue
        sec = v["avg_ns"] / 1e9
        rb = fe[k]["FETCH_SIZE"] * 1024.0 * ff
        wb = wr[k]["WRITE_SIZE"] * 1024.0 * wf
        q = sq.get(k, {})
        mfma = {t: 512.0 * q.get(f"SQ_INSTS_VALU_MFMA_MOPS_{t}", 0.0) for t in PEAK_TFLOPS}
        mt = max(mfma, key=mfma.get)
        mflops = mfma[mt]
        f64 = 64.0 * (q.get("SQ_INSTS_VALU_ADD_F64", 0.0) + q.get("SQ_INSTS_VALU_MUL_F64", 0.0) +
                      2.0 * q.get("SQ_INSTS_VALU_FMA_F64", 0.0))
        gbs = (rb + wb) / sec / 1e9
        row = {"avg_ns": v["avg_ns"], "fetch_bytes": rb, "write_bytes": wb,
               "hbm_gbps": gbs, "mfma_type": mt if mflops else None,
               "mfma_tflops": mflops / sec / 1e12, "fp64_valu_tflops": f64 / sec / 1e12,
               "mfma_insts": q.get("SQ_INSTS_MFMA")}
        table[k] = row
        mf = (f"{row['mfma_tflops']:.1f} ({mt}) | {row['mfma_tflops'] / PEAK_TFLOPS[mt]:.1%}"
              if mflops else "- | -")
        print(f"| `{k}` | {v['avg_ns'] / 1e6:.3f} | {rb / 1e9:.2f} | {wb / 1e9:.2f} | {gbs:.0f} "
              f"| {gbs / HBM_GBPS:.1%} | {mf} | {row['fp64_valu_tflops']:.2f} "
              f"| {row['fp64_valu_tflops'] / PEAK_TFLOPS['F64']:.1%} |")
    print("\nMachine-readable (per dispatch, bytes already scaled):\n\n```json")
    print(json.dumps({"fetch_factor": ff, "write_factor": wf, "kernels": table}, indent=1))
    print("```")


def load_table(path):
    """The JSON block of a `counters` summary, or None."""
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"```json\n(.*?)\n```", text, re.S)
    return json.loads(m.group(1)) if m else None


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--command", required=True)
    s.add_argument("--top", type=int, default=40)
    c = sub.add_parser("counters")
    for k in ("stats", "fetch", "write", "copy-fetch", "copy-write", "sq", "command"):
        c.add_argument("--" + k, required=True)
    a = ap.parse_args()
    cmd_stats(a) if a.cmd == "stats" else cmd_counters(a)


if __name__ == "__main__":
    main()
