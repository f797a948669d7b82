"""Summaries of rocprofv3 runs of bench.py: per-kernel average durations and HBM traffic.

Usage (after three profiling runs of the same bench.py command; each summary records them):

    python -m fftso3_amd.profiling STATS_DIR FETCH_DIR WRITE_DIR OUT_PREFIX [--copy-bytes N]
        [--mfma-dir DIR] [--stats-md FILE]

STATS_DIR holds a `rocprofv3 --kernel-trace --stats --output-format csv` run, FETCH_DIR and
WRITE_DIR `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` runs.  FETCH_SIZE and WRITE_SIZE
are read in KiB per dispatch and scaled by the factor that makes the largest device copy of the
run (bench.py --calibrate-copy: a copy of N bytes) read N bytes and write N bytes.  Writes
OUT_PREFIX.md (a short table) and OUT_PREFIX.json ({kernel: {avg_ms, fetch_bytes, write_bytes,
traffic_bytes}}).  With --mfma-dir (a `rocprofv3 --pmc SQ_INSTS_VALU_MFMA_MOPS_F32` run) each
kernel also gets its FP32 MFMA work per launch (the counter counts units of 512 flops) and the
rate that makes over the kernel's average duration.  --stats-md writes the --stats table (calls,
average / min / max duration, share of GPU time) as markdown.
"""
import argparse
import collections
import csv
import glob
import json
import os


def _csvs(d, pattern):
    return sorted(glob.glob(os.path.join(d, "**", pattern), recursive=True))


def kernel_stats(d):
    """{kernel name: (calls, average ns)} from a --stats run."""
    out = {}
    for f in _csvs(d, "*kernel_stats.csv"):
        for r in csv.DictReader(open(f)):
            out[r["Name"]] = (int(r["Calls"]), float(r["AverageNs"]))
    return out


def counter_per_dispatch(d, counter):
    """{kernel name: [value per dispatch]} of one counter from a --pmc run."""
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    for f in _csvs(d, "*counter_collection.csv"):
        for r in csv.DictReader(open(f)):
            if r["Counter_Name"] != counter:
                continue
            per[r["Kernel_Name"]][int(r["Dispatch_Id"])] += float(r["Counter_Value"])
    return {k: list(v.values()) for k, v in per.items()}


def kernel_stats_full(d):
    """Rows of a --stats run: (name, calls, avg ns, min ns, max ns, percent)."""
    rows = []
    for f in _csvs(d, "*kernel_stats.csv"):
        for r in csv.DictReader(open(f)):
            rows.append((r["Name"], int(r["Calls"]), float(r["AverageNs"]), float(r["MinNs"]), float(r["MaxNs"]),
                         float(r["Percentage"])))
    return sorted(rows, key=lambda x: -x[5])


def stats_markdown(d, command):
    lines = ["Command:", "", "```", command, "```", "",
             "| kernel | calls | avg ms | min ms | max ms | % of GPU time |", "|---|---|---|---|---|---|"]
    for n, c, a, lo, hi, pct in kernel_stats_full(d):
        lines.append("| %s | %d | %.3f | %.3f | %.3f | %.2f |" % (short(n), c, a * 1e-6, lo * 1e-6, hi * 1e-6, pct))
    return "\n".join(lines) + "\n"


def short(name):
    return name.split("(")[0].replace("void ", "")


def summarize(stats_dir, fetch_dir, write_dir, copy_bytes):
    stats = kernel_stats(stats_dir)
    fetch = counter_per_dispatch(fetch_dir, "FETCH_SIZE")
    write = counter_per_dispatch(write_dir, "WRITE_SIZE")
    copies_f = [v for k, vs in fetch.items() if "copyBuffer" in k for v in vs]
    copies_w = [v for k, vs in write.items() if "copyBuffer" in k for v in vs]
    if not copies_f or not copies_w:
        raise SystemExit("no device copy in the PMC runs: run bench.py with --calibrate-copy")
    f_fetch = copy_bytes / (max(copies_f) * 1024.0)
    f_write = copy_bytes / (max(copies_w) * 1024.0)
    table = {}
    for k in sorted(set(fetch) | set(write)):
        if "copyBuffer" in k:
            continue
        fv, wv = fetch.get(k, [0.0]), write.get(k, [0.0])
        fb = sum(fv) / len(fv) * 1024.0 * f_fetch
        wb = sum(wv) / len(wv) * 1024.0 * f_write
        calls, avg_ns = stats.get(k, (0, float("nan")))
        table[short(k)] = {"kernel": k, "calls": calls, "avg_ms": avg_ns * 1e-6, "fetch_bytes": fb,
                           "write_bytes": wb, "traffic_bytes": fb + wb}
    return {"fetch_scale": f_fetch, "write_scale": f_write, "copy_bytes": copy_bytes, "kernels": table}


def add_mfma(s, mfma_dir, counter="SQ_INSTS_VALU_MFMA_MOPS_F32"):
    """MFMA flops per launch (the MOPS counter counts units of 512 flops) and their rate."""
    mops = counter_per_dispatch(mfma_dir, counter)
    for k, vs in mops.items():
        v = s["kernels"].get(short(k))
        if v is None:
            continue
        v["mfma_f32_flops"] = sum(vs) / len(vs) * 512.0
        v["mfma_f32_tflops"] = v["mfma_f32_flops"] / (v["avg_ms"] * 1e-3) / 1e12


def to_markdown(s, command, mfma_label="FP32"):
    mf = any("mfma_f32_flops" in v for v in s["kernels"].values())
    lines = ["| kernel | calls | avg ms | HBM read (GB) | HBM write (GB) | traffic (GB) |" +
             (" %s MFMA GFLOP | MFMA TFLOP/s |" % mfma_label if mf else ""),
             "|---|---|---|---|---|---|" + ("---|---|" if mf else "")]
    for k, v in sorted(s["kernels"].items(), key=lambda kv: -(kv[1]["avg_ms"] * max(kv[1]["calls"], 1))):
        row = "| %s | %d | %.3f | %.3f | %.3f | %.3f |" % (
            k, v["calls"], v["avg_ms"], v["fetch_bytes"] / 1e9, v["write_bytes"] / 1e9, v["traffic_bytes"] / 1e9)
        if mf:
            row += " %.2f | %.1f |" % (v.get("mfma_f32_flops", 0.0) / 1e9, v.get("mfma_f32_tflops", 0.0))
        lines.append(row)
    head = ["Command:", "", "```", command, "```", "",
            "FETCH_SIZE scaled by %.3f and WRITE_SIZE by %.3f (a %d-byte device copy read and written)." % (
                s["fetch_scale"], s["write_scale"], s["copy_bytes"]), ""]
    return "\n".join(head + lines) + "\n"


def with_text(body, title, notes):
    """The summary with a title line above and explanatory notes below."""
    out = ("# " + title + "\n\n" if title else "") + body
    if notes:
        out += "\n" + notes.strip() + "\n"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("stats_dir")
    ap.add_argument("fetch_dir")
    ap.add_argument("write_dir")
    ap.add_argument("out_prefix")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--command", default="")
    ap.add_argument("--mfma-dir", default=None)
    ap.add_argument("--stats-md", default=None)
    ap.add_argument("--mfma-counter", default="SQ_INSTS_VALU_MFMA_MOPS_F32")
    ap.add_argument("--mfma-label", default="FP32")
    ap.add_argument("--title", default="", help="title line of the traffic summary")
    ap.add_argument("--stats-title", default="", help="title line of the --stats summary")
    ap.add_argument("--notes", default="", help="file whose text goes below the traffic table")
    ap.add_argument("--stats-notes", default="", help="file whose text goes below the --stats table")
    a = ap.parse_args()
    read = lambda path: open(path).read() if path else ""
    s = summarize(a.stats_dir, a.fetch_dir, a.write_dir, a.copy_bytes)
    if a.mfma_dir:
        add_mfma(s, a.mfma_dir, a.mfma_counter)
    if a.stats_md:
        with open(a.stats_md, "w") as f:
            f.write(with_text(stats_markdown(a.stats_dir, a.command), a.stats_title, read(a.stats_notes)))
    with open(a.out_prefix + ".json", "w") as f:
        json.dump(s, f, indent=1)
    md = with_text(to_markdown(s, a.command, a.mfma_label), a.title, read(a.notes))
    with open(a.out_prefix + ".md", "w") as f:
        f.write(md)
    print(md)


if __name__ == "__main__":
    main()
