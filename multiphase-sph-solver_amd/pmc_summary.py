"""Summarises rocprofv3 runs of bench.py into a per-kernel table and profiles/pmc_<workload>.json.

    python multiphase-sph-solver_amd/pmc_summary.py <kt_dir> <fetch_dir> <write_dir> <cfetch_dir> <cwrite_dir> \
        <workload> <out.md> <out.json>
    python multiphase-sph-solver_amd/pmc_summary.py --counters <out.md> <pmc_dir> [<pmc_dir> ...]

Inputs are the output directories of: rocprofv3 --kernel-trace --stats (bench), --pmc FETCH_SIZE
(bench), --pmc WRITE_SIZE (bench), and the same two PMC passes over `bench.py --copy-calibration`
(three device copies of 1 GiB), whose counts give the bytes-per-count scale of each counter.
"""
import collections
import csv
import glob
import json
import os
import sys


def _csv(d, suffix):
    f = glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True)
    return list(csv.DictReader(open(f[0]))) if f else []


def counters(d):
    agg = collections.defaultdict(list)
    meta = {}
    for r in _csv(d, "counter_collection.csv"):
        agg[r["Kernel_Name"]].append(float(r["Counter_Value"]))
        meta[r["Kernel_Name"]] = (int(r["VGPR_Count"]), int(r["Accum_VGPR_Count"]), int(r["Scratch_Size"]))
    return agg, meta


def short(name):
    """Pass name as the bench reports it; the neighbour batch width (third template argument of
    the tunable passes) is kept as a suffix, e.g. strain/w2, and so is the search's list staging
    depth, e.g. search/s16."""
    s = name.split("(")[0].replace("void ", "").replace("sphg::", "")
    base = s.split("<")[0].replace("k_", "")
    if "<" in s:
        args = [a.strip() for a in s[s.index("<") + 1:s.rindex(">")].split(",")]
        if base in ("density", "norm", "strain", "xsph") and len(args) == 3:
            return f"{base}/w{args[2]}"
        if base == "search" and len(args) == 2:
            return f"{base}/s{args[1]}"
    return base


def main(kt, fe, wr, cfe, cwr, workload, out_md, out_json):
    stats = _csv(kt, "kernel_stats.csv")
    f, meta = counters(fe)
    w, _ = counters(wr)
    cf, _ = counters(cfe)
    cw, _ = counters(cwr)
    gib_kib = 1024.0 * 1024.0
    copy_name = [k for k in cf if "copyBuffer" in k][0]
    fscale = gib_kib / (sum(cf[copy_name]) / len(cf[copy_name]))
    wscale = gib_kib / (sum(cw[copy_name]) / len(cw[copy_name]))
    rows = []
    js = {"workload": workload, "fetch_scale": fscale, "write_scale": wscale, "kernels": {}}
    for r in stats:
        name = r["Name"]
        if "sphg" not in name:
            continue
        k = short(name)
        kname = "norm_eos" + k[4:] if k.startswith("norm") else k
        fk = f.get(name, [0.0])
        wk = w.get(name, [0.0])
        fb = sum(fk) / len(fk) * 1024 * fscale
        wb = sum(wk) / len(wk) * 1024 * wscale
        vg = meta.get(name, (0, 0, 0))
        rows.append((kname, int(r["Calls"]), float(r["AverageNs"]) / 1e3, float(r["Percentage"]), fb, wb, vg))
        js["kernels"][kname] = {"avg_us": float(r["AverageNs"]) / 1e3, "fetch_bytes_per_launch": fb,
                                "write_bytes_per_launch": wb, "hbm_bytes_per_launch": fb + wb,
                                "vgpr": vg[0], "agpr": vg[1], "scratch": vg[2]}
    with open(out_json, "w") as fo:
        json.dump(js, fo, indent=1)
    with open(out_md, "a") as fo:
        fo.write(f"\nFETCH_SIZE scale {fscale:.4f}, WRITE_SIZE scale {wscale:.4f} "
                 f"(1 GiB device copy read back as {gib_kib / fscale:.0f} / {gib_kib / wscale:.0f} KiB)\n\n")
        fo.write("| kernel | calls | avg us | % time | HBM fetch MB/launch | HBM write MB/launch | VGPR | scratch |\n")
        fo.write("|---|---|---|---|---|---|---|---|\n")
        for k, c, us, pct, fb, wb, vg in rows:
            fo.write(f"| {k} | {c} | {us:.1f} | {pct:.1f} | {fb / 1e6:.1f} | {wb / 1e6:.1f} | {vg[0]} | {vg[2]} |\n")


def counter_table(out_md, dirs):
    """Per-kernel average of every counter in the given --pmc output directories, appended to
    out_md as one table (rows: step kernels, columns: counters)."""
    vals = collections.defaultdict(dict)
    for d in dirs:
        per = collections.defaultdict(lambda: collections.defaultdict(list))
        for r in _csv(d, "counter_collection.csv"):
            if "sphg" in r["Kernel_Name"]:
                per[short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
        for k, cs in per.items():
            for c, v in cs.items():
                vals[k][c] = sum(v) / len(v)
    cols = sorted({c for v in vals.values() for c in v})
    with open(out_md, "a") as fo:
        fo.write("\n| kernel | " + " | ".join(cols) + " |\n")
        fo.write("|---" * (len(cols) + 1) + "|\n")
        for k in sorted(vals):
            fo.write(f"| {k} | " + " | ".join(f"{vals[k].get(c, float('nan')):.4g}" for c in cols) + " |\n")


if __name__ == "__main__":
    if sys.argv[1] == "--counters":
        counter_table(sys.argv[2], sys.argv[3:])
    else:
        main(*sys.argv[1:9])
