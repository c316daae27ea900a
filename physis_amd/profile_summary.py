"""Summarises rocprofv3 output.

    python3 -m physis_amd.profile_summary FETCH_DIR WRITE_DIR KTRACE_DIR OUT.md
    python3 -m physis_amd.profile_summary --ktrace KTRACE_DIR OUT.md COMMAND
    python3 -m physis_amd.profile_summary --unet KTRACE_DIR OUT.md COMMAND
    python3 -m physis_amd.profile_summary --counters DIR... OUT.md COMMAND
    python3 -m physis_amd.profile_summary --counters-unet DIR... OUT.md COMMAND

The first form summarises physis_amd.profile_kernels (HBM traffic of the
hot-path kernels); --ktrace tabulates the kernel statistics of any command;
--unet sets the U-Net convolution kernels of physis_amd.profile_unet
against their roofline.

FETCH_DIR / WRITE_DIR hold the *_counter_collection.csv of a
`--pmc FETCH_SIZE` and a `--pmc WRITE_SIZE` run, KTRACE_DIR the
*_kernel_stats.csv of a `--kernel-trace --stats` run of the same workload.
Counter values are converted to bytes with the factor that makes the
512 MiB device-to-device copy read and write exactly 512 MiB.  Writes a
Markdown table (OUT.md) and, next to it, diffusion7_512_traffic.json with the
7-point run kernel's HBM bytes per launch (read by bench.py).
"""
from __future__ import annotations

import csv
import glob
import json
import os
import sys
from collections import defaultdict

COPY_BYTES = 512 * 2 ** 20
KERNELS = [  # (label, name fragment, algorithmic bytes per launch)
    # Two-sweep kernels: the minimum HBM traffic of a launch, which reads
    # its inputs once and writes its output once for two sweeps.
    ("7-pt clamped two-sweep kernel, 512^3 (2 sweeps per launch, "
     "minimum bytes: read u, write u'')",
     "__PSStencilRun2_diffusion7_kernel<", 8 * 512 ** 3),
    ("7-pt clamped run kernel, 512^3", "__PSStencilRun_diffusion7_kernel<",
     8 * 512 ** 3),
    ("7-pt periodic two-sweep kernel, 512^3 (2 sweeps per launch, minimum "
     "bytes)", "__PSStencilRun2_diffusion7_periodic_kernel<", 8 * 512 ** 3),
    ("19-pt Himeno two-sweep kernel, 512x256x256 (2 sweeps per launch, "
     "minimum bytes: 12 fields + p read, p'' written)",
     "__PSStencilRun2_jacobi_kernel<", 56 * 512 * 256 * 256),
    ("19-pt Himeno run kernel, 512x256x256", "__PSStencilRun_jacobi_kernel<",
     56 * 512 * 256 * 256),
    ("PSReduce SUM stage 1, 512^3", "ReduceStage1<float, 2>", 4 * 512 ** 3),
    ("BF16 two-sweep inner Jacobi pass, 512^3 (2 sweeps per launch, "
     "minimum bytes)", "TwoSweepKernel<physis_hip::bf16_t", 28 * 512 ** 3),
    ("BF16 inner Jacobi sweep, 512^3", "SweepKernel<physis_hip::bf16_t",
     28 * 512 ** 3),
    ("FP64 update + residual (BF16 E/RHS), 512^3",
     "ResidualKernel<double, physis_hip::bf16_t, 2, true, true>",
     68 * 512 ** 3),
]


def _counter(d: str, name: str):
    f = glob.glob(os.path.join(d, "*_counter_collection.csv"))[0]
    vals = defaultdict(list)
    for r in csv.DictReader(open(f)):
        if r["Counter_Name"] == name:
            vals[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return vals


def _matches(name: str, frag: str) -> bool:
    """frag names the kernel itself (it starts the name or follows a space or
    a namespace separator), not a longer name that ends with it."""
    import re
    return re.search(r"(^|[\s:])" + re.escape(frag), name) is not None


def _pick(vals, frag):
    ks = [k for k in vals if _matches(k, frag)]
    if not ks:
        raise SystemExit(f"kernel {frag!r}: no match")
    k = max(ks, key=lambda k: len(vals[k]))
    v = vals[k]
    return k, sum(v) / len(v), len(v)


def main(fetch_dir, write_dir, kt_dir, out_md) -> None:
    fetch = _counter(fetch_dir, "FETCH_SIZE")
    write = _counter(write_dir, "WRITE_SIZE")
    copy = [k for k in write if "copyBuffer" in k][0]
    big_w = sorted(write[copy])[-3:]
    big_f = sorted(fetch[copy])[-3:]
    ws = COPY_BYTES / (sum(big_w) / len(big_w))
    fs = COPY_BYTES / (sum(big_f) / len(big_f))
    kt = {}
    for f in glob.glob(os.path.join(kt_dir, "*_kernel_stats.csv")):
        for r in csv.DictReader(open(f)):
            kt[r["Name"]] = (float(r["AverageNs"]), int(r["Calls"]))
    rows = []
    traffic = {}
    for label, frag, alg in KERNELS:
        name, f, n = _pick(fetch, frag)
        _, w, _ = _pick(write, frag)
        fb, wb = f * fs, w * ws
        cands = [(k, v) for k, v in kt.items() if _matches(k, frag)]
        # several instantiations (vector width, reduction operator): the
        # one launched most often
        cands.sort(key=lambda kv: -kv[1][1])
        avg_ns = cands[0][1][0] if cands else float("nan")
        rows.append((label, fb, wb, alg, avg_ns, n))
        if "diffusion7_kernel<" in frag:
            traffic[name.split("(")[0].replace("void ", "").split("<")[0]] = {
                "hbm_bytes_per_launch": int(fb + wb),
                "algorithmic_bytes_per_launch": alg}
    lines = [
        "# HBM traffic of the hot-path kernels (MI355X)",
        "",
        "Workload: `python3 -m physis_amd.profile_kernels`. Commands:",
        "",
        "```",
        "rocprofv3 --kernel-trace --stats -d OUT/kt -o kt --output-format csv "
        "-- python3 -m physis_amd.profile_kernels",
        "rocprofv3 --pmc FETCH_SIZE -d OUT/fetch -o fetch --output-format csv "
        "-- python3 -m physis_amd.profile_kernels",
        "rocprofv3 --pmc WRITE_SIZE -d OUT/write -o write --output-format csv "
        "-- python3 -m physis_amd.profile_kernels",
        "python3 -m physis_amd.profile_summary OUT/fetch OUT/write OUT/kt "
        "profiles/hot_kernels_traffic.md",
        "```",
        "",
        f"Counter units: the 512 MiB device-to-device copy reads "
        f"{sum(big_f)/len(big_f):.0f} FETCH_SIZE units and writes "
        f"{sum(big_w)/len(big_w):.0f} WRITE_SIZE units, so 1 FETCH_SIZE unit "
        f"= {fs:.1f} B and 1 WRITE_SIZE unit = {ws:.1f} B here.",
        "",
        "| kernel | HBM read MB | HBM write MB | algorithmic MB | traffic / "
        "algorithmic | avg time (us) | achieved algorithmic GB/s | % of 8 TB/s |",
        "|---|---|---|---|---|---|---|---|",
    ]
    for label, fb, wb, alg, ns, n in rows:
        gbs = alg / (ns * 1e-9) / 1e9 if ns == ns else float("nan")
        lines.append(f"| {label} | {fb/1e6:.1f} | {wb/1e6:.1f} | {alg/1e6:.1f} "
                     f"| {(fb+wb)/alg:.3f} | {ns/1e3:.1f} | {gbs:.0f} | "
                     f"{gbs/8000*100:.1f} |")
    lines.append("")
    os.makedirs(os.path.dirname(os.path.abspath(out_md)), exist_ok=True)
    open(out_md, "w").write("\n".join(lines))
    jpath = os.path.join(os.path.dirname(os.path.abspath(out_md)),
                         "diffusion7_512_traffic.json")
    json.dump({"kernels": traffic,
               "fetch_bytes_per_unit": fs, "write_bytes_per_unit": ws,
               "source": os.path.basename(out_md)},
              open(jpath, "w"), indent=1)
    print("\n".join(lines))


def ktrace_md(kt_dir: str, out_md: str, command: str) -> None:
    """Kernel-trace statistics of any command as a Markdown table."""
    rows = []
    for f in glob.glob(os.path.join(kt_dir, "*_kernel_stats.csv")):
        for r in csv.DictReader(open(f)):
            rows.append((r["Name"], int(r["Calls"]), float(r["AverageNs"]),
                         float(r["MinNs"]), float(r["MaxNs"]),
                         float(r["Percentage"])))
    lines = ["# Kernel trace (rocprofv3 --kernel-trace --stats, MI355X)", "",
             "```", command, "```", "",
             "| kernel | calls | avg us | min us | max us | % time |",
             "|---|---|---|---|---|---|"]
    for n, c, a, lo, hi, pct in sorted(rows, key=lambda r: -r[5]):
        # drop the parameter list (the last top-level parenthesis group);
        # names may contain "(anonymous namespace)"
        name = n.replace("(anonymous namespace)", "{anon}")
        name = name.split("(")[0].replace("{anon}", "(anonymous namespace)")
        name = name.replace("|", "/")
        lines.append(f"| `{name}` | {c} | {a/1e3:.1f} | {lo/1e3:.1f} | "
                     f"{hi/1e3:.1f} | {pct:.2f} |")
    lines.append("")
    open(out_md, "w").write("\n".join(lines))
    print("\n".join(lines))


# U-Net convolution shapes (input, output channels) -> pyramid level, and
# positions per launch at level 1..3 for batch 8 of 128^3 problems
UNET_LEVEL = {(8, 8): 1, (24, 8): 1, (8, 24): 1, (8, 16): 2, (16, 8): 2,
              (16, 16): 2, (48, 16): 2, (16, 48): 2, (16, 32): 3,
              (32, 16): 3, (32, 32): 3}
UNET_POS = {1: 8 * 128 ** 3, 2: 8 * 32 ** 3, 3: 8 * 8 ** 3}
MFMA_BF16_PEAK = 2.5e15  # dense BF16 FLOP/s
HBM_PEAK = 8e12          # B/s


def unet_md(kt_dir: str, out_md: str, command: str) -> None:
    """Convolution kernels of physis_amd.profile_unet against their
    roofline: useful flops 2*27*Cin*Cout per position, algorithmic bytes
    2*(Cin + Cout) per position (BF16 input read once, output or the
    output gradient read/written once), time bound max(flops / MFMA peak,
    bytes / HBM peak)."""
    import re
    rows = []
    for f in glob.glob(os.path.join(kt_dir, "*_kernel_stats.csv")):
        for r in csv.DictReader(open(f)):
            m = re.search(r"ConvZ?Kernel<(\d+), (\d+), (true|false)",
                          r["Name"])
            mp = re.search(r"ConvZPairKernel<(\d+), (true|false)", r["Name"])
            w = re.search(r"Wgrad(?:Lds|Z)?Kernel<(\d+), (\d+)",
                          r["Name"])
            if m:
                cin, cout = int(m.group(1)), int(m.group(2))
                what = "forward" if m.group(3) == "true" else "backward data"
            elif mp:  # 8 output channels, two rows per wave
                cin, cout = int(mp.group(1)), 8
                what = "forward" if mp.group(2) == "true" else "backward data"
            elif w:
                cin, cout = int(w.group(1)), int(w.group(2))
                what = "backward filter"
            else:
                continue
            lvl = UNET_LEVEL[(cin, cout)]
            pos = UNET_POS[lvl]
            flops = 2 * 27 * cin * cout * pos
            byts = 2 * (cin + cout) * pos
            t = float(r["AverageNs"]) * 1e-9
            tb = max(flops / MFMA_BF16_PEAK, byts / HBM_PEAK)
            rows.append((what, cin, cout, lvl, int(r["Calls"]), t, flops,
                         byts, tb, float(r["Percentage"])))
    lines = ["# U-Net convolution kernels against their roofline (MI355X)", "",
             "```", command, "```", "",
             "Batch 8 of 128^3 problems: level 1 = 8 x 128^3, level 2 = "
             "8 x 32^3, level 3 = 8 x 8^3 positions.  Useful flops "
             "2*27*Cin*Cout per position (MFMA padding not counted); "
             "algorithmic bytes 2*(Cin+Cout) per position; roofline time = "
             "max(flops / 2.5 PFLOP/s, bytes / 8 TB/s).", "",
             "| kernel | Cin -> Cout | level | calls | avg us | TFLOP/s | "
             "% of MFMA peak | GB/s | % of HBM peak | roofline us | "
             "% of roofline | % of profile time |",
             "|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for (what, cin, cout, lvl, calls, t, fl, by, tb,
         pct) in sorted(rows, key=lambda r: (-r[9])):
        lines.append(
            f"| {what} | {cin} -> {cout} | {lvl} | {calls} | {t*1e6:.1f} | "
            f"{fl/t/1e12:.1f} | {fl/t/MFMA_BF16_PEAK*100:.2f} | "
            f"{by/t/1e9:.0f} | {by/t/HBM_PEAK*100:.1f} | {tb*1e6:.1f} | "
            f"{tb/t*100:.1f} | {pct:.1f} |")
    lines.append("")
    open(out_md, "w").write("\n".join(lines))
    print("\n".join(lines))


STENCIL_FRAGS = [
    ("two-sweep 7-pt (clamped)", "__PSStencilRun2_diffusion7_kernel<"),
    ("one-sweep 7-pt (clamped)", "__PSStencilRun_diffusion7_kernel<"),
    ("two-sweep 7-pt (periodic)", "__PSStencilRun2_diffusion7_periodic_kernel<"),
    ("512 MiB copy", "copyBuffer")]
# the 19-point kernels of physis_amd.profile_kernels (Himeno 512x256x256,
# precision tier 512^3)
POISSON_FRAGS = [
    ("Himeno two-sweep", "__PSStencilRun2_jacobi_kernel<"),
    ("BF16 inner two-sweep", "TwoSweepKernel<physis_hip::bf16_t"),
    ("FP64 update + residual", "ResidualKernel<double, physis_hip::bf16_t, 2, true"),
    ("512 MiB copy", "copyBuffer")]
# the level-1 kernels of physis_amd.profile_unet (batch 8 of 128^3)
UNET_FRAGS = [
    ("forward 8->8", "ConvZPairKernel<8, true"),
    ("forward 24->8", "ConvZPairKernel<24, true"),
    ("backward data 8->24", "ConvZKernel<8, 24, false"),
    ("backward filter 8->8", "WgradZKernel<8, 8,"),
    ("backward filter 24->8", "WgradZKernel<24, 8,")]


def counters_md(dirs, out_md: str, command: str, frags=None,
                title="the 7-point run kernels") -> None:
    """Per-launch averages of the instruction / stall counters of
    `rocprofv3 --pmc` runs (one directory per counter group) for the
    kernels of `frags` (label, kernel-name fragment), with derived rates:
    busy and wait fractions of the wave cycles, instructions per wave, LDS
    bank conflicts per LDS instruction."""
    frags = STENCIL_FRAGS if frags is None else frags
    vals = defaultdict(dict)
    for d in dirs:
        for f in glob.glob(os.path.join(d, "**", "*_counter_collection.csv"),
                           recursive=True):
            acc = defaultdict(list)
            for r in csv.DictReader(open(f)):
                for label, frag in frags:
                    if _matches(r["Kernel_Name"], frag) or (
                            frag == "copyBuffer" and frag in r["Kernel_Name"]):
                        acc[(label, r["Counter_Name"])].append(
                            float(r["Counter_Value"]))
            for (label, c), v in acc.items():
                vals[label][c] = sum(v) / len(v)
    names = sorted({c for v in vals.values() for c in v})
    lines = [f"# Instruction and stall counters of {title} (MI355X)", "",
             "```", command, "```", "",
             "Per-launch averages (rocprofv3 sums each counter over the "
             "device for a dispatch).", "",
             "| counter | " + " | ".join(l for l, _ in frags if l in vals) + " |",
             "|---|" + "---|" * len([1 for l, _ in frags if l in vals])]
    for c in names:
        lines.append(f"| {c} | " + " | ".join(
            f"{vals[l].get(c, float('nan')):.4g}" for l, _ in frags
            if l in vals) + " |")
    lines += ["", "Derived:", ""]
    for l, _ in frags:
        v = vals.get(l)
        if not v:
            continue
        parts = []
        wc = v.get("SQ_WAVE_CYCLES")
        waves = v.get("SQ_WAVES")
        if wc:
            for k in ("SQ_BUSY_CYCLES", "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY",
                      "SQ_ACTIVE_INST_ANY", "SQ_ACTIVE_INST_VALU",
                      "SQ_ACTIVE_INST_LDS", "SQ_WAIT_INST_LDS"):
                if k in v:
                    parts.append(f"{k} / SQ_WAVE_CYCLES = {v[k] / wc:.3f}")
        if waves:
            for k in ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS",
                      "SQ_INSTS_MFMA", "SQ_INSTS_VMEM_RD",
                      "SQ_INSTS_VMEM_WR"):
                if k in v:
                    parts.append(f"{k} per wave = {v[k] / waves:.1f}")
        if v.get("SQ_LDS_BANK_CONFLICT") is not None and \
                v.get("SQ_ACTIVE_INST_LDS"):
            parts.append("SQ_LDS_BANK_CONFLICT / SQ_ACTIVE_INST_LDS = "
                         f"{v['SQ_LDS_BANK_CONFLICT'] / v['SQ_ACTIVE_INST_LDS']:.3f}")
        lines.append(f"* {l}: " + "; ".join(parts))
    lines.append("")
    open(out_md, "w").write("\n".join(lines))
    print("\n".join(lines))


if __name__ == "__main__":
    if len(sys.argv) >= 5 and sys.argv[1] == "--counters":
        counters_md(sys.argv[2:-2], sys.argv[-2], sys.argv[-1])
        sys.exit(0)
    if len(sys.argv) >= 5 and sys.argv[1] == "--counters-19pt":
        counters_md(sys.argv[2:-2], sys.argv[-2], sys.argv[-1], POISSON_FRAGS,
                    "the 19-point kernels")
        sys.exit(0)
    if len(sys.argv) >= 5 and sys.argv[1] == "--counters-unet":
        counters_md(sys.argv[2:-2], sys.argv[-2], sys.argv[-1], UNET_FRAGS,
                    "the level-1 U-Net convolution kernels")
        sys.exit(0)
    if len(sys.argv) == 5 and sys.argv[1] == "--unet":
        unet_md(*sys.argv[2:])
    elif len(sys.argv) == 5 and sys.argv[1] == "--ktrace":
        ktrace_md(*sys.argv[2:])
    elif len(sys.argv) == 5:
        main(*sys.argv[1:])
    else:
        raise SystemExit(__doc__)
