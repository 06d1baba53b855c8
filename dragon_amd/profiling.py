"""Reading rocprofv3 output for this library's kernels.

    python -m dragon_amd.profiling stats  KT_DIR OUT.md --command "..."
    python -m dragon_amd.profiling traffic FETCH_DIR WRITE_DIR OUT.json --steps 3 --forms fp64,mixed
    python -m dragon_amd.profiling valu PMC_DIR OUT.md --command "..." [--json OUT.json --steps 3]

``stats`` turns a ``rocprofv3 --kernel-trace --stats --output-format csv``
directory into a Markdown table (per kernel: calls, average and total device
time, the kernel class the library's own HIP-event timing reports it under).

``traffic`` turns two ``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE``
runs of ``bench.py`` (kernel filter ``dg::|elementwise``) into HBM bytes per
step for each precision form (run bench.py with --no-batch).  The counters
are scaled by what they read for bench.py's copy probe, torch's negation
kernel over 1 GiB (1 GiB read, 1 GiB written): factor = known bytes /
counted bytes.

``valu`` turns one ``rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU SQ_INSTS_VMEM_RD
SQ_INSTS_SALU SQ_WAVE_CYCLES SQ_WAIT_ANY SQ_ACTIVE_INST_VALU SQ_BUSY_CYCLES``
run into a table per kernel class and form: waves, instructions per wave,
device time per launch (from the counter run's own timestamps) and the VALU
issue share, SQ_INSTS_VALU x 4 cycles (a wave64 instruction on a 16-lane
SIMD) over time x 2.4 GHz x 1024 SIMDs -- a lower bound when the engine
clock runs below its 2.4 GHz peak.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import re

# demangled kernel name -> the kernel class of dg_kernel_time()
_CLASSES = [
    (r"k_sweep_(march|wave)<\d, (true|false), (double|float), (\d), false>", lambda m: "pred_sweep_" + "xyz"[int(m.group(4))]),
    (r"k_sweep_(march|wave)<\d, (true|false), (double|float), (\d), true>", lambda m: "flux_sweep_" + "xyz"[int(m.group(4))]),
    (r"k_emf[23]<true>", "emf_pred"),
    (r"k_emf[23]<false>", "emf"),
    (r"k_emf3_tiled\(", "emf (predictor + final)"),  # one kernel for both EMFs of the 3D step
    (r"k_faraday<", "faraday (half + final)"),
    (r"k_whalf<", "whalf"),
    (r"k_apply1?<", "apply"),
    (r"k_cfl<", "cfl"),
    (r"k_bc_(cells|faces)", "boundary"),
    (r"k_half<", "half_states"),
    (r"k_flux<", "flux (reference-structured path)"),
    (r"k_(correct|ctu_flux|ctu_mhd)<", "ctu_correct"),
    (r"k_fill_u64", "failure-word reset"),
    (r"k_init_B<", "init (initialize_B_fields)"),
    (r"k_divb<", "diagnostic (div B)"),
    (r"k_(aos_to_soa|soa_to_aos)", "layout transfer"),
    (r"kb_(\w+)", lambda m: "batch " + m.group(1)),
]
STEP_CLASSES_EXCLUDED = ("init (initialize_B_fields)", "diagnostic (div B)", "layout transfer")


def kernel_class(name: str) -> str:
    for pat, cls in _CLASSES:
        m = re.search(pat, name)
        if m:
            return cls(m) if callable(cls) else cls
    return "other"


def _find(d, suffix):
    files = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not files:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    return files


def _col(row, *names):
    for n in names:
        if n in row:
            return row[n]
    raise KeyError(names)


def stats_table(kt_dir: str, command: str) -> str:
    rows = []
    for f in _find(kt_dir, "kernel_stats.csv"):
        for r in csv.DictReader(open(f)):
            rows.append(r)
    lines = [f"Command: `{command}`", "",
             "| kernel | class | calls | avg us | total ms | % |", "|---|---|---|---|---|---|"]
    for r in rows:
        name = _col(r, "Name", "KernelName", "Kernel_Name")
        calls = int(_col(r, "Calls"))
        tot = float(_col(r, "TotalDurationNs"))
        avg = float(_col(r, "AverageNs"))
        pct = float(_col(r, "Percentage"))
        short = name if len(name) < 110 else name[:107] + "..."
        lines.append(f"| `{short}` | {kernel_class(name)} | {calls} | {avg / 1e3:.1f} | {tot / 1e6:.2f} | {pct:.2f} |")
    return "\n".join(lines) + "\n"


def _pmc_records(d, counter):
    recs = {}
    for f in _find(d, "counter_collection.csv"):
        for r in csv.DictReader(open(f)):
            if _col(r, "Counter_Name") != counter:
                continue
            did = int(_col(r, "Dispatch_Id"))
            name = _col(r, "Kernel_Name")
            rec = recs.setdefault(did, [name, 0.0])
            rec[1] += float(_col(r, "Counter_Value"))
    return [(did, v[0], v[1]) for did, v in sorted(recs.items())]


def _copy_factor(recs, nbytes=1 << 30):
    """bench.py's copy probe is torch's negation kernel (1 GiB each way)."""
    vals = sorted(v for _, n, v in recs if re.search(r"neg", n, re.I))
    if not vals:
        raise RuntimeError("copy probe kernel not found in the counter run")
    kib = vals[len(vals) // 2]  # median over the probe's calls
    return nbytes / (kib * 1024.0), kib


def _split_forms(recs, forms):
    """Each form starts with its own initialize_B_fields launch."""
    starts = [i for i, (_, n, _) in enumerate(recs) if "k_init_B<" in n]
    if len(starts) != len(forms):
        raise RuntimeError(f"expected {len(forms)} init launches, saw {len(starts)}")
    out = {}
    for q, f in enumerate(forms):
        hi = starts[q + 1] if q + 1 < len(starts) else len(recs)
        out[f] = recs[starts[q]:hi]
    return out


def traffic(fetch_dir, write_dir, steps, forms, commands=()):
    fr = _pmc_records(fetch_dir, "FETCH_SIZE")
    wr = _pmc_records(write_dir, "WRITE_SIZE")
    ff, fkib = _copy_factor(fr)
    wf, wkib = _copy_factor(wr)
    out = {"unit": "bytes", "steps_per_form": steps, "commands": list(commands),
           "copy_probe": {"bytes_each_way": 1 << 30, "fetch_kib": fkib, "write_kib": wkib,
                          "fetch_factor": ff, "write_factor": wf}}
    fsplit, wsplit = _split_forms(fr, forms), _split_forms(wr, forms)
    for f in forms:
        per_class = collections.defaultdict(lambda: [0.0, 0.0])
        for _, n, v in fsplit[f]:
            c = kernel_class(n)
            if c not in STEP_CLASSES_EXCLUDED:
                per_class[c][0] += v * 1024.0 * ff / steps
        for _, n, v in wsplit[f]:
            c = kernel_class(n)
            if c not in STEP_CLASSES_EXCLUDED:
                per_class[c][1] += v * 1024.0 * wf / steps
        fetch = sum(v[0] for v in per_class.values())
        write = sum(v[1] for v in per_class.values())
        out[f] = {"bytes_per_step": fetch + write, "fetch_bytes_per_step": fetch, "write_bytes_per_step": write,
                  "per_class": {k: {"fetch": v[0], "write": v[1]} for k, v in sorted(per_class.items())}}
    return out


VALU_CYCLES, CLOCK_HZ, SIMDS = 4, 2.4e9, 1024


def _valu_agg(pmc_dir: str):
    disp = {}
    for f in _find(pmc_dir, "counter_collection.csv"):
        for r in csv.DictReader(open(f)):
            did = int(_col(r, "Dispatch_Id"))
            d = disp.setdefault(did, {"name": _col(r, "Kernel_Name"), "ns": int(_col(r, "End_Timestamp")) - int(_col(r, "Start_Timestamp"))})
            c = _col(r, "Counter_Name")
            d[c] = d.get(c, 0.0) + float(_col(r, "Counter_Value"))
    agg = collections.OrderedDict()
    for did in sorted(disp):
        d = disp[did]
        cls = kernel_class(d["name"])
        if cls in STEP_CLASSES_EXCLUDED or cls == "other" or cls.startswith("batch"):
            continue
        form = "fp64" if "p64::" in d["name"] else "mixed" if "p32::" in d["name"] else "?"
        a = agg.setdefault((form, cls), collections.defaultdict(float))
        a["launches"] += 1
        a["ns"] += d["ns"]
        for k, v in d.items():
            if k.startswith("SQ_"):
                a[k] += v
    return agg


def valu_per_step(pmc_dir: str, steps: int, command: str = "") -> dict:
    """Wave-level VALU instructions per step of each form (the step's kernel
    classes summed; the counted run advanced each form `steps` steps)."""
    out = {}
    for (form, cls), a in _valu_agg(pmc_dir).items():
        f = out.setdefault(form, {"valu_insts_per_step": 0.0, "kernel_ns_per_step": 0.0, "steps": steps, "command": command})
        f["valu_insts_per_step"] += a.get("SQ_INSTS_VALU", 0.0) / steps
        f["kernel_ns_per_step"] += a["ns"] / steps
    for f in out.values():
        f["valu_issue_share_counter_run"] = round(f["valu_insts_per_step"] * VALU_CYCLES / (f["kernel_ns_per_step"] * 1e-9 * CLOCK_HZ * SIMDS), 4)
    return out


def valu_table(pmc_dir: str, command: str) -> str:
    agg = _valu_agg(pmc_dir)
    lines = [f"Command: `{command}`", "",
             "| form | kernel class | launches | waves/launch | VALU/wave | VMEM rd/wave | SALU/wave | ms/launch | VALU issue share | wave cycles waiting |",
             "|---|---|---|---|---|---|---|---|---|---|"]
    for (form, cls), a in agg.items():
        n, w = a["launches"], a.get("SQ_WAVES", 0.0)
        if not w:
            continue
        share = a.get("SQ_INSTS_VALU", 0.0) * VALU_CYCLES / (a["ns"] * 1e-9 * CLOCK_HZ * SIMDS)
        wait = a.get("SQ_WAIT_ANY", 0.0) / a["SQ_WAVE_CYCLES"] if a.get("SQ_WAVE_CYCLES") else float("nan")
        lines.append(f"| {form} | {cls} | {int(n)} | {w / n:.0f} | {a.get('SQ_INSTS_VALU', 0.0) / w:.0f} | "
                     f"{a.get('SQ_INSTS_VMEM_RD', 0.0) / w:.0f} | {a.get('SQ_INSTS_SALU', 0.0) / w:.0f} | "
                     f"{a['ns'] / n / 1e6:.2f} | {share:.2f} | {wait:.2f} |")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m dragon_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a1 = sub.add_parser("stats")
    a1.add_argument("kt_dir")
    a1.add_argument("out")
    a1.add_argument("--command", default="")
    a2 = sub.add_parser("traffic")
    a2.add_argument("fetch_dir")
    a2.add_argument("write_dir")
    a2.add_argument("out")
    a2.add_argument("--steps", type=int, required=True, help="steps per form in the counted run (warmup + timed)")
    a2.add_argument("--forms", default="fp64,mixed")
    a2.add_argument("--command", action="append", default=[], help="the counter runs' commands, recorded in the output")
    a3 = sub.add_parser("valu")
    a3.add_argument("pmc_dir")
    a3.add_argument("out")
    a3.add_argument("--command", default="")
    a3.add_argument("--json", help="also write VALU instructions per step per form (bench.py reads it)")
    a3.add_argument("--steps", type=int, default=0, help="steps per form in the counted run (warmup + timed), for --json")
    a = ap.parse_args(argv)
    if a.cmd == "valu":
        with open(a.out, "w") as f:
            f.write(valu_table(a.pmc_dir, a.command))
        if a.json:
            if a.steps <= 0:
                ap.error("--json needs --steps")
            with open(a.json, "w") as f:
                json.dump(valu_per_step(a.pmc_dir, a.steps, a.command), f, indent=1)
    elif a.cmd == "stats":
        text = stats_table(a.kt_dir, a.command)
        with open(a.out, "w") as f:
            f.write(text)
    else:
        res = traffic(a.fetch_dir, a.write_dir, a.steps, a.forms.split(","), a.command)
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
