"""Summaries of rocprofv3 CSV output for the committed profiles.

`python -m imagebasedspatiotemporalinterpolation_amd.profile_summary DIR` reads
the runs under DIR (made with `--output-format csv`):

  trace/trace_kernel_stats.csv           --kernel-trace --stats of bench.py
  fetch/fetch_counter_collection.csv     --pmc FETCH_SIZE of bench.py
  write/write_counter_collection.csv     --pmc WRITE_SIZE of bench.py
  cfetch/cf_counter_collection.csv       --pmc FETCH_SIZE of bench.py --copy-probe
  cwrite/cw_counter_collection.csv       --pmc WRITE_SIZE of bench.py --copy-probe
  sq/sq_counter_collection.csv           --pmc SQ_WAVES SQ_INSTS_VALU ... of bench.py
  atom/atom_counter_collection.csv       --pmc TCC_ATOMIC_sum TCC_EA0_ATOMIC_sum of bench.py
                                         (optional: L2 atomic requests, and those sent on
                                         to memory)

and writes the per-kernel counter file bench.py reads for its roofline
(--out, default profiles/counters_1080p.json): per launch, the HBM traffic
FETCH_SIZE * fetch_scale + WRITE_SIZE * write_scale (KiB -> bytes; the scales
make a plain device copy of known size read exactly its bytes), waves and
instruction counts (all waves). It prints the markdown tables of the profile
notes.
"""
from __future__ import annotations

import argparse
import csv
import json
import re
from collections import defaultdict
from pathlib import Path

# the kernels of the per-frame path whose counters the bench looks up
KERNELS = re.compile(r"(clear_splat_kernel|reproject_splat_kernel|reproject_splat_binned_kernel|"
                     r"resolve_kernel<[^>]*>|clear_batch_kernel|splat_batch_kernel|resolve_batch_kernel<[^>]*>)")
COPY_BYTES = 64 * 1024 * 1024 * 4  # bench.py --copy-probe: one 256 MiB copy, repeated
SQ = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR", "SQ_INSTS_LDS")


def short_name(full: str) -> str | None:
    """`resolve_kernel<1, false, 2, 16, 16>` etc. from a demangled kernel name."""
    m = KERNELS.search(full)
    return m.group(1) if m else None


def counters(path: Path) -> dict:
    """{kernel: {counter: [value per dispatch]}} of one --pmc CSV (values of a
    dispatch's counter summed over its rows)."""
    per = defaultdict(lambda: defaultdict(float))
    names = {}
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            key = (r["Dispatch_Id"], r["Counter_Name"])
            per[key]["v"] += float(r["Counter_Value"])
            names[key] = r["Kernel_Name"]
    out = defaultdict(lambda: defaultdict(list))
    for (disp, ctr), v in per.items():
        out[names[(disp, ctr)]][ctr].append(v["v"])
    return out


def copy_scale(path: Path, counter: str) -> float:
    """Bytes per KiB reading that makes the copy probe's largest copy kernel
    move exactly COPY_BYTES."""
    best = 0.0
    for name, cs in counters(path).items():
        if "copy" in name.lower() or "elementwise" in name.lower():
            for v in cs.get(counter, []):
                best = max(best, v)
    return COPY_BYTES / (best * 1024.0) if best else 1.0


def isolated(trace_csv: Path) -> dict:
    """{kernel name: (launches, average ns)} over the launches that ran alone:
    no other kernel's interval overlaps theirs (the bench also renders frames
    on two streams at once, which stretches the kernels that overlap)."""
    iv = []
    with open(trace_csv, newline="") as f:
        for r in csv.DictReader(f):
            iv.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    iv.sort()
    out = defaultdict(list)
    reach = 0  # latest end among the earlier launches
    for i, (a, b, name) in enumerate(iv):
        nxt = iv[i + 1][0] if i + 1 < len(iv) else b
        if reach <= a and b <= nxt:
            out[name].append(b - a)
        reach = max(reach, b)
    return {k: (len(v), sum(v) / len(v)) for k, v in out.items()}


def summarise(root: Path, scales: tuple | None = None, workload: str | None = None) -> dict:
    """Per-kernel summary of the runs under root. scales = (fetch, write)
    byte scales taken from an earlier calibration instead of the copy runs."""
    if scales is None:
        fetch_scale = round(copy_scale(root / "cfetch/cf_counter_collection.csv", "FETCH_SIZE"), 4)
        write_scale = round(copy_scale(root / "cwrite/cw_counter_collection.csv", "WRITE_SIZE"), 4)
    else:
        fetch_scale, write_scale = scales
    ks: dict = {}

    def entry(full):
        s = short_name(full)
        if s is None:
            return None
        return ks.setdefault(s, {})

    for name, cs in counters(root / "fetch/fetch_counter_collection.csv").items():
        e = entry(name)
        if e is not None and cs.get("FETCH_SIZE"):
            v = cs["FETCH_SIZE"]
            e["launches"] = len(v)
            e["fetch_kib"] = round(sum(v) / len(v), 1)
    for name, cs in counters(root / "write/write_counter_collection.csv").items():
        e = entry(name)
        if e is not None and cs.get("WRITE_SIZE"):
            v = cs["WRITE_SIZE"]
            e["write_kib"] = round(sum(v) / len(v), 1)
    for name, cs in counters(root / "sq/sq_counter_collection.csv").items():
        e = entry(name)
        if e is None:
            continue
        for c in SQ:
            v = cs.get(c)
            if v:
                e[c[3:].lower() if c != "SQ_WAVES" else "waves"] = int(round(sum(v) / len(v)))
    atom = root / "atom/atom_counter_collection.csv"
    if atom.exists():
        for name, cs in counters(atom).items():
            e = entry(name)
            if e is None:
                continue
            for c, key in (("TCC_ATOMIC_sum", "l2_atomics"), ("TCC_EA0_ATOMIC_sum", "memory_atomics")):
                v = cs.get(c)
                if v:
                    e[key] = int(round(sum(v) / len(v)))
    for e in ks.values():
        if "fetch_kib" in e and "write_kib" in e:
            e["traffic"] = int(round((e["fetch_kib"] * fetch_scale + e["write_kib"] * write_scale) * 1024))
    with open(root / "trace/trace_kernel_stats.csv", newline="") as f:
        for r in csv.DictReader(f):
            e = entry(r["Name"])
            if e is not None:
                e["trace_calls"] = int(r["Calls"])
                e["trace_avg_ns"] = int(float(r["AverageNs"]))
    for name, (n, avg) in isolated(root / "trace/trace_kernel_trace.csv").items():
        e = entry(name)
        if e is not None and n:
            e["trace_isolated_calls"] = n
            e["trace_isolated_avg_ns"] = int(avg)
    # rename the SQ fields to the names bench.py reads
    for e in ks.values():
        for a, b in (("insts_valu", "valu_insts"), ("insts_salu", "salu_insts"), ("insts_vmem_rd", "vmem_rd_insts"),
                     ("insts_vmem_wr", "vmem_wr_insts"), ("insts_lds", "lds_insts")):
            if a in e:
                e[b] = e.pop(a)
    return {
        "workload": workload or "bench.py 1080p workload: 1920x1080 + 2 aux 960x540, soft shadows, 7 frames per "
                                "keyframe pair, search <= 32",
        "units": "traffic: HBM bytes per launch = FETCH_SIZE*fetch_scale + WRITE_SIZE*write_scale (KiB -> B); "
                 "instruction counts per launch (all waves)",
        "fetch_scale": fetch_scale,
        "write_scale": write_scale,
        "kernels": dict(sorted(ks.items())),
    }


def tables(d: dict, simds: int = 1024, clock_hz: float = 2.4e9) -> str:
    rows = ["| Kernel | Launches | Trace avg us (alone) | Traffic MB | SQ_WAVES | VALU / wave | SALU / wave "
            "| VMEM rd / wave | LDS / wave | VALU issue us | VALU issue / duration |",
            "|---|---|---|---|---|---|---|---|---|---|---|"]
    atoms = [(k, e) for k, e in d["kernels"].items() if "l2_atomics" in e]
    for k, e in d["kernels"].items():
        if "waves" not in e:  # traced only (no counter run launched it)
            continue
        w = e["waves"]
        dur = e.get("trace_isolated_avg_ns", e.get("trace_avg_ns", 0)) / 1e3
        issue = e.get("valu_insts", 0) * 4 / (simds * clock_hz) * 1e6
        rows.append(f"| `{k}` | {e.get('launches', '')} | {dur:.1f} | {e.get('traffic', 0) / 1e6:.1f} | {w} "
                    f"| {e.get('valu_insts', 0) / w:.0f} | {e.get('salu_insts', 0) / w:.0f} "
                    f"| {e.get('vmem_rd_insts', 0) / w:.1f} | {e.get('lds_insts', 0) / w:.1f} | {issue:.1f} "
                    f"| {issue / dur if dur else 0:.2f} |")
    if atoms:
        rows += ["", "| Kernel | L2 atomic requests (TCC_ATOMIC_sum) | sent on to memory (TCC_EA0_ATOMIC_sum) |",
                 "|---|---|---|"]
        rows += [f"| `{k}` | {e['l2_atomics']} | {e.get('memory_atomics', 0)} |" for k, e in atoms]
    return "\n".join(rows)


def _main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("root", help="directory holding the rocprofv3 runs")
    ap.add_argument("--out", default=str(Path(__file__).resolve().parent.parent / "profiles" / "counters_1080p.json"))
    ap.add_argument("--scales-from", help="take the FETCH/WRITE byte scales from this counter file "
                                          "instead of copy runs under root")
    ap.add_argument("--workload", help="workload description stored in the output")
    a = ap.parse_args(argv)
    scales = None
    if a.scales_from:
        ref = json.loads(Path(a.scales_from).read_text())
        scales = (ref["fetch_scale"], ref["write_scale"])
    d = summarise(Path(a.root), scales, a.workload)
    Path(a.out).write_text(json.dumps(d, indent=1) + "\n")
    print(f"fetch_scale {d['fetch_scale']} write_scale {d['write_scale']}")
    print(tables(d))


if __name__ == "__main__":
    _main()
