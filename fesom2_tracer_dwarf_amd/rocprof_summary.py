"""Summaries of rocprofv3 CSV output for the tracer-advection kernels.

    python -m fesom2_tracer_dwarf_amd.rocprof_summary stats <kernel_stats.csv>
    python -m fesom2_tracer_dwarf_amd.rocprof_summary traffic <fetch.csv> <write.csv> <copy_fetch.csv> <copy_write.csv>
    python -m fesom2_tracer_dwarf_amd.rocprof_summary counters <counter_collection.csv> [kernel-name-prefix ...]

`stats` prints the per-kernel duration table of a `--kernel-trace --stats`
run. `traffic` turns two `--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs of the
benchmark, plus the same two counters read on a plain device-to-device copy
of known size (`bench.py --calibrate-copy`), into HBM bytes per dispatch:
each counter is scaled by (bytes the copy moved) / (what the counter read for
it), so the figures do not depend on the counter's unit or sampling.

`counters` prints, per kernel (optionally only those whose short name starts
with one of the prefixes), the dispatch count and the mean per dispatch of
every counter in a `--pmc` run.

`read_traffic(path)` parses a `traffic` table back (the benchmark reports the
dominant kernel's measured bytes from it).
"""
from __future__ import annotations

import csv
import re
import sys
from collections import defaultdict
from pathlib import Path

COPY_BYTES = 1 << 30  # bench.py --calibrate-copy: 2^27 float64 each way per launch


def short_name(kernel: str) -> str:
    kernel = kernel.replace("(anonymous namespace)::", "")
    m = re.match(r"void fta::(\w+)<([^>]*)>", kernel)
    if m:
        return f"{m.group(1)}<{m.group(2).replace('fta::', '')}>"
    m = re.match(r"(?:void )?(?:fta::)?(\w+)", kernel)
    return m.group(1) if m else kernel[:48]


def kernel_stats(path: str) -> list[tuple[str, int, float, float]]:
    rows = []
    for r in csv.DictReader(open(path)):
        rows.append((short_name(r["Name"]), int(r["Calls"]), float(r["AverageNs"]) / 1e3,
                     float(r["Percentage"])))
    return rows


def counter_means(path: str, counter: str) -> dict[str, float]:
    acc = defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == counter:
            acc[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def copy_scale(path: str, counter: str) -> float:
    vals = [float(r["Counter_Value"]) for r in csv.DictReader(open(path))
            if r["Counter_Name"] == counter and "copyBuffer" in r["Kernel_Name"]]
    if not vals:
        raise SystemExit(f"no copy dispatch with {counter} in {path}")
    return COPY_BYTES / (sum(vals) / len(vals))


def traffic(fetch: str, write: str, cfetch: str, cwrite: str) -> tuple[dict, float, float]:
    sf, sw = copy_scale(cfetch, "FETCH_SIZE"), copy_scale(cwrite, "WRITE_SIZE")
    f, w = counter_means(fetch, "FETCH_SIZE"), counter_means(write, "WRITE_SIZE")
    out = {}
    for k in sorted(set(f) | set(w)):
        if not k.startswith("k_"):
            continue
        out[k] = (f.get(k, 0.0) * sf, w.get(k, 0.0) * sw)
    return out, sf, sw


def read_traffic(path: str | Path) -> dict[str, dict[str, float]]:
    """{kernel_base_name: {precision tag: bytes per dispatch}} from a traffic table."""
    res: dict[str, dict[str, float]] = defaultdict(dict)
    p = Path(path)
    if not p.exists():
        return {}
    for line in p.read_text().splitlines():
        m = re.match(r"\|\s*(k_\w+)<([^|]*)>\s*\|\s*([\d.eE+]+)\s*\|\s*([\d.eE+]+)\s*\|\s*([\d.eE+]+)\s*\|", line)
        if not m:
            continue
        prec = "fp64" if m.group(2).startswith("double") else "fp32"
        if "__half" in m.group(2):
            prec = "fp32_h16"
        res[m.group(1)].setdefault(prec, float(m.group(5)) * 1e9)
    return dict(res)


def main(argv: list[str]) -> None:
    if len(argv) >= 2 and argv[0] == "stats":
        print("| kernel | calls | mean us | % of GPU time |")
        print("|---|---|---|---|")
        for k, n, us, pct in kernel_stats(argv[1]):
            print(f"| {k} | {n} | {us:.1f} | {pct:.2f} |")
        return
    if len(argv) == 5 and argv[0] == "traffic":
        t, sf, sw = traffic(*argv[1:])
        print(f"scale: FETCH_SIZE x {sf:.4g} B/unit, WRITE_SIZE x {sw:.4g} B/unit (from the 1 GiB copy)\n")
        print("| kernel | read GB | write GB | total GB |")
        print("|---|---|---|---|")
        for k, (r, w) in t.items():
            print(f"| {k} | {r / 1e9:.4f} | {w / 1e9:.4f} | {(r + w) / 1e9:.4f} |")
        return
    if len(argv) >= 2 and argv[0] == "counters":
        acc, names = defaultdict(lambda: defaultdict(list)), []
        for r in csv.DictReader(open(argv[1])):
            k = short_name(r["Kernel_Name"])
            if len(argv) > 2 and not any(k.startswith(p) for p in argv[2:]):
                continue
            if r["Counter_Name"] not in names:
                names.append(r["Counter_Name"])
            acc[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
        print("| kernel | dispatches | " + " | ".join(names) + " |")
        print("|---|---|" + "---|" * len(names))
        for k, c in acc.items():
            n = max(len(v) for v in c.values())
            vals = [sum(c[x]) / len(c[x]) if c[x] else float("nan") for x in names]
            print(f"| {k} | {n} | " + " | ".join(f"{v:.4g}" for v in vals) + " |")
        return
    raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
