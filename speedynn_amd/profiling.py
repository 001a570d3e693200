"""Summaries of rocprofv3 output for the committed profiles.

    python -m speedynn_amd.profiling stats <kernel_stats.csv> --title T --command C
    python -m speedynn_amd.profiling pmc <counter_collection.csv> [--kernel REGEX]
    python -m speedynn_amd.profiling traffic --kernel NAME --workload W \
        --fetch <csv> --write <csv> --copy-fetch <csv> --copy-write <csv> \
        --copy-bytes B --command-fetch C1 --command-write C2

`stats` turns a `--kernel-trace --stats` table into markdown; `pmc` prints the
mean of each counter per dispatch and kernel; `traffic` writes the HBM bytes
per launch of one kernel: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each
scaled by what the same counter reads for a device copy of known size (a
copy of B bytes reads and writes B bytes), as a markdown file whose line
`traffic_bytes_per_launch: N` bench.py reads.
"""
from __future__ import annotations

import argparse
import collections
import csv
import re
import sys


def _short(name: str) -> str:
    return name.split("(")[0].replace("void ", "").strip()


def stats_markdown(path: str, title: str, command: str, top: int = 16) -> str:
    rows = list(csv.DictReader(open(path)))
    out = [f"# {title}", "", "Command (repository root, one MI355X):", "", "```", command, "```",
           "", "| kernel | calls | total ms | avg us | min us | max us | % |",
           "|---|---|---|---|---|---|---|"]
    for r in rows[:top]:
        out.append(f"| `{_short(r['Name'])}` | {r['Calls']} | "
                   f"{float(r['TotalDurationNs']) / 1e6:.2f} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                   f"{float(r['Percentage']):.2f} |")
    return "\n".join(out) + "\n"


def pmc_means(path: str, kernel: str | None = None) -> dict:
    """{(kernel, counter): (mean per dispatch, dispatches)} from a counter CSV."""
    tot = collections.defaultdict(float)
    cnt = collections.Counter()
    pat = re.compile(kernel) if kernel else None
    for r in csv.DictReader(open(path)):
        name = _short(r["Kernel_Name"])
        if pat and not pat.search(name):
            continue
        key = (name, r["Counter_Name"])
        tot[key] += float(r["Counter_Value"])
        cnt[key] += 1
    return {k: (tot[k] / cnt[k], cnt[k]) for k in tot}


def _one(means: dict, counter: str) -> tuple[str, float, int]:
    """Mean per dispatch over every kernel the pattern matched (dispatch-weighted)."""
    hits = [(k[0], v[0], v[1]) for k, v in means.items() if k[1] == counter]
    if not hits:
        raise SystemExit(f"no kernel with {counter}")
    total = sum(m * c for _, m, c in hits)
    count = sum(c for _, _, c in hits)
    return " + ".join(sorted(h[0] for h in hits)), total / count, count


def traffic_markdown(a) -> str:
    fk, fetch, fn = _one(pmc_means(a.fetch, a.kernel), "FETCH_SIZE")
    wk, write, wn = _one(pmc_means(a.write, a.kernel), "WRITE_SIZE")
    _, cfetch, _ = _one(pmc_means(a.copy_fetch, a.copy_kernel), "FETCH_SIZE")
    _, cwrite, _ = _one(pmc_means(a.copy_write, a.copy_kernel), "WRITE_SIZE")
    kib = a.copy_bytes / 1024.0
    ffac, wfac = kib / cfetch, kib / cwrite  # true KiB per counted KiB
    read_b = fetch * ffac * 1024.0
    write_b = write * wfac * 1024.0
    lines = [
        f"# HBM traffic: {fk} ({a.workload})", "",
        "FETCH_SIZE and WRITE_SIZE (KiB per dispatch, rocprofv3 --pmc, one counter per run),",
        "each scaled by the same counter on a device copy of known size.", "",
        "Commands (repository root, one MI355X):", "", "```", a.command_fetch, a.command_write,
        a.command_copy or "", "```", "",
        "| quantity | value |", "|---|---|",
        f"| dispatches (FETCH run / WRITE run) | {fn} / {wn} |",
        f"| FETCH_SIZE mean per dispatch | {fetch:.1f} KiB |",
        f"| WRITE_SIZE mean per dispatch | {write:.1f} KiB |",
        f"| copy of {a.copy_bytes} B: FETCH_SIZE / WRITE_SIZE | {cfetch:.1f} / {cwrite:.1f} KiB |",
        f"| scale factors (true / counted) | {ffac:.4f} / {wfac:.4f} |",
        f"| read bytes per launch | {read_b:.4e} |",
        f"| write bytes per launch | {write_b:.4e} |", "",
        f"traffic_bytes_per_launch: {read_b + write_b:.6e}", "",
    ]
    return "\n".join(lines)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--title", required=True)
    s.add_argument("--command", required=True)
    p = sub.add_parser("pmc")
    p.add_argument("csv")
    p.add_argument("--kernel")
    t = sub.add_parser("traffic")
    t.add_argument("--kernel", required=True)
    t.add_argument("--workload", required=True)
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-fetch", required=True)
    t.add_argument("--copy-write", required=True)
    t.add_argument("--copy-kernel", default=None)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--command-fetch", required=True)
    t.add_argument("--command-write", required=True)
    t.add_argument("--command-copy", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        sys.stdout.write(stats_markdown(a.csv, a.title, a.command))
    elif a.cmd == "pmc":
        for (k, c), (v, n) in sorted(pmc_means(a.csv, a.kernel).items()):
            print(f"{k}\t{c}\t{v:.3f}\t{n}")
    else:
        sys.stdout.write(traffic_markdown(a))
    return 0


if __name__ == "__main__":
    sys.exit(main())
