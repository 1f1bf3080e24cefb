"""Summaries of rocprofv3 CSV output for this library's kernels.

    python -m ennemi_amd.profiling stats  <kernel_stats.csv>  [title]
    python -m ennemi_amd.profiling traffic <fetch counter_collection.csv> <write counter_collection.csv>
    python -m ennemi_amd.profiling counters <counter_collection.csv> [kernel substring ...]

`stats` prints a Markdown table of per-kernel calls and durations from a
`rocprofv3 --kernel-trace --stats --output-format csv` run.  `traffic`
prints JSON: per kernel, the average FETCH_SIZE and WRITE_SIZE per dispatch
(KiB as the counters report them), each scaled by what the same counter
reads for a plain device-to-device copy of known size in the same run (the
`--copy-calib` step of bench.py: 1 GiB read and 1 GiB written), and their
sum in bytes per launch.  `counters` prints a Markdown table of the average
value per dispatch of every counter of a `--pmc` run, per kernel, and for the
SQ counters the shares of the waves' resident cycles (SQ_WAVE_CYCLES) spent
waiting on anything, memory included (SQ_WAIT_ANY), waiting for an
instruction to issue (SQ_WAIT_INST_ANY) or issuing VALU work
(SQ_ACTIVE_INST_VALU); TCC_HIT / TCC_MISS give the L2 hit rate.
"""

from __future__ import annotations

import collections
import csv
import json
import sys

COPY_BYTES = 1 << 30


def _short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "").replace("void ", "")
    depth, out = 0, []
    for ch in name:
        if ch == "(" and depth == 0:
            break
        out.append(ch)
        depth += ch == "<"
        depth -= ch == ">"
    return "".join(out).strip()


def kernel_stats(path: str, title: str = "") -> str:
    rows = list(csv.DictReader(open(path)))
    lines = []
    if title:
        lines += [f"### {title}", ""]
    lines += ["| kernel | calls | avg ms | total ms | % |", "|---|---:|---:|---:|---:|"]
    for r in rows:
        lines.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.3f} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.1f} |")
    return "\n".join(lines)


def _per_kernel(path: str, counter: str):
    """Average value per dispatch of each kernel, and the largest single
    dispatch value of each."""
    tot = collections.defaultdict(float)
    top = collections.defaultdict(float)
    cnt = collections.Counter()
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        k = _short(r["Kernel_Name"])
        v = float(r["Counter_Value"])
        tot[k] += v
        top[k] = max(top[k], v)
        cnt[k] += 1
    return {k: tot[k] / cnt[k] for k in tot}, dict(top)


def _copy_scale(top: dict) -> float:
    """Bytes per counter unit, from the 1 GiB device copy (the largest
    single copy dispatch in the run)."""
    copies = {k: v for k, v in top.items() if "copyBuffer" in k}
    if not copies or max(copies.values()) <= 0:
        raise SystemExit("no copy kernel in the counter file (run bench.py with --copy-calib)")
    return COPY_BYTES / max(copies.values())


def traffic(fetch_path: str, write_path: str) -> dict:
    fetch, fetch_top = _per_kernel(fetch_path, "FETCH_SIZE")
    write, write_top = _per_kernel(write_path, "WRITE_SIZE")
    fscale = _copy_scale(fetch_top)
    wscale = _copy_scale(write_top)
    out = {"calibration": {"fetch_bytes_per_unit": fscale, "write_bytes_per_unit": wscale,
                           "copy_bytes": COPY_BYTES}}
    for k in sorted(set(fetch) | set(write)):
        if "copyBuffer" in k or "elementwise" in k:
            continue
        fb = fetch.get(k, 0.0) * fscale
        wb = write.get(k, 0.0) * wscale
        out[k] = {"fetch_size_kib": fetch.get(k, 0.0), "write_size_kib": write.get(k, 0.0),
                  "fetch_bytes": fb, "write_bytes": wb, "hbm_bytes_per_launch": fb + wb}
    return out


def counters(path: str, only=()) -> str:
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    cnt = collections.defaultdict(collections.Counter)
    names = []
    for r in csv.DictReader(open(path)):
        k = _short(r["Kernel_Name"])
        if only and not any(o in k for o in only):
            continue
        c = r["Counter_Name"]
        if c not in names:
            names.append(c)
        per[k][c] += float(r["Counter_Value"])
        cnt[k][c] += 1
    lines = ["| kernel | " + " | ".join(names) + " | derived |", "|---|" + "---:|" * (len(names) + 1)]
    for k in sorted(per):
        avg = {c: per[k][c] / max(cnt[k][c], 1) for c in names}
        derived = []
        wave = avg.get("SQ_WAVE_CYCLES", 0.0)
        if wave > 0:
            if "SQ_WAIT_ANY" in avg:
                derived.append(f"waiting {avg['SQ_WAIT_ANY'] / wave:.2f}")
            if "SQ_WAIT_INST_ANY" in avg:
                derived.append(f"wait {avg['SQ_WAIT_INST_ANY'] / wave:.2f}")
            if "SQ_ACTIVE_INST_VALU" in avg:
                derived.append(f"VALU {avg['SQ_ACTIVE_INST_VALU'] / wave:.2f}")
        hit, miss = avg.get("TCC_HIT_sum", avg.get("TCC_HIT")), avg.get("TCC_MISS_sum", avg.get("TCC_MISS"))
        if hit is not None and miss is not None and hit + miss > 0:
            derived.append(f"L2 hit {hit / (hit + miss):.2f}")
        lines.append(f"| `{k}` | " + " | ".join(f"{avg[c]:.4g}" for c in names) + " | " +
                     ", ".join(derived) + " |")
    return "\n".join(lines)


def main(argv) -> int:
    if len(argv) >= 2 and argv[0] == "stats":
        print(kernel_stats(argv[1], argv[2] if len(argv) > 2 else ""))
        return 0
    if len(argv) == 3 and argv[0] == "traffic":
        print(json.dumps(traffic(argv[1], argv[2]), indent=1))
        return 0
    if len(argv) >= 2 and argv[0] == "counters":
        print(counters(argv[1], tuple(argv[2:])))
        return 0
    print(__doc__, file=sys.stderr)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
