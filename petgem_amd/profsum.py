"""Text summaries of rocprofv3 CSV output (kernel statistics, HBM traffic).

``stats``: the top kernels of a ``--kernel-trace --stats`` run
(``<prefix>_kernel_stats.csv``) as a markdown table.

``traffic``: HBM bytes per launch of every kernel from two ``--pmc`` runs of the
same program, one collecting ``FETCH_SIZE`` and one ``WRITE_SIZE``
(``<prefix>_counter_collection.csv``, KiB per dispatch).  Each counter is
scaled by what it reads for the program's device copy of a known size
(``--copy-bytes`` bytes read and written: the largest ``copyBuffer``
dispatch), since the counters do not see every memory channel equally.

Usage::

    python -m petgem_amd.profsum stats prof/kb_kernel_stats.csv --top 20
    python -m petgem_amd.profsum traffic --fetch F.csv --write W.csv --copy-bytes 1073741824
"""
import argparse
import csv
import re
import sys
from collections import defaultdict


def short_name(name):
    """`void pg::k_spmv_er<16, 4, 2, 1, HIP_vector_type<float, 2u>, ...>(args)` ->
    `k_spmv_er<16,4,2,1,float2,float2>`."""
    s = re.sub(r"^void ", "", name.strip())
    depth, cut = 0, len(s)
    for i, ch in enumerate(s):  # drop the argument list (first '(' outside <>)
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    s = s[:cut]
    s = re.sub(r"HIP_vector_type<(float|double|int), (\d)u>", r"\g<1>\2", s)
    s = re.sub(r"\bpg::((batch|eop)::)?", "", s)
    return s.replace(" ", "")


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _key(row, *cands):
    low = {k.lower(): k for k in row}
    for c in cands:
        if c.lower() in low:
            return low[c.lower()]
    raise KeyError(f"none of {cands} in {list(row)}")


def stats(path, top):
    rows = _rows(path)
    kn, kc, kt, ka = (_key(rows[0], "Name", "KernelName"), _key(rows[0], "Calls"),
                      _key(rows[0], "TotalDurationNs"), _key(rows[0], "AverageNs"))
    total = sum(float(r[kt]) for r in rows)
    rows.sort(key=lambda r: -float(r[kt]))
    out = ["| kernel | calls | avg (us) | total (ms) | % |", "|---|---:|---:|---:|---:|"]
    for r in rows[:top]:
        out.append(f"| `{short_name(r[kn])}` | {int(float(r[kc]))} | {float(r[ka]) / 1e3:.1f} | "
                   f"{float(r[kt]) / 1e6:.2f} | {100 * float(r[kt]) / total:.1f} |")
    return "\n".join(out)


def _per_kernel(path):
    rows = _rows(path)
    kn = _key(rows[0], "Kernel_Name", "KernelName", "kernel_name")
    kv = _key(rows[0], "Counter_Value", "CounterValue", "value")
    kg = _key(rows[0], "Grid_Size", "GridSize", "grid_size")
    agg = defaultdict(list)
    copy = defaultdict(list)
    for r in rows:
        v = float(r[kv])
        agg[short_name(r[kn])].append(v)
        if "copyBuffer" in r[kn]:
            copy[int(float(r[kg]))].append(v)
    return agg, copy


def traffic(fetch_csv, write_csv, copy_bytes, match):
    fa, fc = _per_kernel(fetch_csv)
    wa, wc = _per_kernel(write_csv)
    grid = max(wc, key=lambda g: sum(wc[g]) / len(wc[g]))  # the big copy
    f_kib = sum(fc[grid]) / len(fc[grid])
    w_kib = sum(wc[grid]) / len(wc[grid])
    sf, sw = copy_bytes / (1024 * f_kib), copy_bytes / (1024 * w_kib)
    out = [f"copy of {copy_bytes} B: FETCH_SIZE {f_kib:.1f} KiB -> factor {sf:.3f}; "
           f"WRITE_SIZE {w_kib:.1f} KiB -> factor {sw:.3f}", "",
           "| kernel | dispatches | FETCH_SIZE (KiB) | WRITE_SIZE (KiB) | HBM bytes / launch |",
           "|---|---:|---:|---:|---:|"]
    for k in sorted(fa):
        if match and not re.search(match, k):
            continue
        f = sum(fa[k]) / len(fa[k])
        w = sum(wa.get(k, [0.0])) / max(1, len(wa.get(k, [])))
        out.append(f"| `{k}` | {len(fa[k])} | {f:.1f} | {w:.1f} | "
                   f"{int(round(1024 * (sf * f + sw * w)))} |")
    return "\n".join(out)


def parse_traffic_table(text):
    """{kernel: HBM bytes per launch} from a table written by ``traffic``."""
    out = {}
    for line in text.splitlines():
        m = re.match(r"\|\s*`([^`]+)`\s*\|\s*\d+\s*\|[^|]*\|[^|]*\|\s*(\d+)\s*\|", line)
        if m:
            out[m.group(1)] = int(m.group(2))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--top", type=int, default=20)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--match", default=None, help="regex on kernel names to keep")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(stats(a.csv, a.top))
    else:
        print(traffic(a.fetch, a.write, a.copy_bytes, a.match))
    return 0


if __name__ == "__main__":
    sys.exit(main())
