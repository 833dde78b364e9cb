"""Summaries of rocprofv3 CSV output: per-kernel time, overlap between the
host threads that launched kernels, and hardware counters per dispatch.

    python -m vhci_amd.tracestats kernels TRACE.csv [--top N]
    python -m vhci_amd.tracestats overlap TRACE.csv
    python -m vhci_amd.tracestats counters COUNTERS.csv [--kernels REGEX]

TRACE.csv is a `--kernel-trace --output-format csv` file (*_kernel_trace.csv),
COUNTERS.csv a `--pmc` file (*_counter_collection.csv).  The serving loop
runs each request slot on its own host thread and HIP stream, so the launching
thread (Thread_Id) names the slot a kernel belongs to; `overlap` reports how
much of the busy time had kernels of two or more slots in flight at once.
Output is Markdown.
"""
import argparse
import csv
import re
import sys

import numpy as np


def short_name(name, width=96):
    """Kernel name without namespaces, return type and parameter list
    (template arguments kept), cut to `width` characters."""
    s = name.strip().replace("(anonymous namespace)::", "")
    if s.startswith("void "):
        s = s[5:]
    depth = 0
    for i, ch in enumerate(s):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            s = s[:i]
            break
    if s.startswith("vhci::"):
        s = s[6:]
    return s if len(s) <= width else s[:width - 3] + "..."


def read_trace(path):
    """(start ns, end ns, thread id, name index) arrays and the name list."""
    names, index = [], {}
    st, en, th, ki = [], [], [], []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            n = short_name(r["Kernel_Name"])
            k = index.get(n)
            if k is None:
                k = index[n] = len(names)
                names.append(n)
            st.append(int(r["Start_Timestamp"]))
            en.append(int(r["End_Timestamp"]))
            th.append(int(r["Thread_Id"]))
            ki.append(k)
    return (np.array(st, dtype=np.int64), np.array(en, dtype=np.int64), np.array(th, dtype=np.int64),
            np.array(ki, dtype=np.int64), names)


def kernel_table(path, top=25):
    st, en, _, ki, names = read_trace(path)
    dur = (en - st).astype(np.float64)
    tot = np.bincount(ki, weights=dur, minlength=len(names))
    cnt = np.bincount(ki, minlength=len(names))
    order = np.argsort(-tot)
    allt = dur.sum()
    out = ["| kernel | calls | total ms | avg us | share |", "|---|---:|---:|---:|---:|"]
    for k in order[:top]:
        avg = tot[k] / cnt[k] / 1e3
        out.append("| `%s` | %d | %.1f | %.1f | %.1f %% |" % (names[k], cnt[k], tot[k] / 1e6, avg,
                                                             100.0 * tot[k] / allt))
    out.append("")
    out.append("%d dispatches of %d kernels, %.1f ms of kernel time" % (len(st), len(names), allt / 1e6))
    return "\n".join(out)


def overlap_table(path):
    """Sweep over kernel start/end events: busy time (union of intervals),
    time with 2+ kernels in flight, and time with kernels of 2+ launching
    threads in flight."""
    st, en, th, _, _ = read_trace(path)
    if len(st) == 0:
        return "no dispatches"
    tids, tix = np.unique(th, return_inverse=True)
    t = np.concatenate([st, en])
    d = np.concatenate([np.ones(len(st), np.int64), -np.ones(len(en), np.int64)])
    who = np.concatenate([tix, tix])
    # ends before starts at equal times: touching kernels do not overlap
    order = np.lexsort((d, t))
    t, d, who = t[order], d[order], who[order]
    per = np.zeros(len(tids), np.int64)
    active = 0
    threads_on = 0
    busy = multi = cross = 0
    max_k = max_t = 0
    prev = int(t[0])
    for i in range(len(t)):
        now = int(t[i])
        dt = now - prev
        if active > 0:
            busy += dt
        if active > 1:
            multi += dt
        if threads_on > 1:
            cross += dt
        prev = now
        w = who[i]
        if d[i] > 0:
            if per[w] == 0:
                threads_on += 1
            per[w] += 1
            active += 1
        else:
            per[w] -= 1
            if per[w] == 0:
                threads_on -= 1
            active -= 1
        max_k = max(max_k, active)
        max_t = max(max_t, threads_on)
    span = int(en.max() - st.min())
    ksum = int((en - st).sum())
    rows = [
        ("dispatches", "%d" % len(st)),
        ("launching threads", "%d" % len(tids)),
        ("span (first start to last end)", "%.3f s" % (span / 1e9)),
        ("busy (at least one kernel in flight)", "%.3f s (%.1f %% of span)" % (busy / 1e9, 100.0 * busy / span)),
        ("sum of kernel durations", "%.3f s (%.2f x busy)" % (ksum / 1e9, ksum / max(busy, 1))),
        ("two or more kernels in flight", "%.3f s (%.1f %% of busy)" % (multi / 1e9, 100.0 * multi / max(busy, 1))),
        ("kernels of two or more threads in flight",
         "%.3f s (%.1f %% of busy)" % (cross / 1e9, 100.0 * cross / max(busy, 1))),
        ("most kernels in flight at once", "%d" % max_k),
        ("most threads with kernels in flight at once", "%d" % max_t),
    ]
    out = ["| quantity | value |", "|---|---|"] + ["| %s | %s |" % r for r in rows]
    return "\n".join(out)


def counter_table(path, kernels=None):
    """Mean counter value per dispatch, per kernel and counter."""
    pat = re.compile(kernels) if kernels else None
    acc = {}
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            n = short_name(r["Kernel_Name"])
            if pat and not pat.search(n):
                continue
            key = (n, r["Counter_Name"])
            s, c = acc.get(key, (0.0, 0))
            acc[key] = (s + float(r["Counter_Value"]), c + 1)
    out = ["| kernel | counter | dispatches | mean per dispatch |", "|---|---|---:|---:|"]
    for (n, c), (s, k) in sorted(acc.items()):
        out.append("| `%s` | %s | %d | %.1f |" % (n, c, k, s / k))
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("what", choices=["kernels", "overlap", "counters"])
    ap.add_argument("path")
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--kernels", default=None, help="regular expression on kernel names (counters)")
    a = ap.parse_args(argv)
    if a.what == "kernels":
        print(kernel_table(a.path, a.top))
    elif a.what == "overlap":
        print(overlap_table(a.path))
    else:
        print(counter_table(a.path, a.kernels))
    return 0


if __name__ == "__main__":
    sys.exit(main())
