"""Text summaries of rocprofv3 result databases (kernel trace and counters).

    python -m pastis_amd.rocprof_summary trace DB [--top N]
    python -m pastis_amd.rocprof_summary counters DB COUNTER [--scale S]
    python -m pastis_amd.rocprof_summary table DB COUNTER [COUNTER ...]

`trace` lists kernels by total duration (calls, total/average ms, share).
`counters` lists the per-dispatch average of one counter per kernel,
multiplied by `--scale` (the calibration factor measured on a copy kernel of
known size).  `table` lists the per-dispatch averages of several counters
collected in one pass, one column per counter.
"""
import argparse
import re
import sqlite3
import sys


def short(name, width=72):
    name = re.sub(r"\(anonymous namespace\)::", "", name)
    name = re.sub(r"\(pa::[^)]*\)|\(pa_pair[^)]*\)|\(unsigned long[^)]*\)", "",
                  name)
    return name[:width]


def trace(db, top):
    c = sqlite3.connect(db)
    rows = list(c.execute("select name, total_calls, total_duration, average, "
                          "percentage from top_kernels"))
    out = ["%-72s %6s %10s %10s %6s" % ("kernel", "calls", "total_ms",
                                         "avg_ms", "%")]
    # top_kernels durations are in microseconds
    for name, calls, tot, avg, pct in rows[:top]:
        out.append("%-72s %6d %10.3f %10.4f %6.2f" %
                   (short(name), calls, tot / 1e3, avg / 1e3, pct))
    return "\n".join(out)


def counters(db, counter, scale):
    c = sqlite3.connect(db)
    rows = list(c.execute(
        "select kernel_name, count(*), avg(value), sum(value) from "
        "counters_collection where counter_name = ? group by kernel_name "
        "order by sum(value) desc", (counter,)))
    out = ["%-72s %6s %16s %16s" % ("kernel", "disp", "avg_" + counter,
                                    "scaled_avg")]
    for name, n, avg, tot in rows:
        out.append("%-72s %6d %16.1f %16.1f" %
                   (short(name), n, avg, avg * scale))
    return "\n".join(out)


def table(db, names, top):
    c = sqlite3.connect(db)
    rows = list(c.execute(
        "select kernel_name, counter_name, count(*), avg(value) from "
        "counters_collection group by kernel_name, counter_name"))
    per = {}
    disp = {}
    for k, n, cnt, avg in rows:
        per.setdefault(k, {})[n] = avg
        disp[k] = max(disp.get(k, 0), cnt)
    order = sorted(per, key=lambda k: -per[k].get(names[0], 0.0))[:top]
    out = ["%-60s %5s " % ("kernel", "disp") +
           " ".join("%18s" % n[:18] for n in names)]
    for k in order:
        out.append("%-60s %5d " % (short(k, 60), disp[k]) +
                   " ".join("%18.1f" % per[k].get(n, float("nan"))
                            for n in names))
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    t.add_argument("--top", type=int, default=20)
    k = sub.add_parser("counters")
    k.add_argument("db")
    k.add_argument("counter")
    k.add_argument("--scale", type=float, default=1.0)
    tb = sub.add_parser("table")
    tb.add_argument("db")
    tb.add_argument("names", nargs="+")
    tb.add_argument("--top", type=int, default=20)
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        print(trace(a.db, a.top))
    elif a.cmd == "table":
        print(table(a.db, a.names, a.top))
    else:
        print(counters(a.db, a.counter, a.scale))


if __name__ == "__main__":
    sys.exit(main())
