"""Per-variant summary of a rocprofv3 run of the large step, against the byte model.

    python -m mpas_dycore_pytorch_amd.profile_summary DIR --kernel k_acoustic_cells \\
        [--level 7 --levels 55 --scalars 1]

``DIR`` is a rocprofv3 output directory (searched recursively) of a run that
launched ``kernel`` only inside whole large steps (``bench.py`` with its
derivative and mixed-precision legs off).  Each dispatch of the kernel is
labelled with its variant by its position in the step's launch order
(``bytes_model.launch_schedule``: the first or a later acoustic substep, the
first or a later RK stage), then:

* from a ``--kernel-trace`` CSV: launches, mean duration and algorithmic GB/s
  per variant, and the launch-weighted whole;
* from ``--pmc FETCH_SIZE`` / ``WRITE_SIZE`` runs: the counters per variant
  (KiB per dispatch, as rocprofv3 reports them), and bytes after dividing by
  what the same counter reads per byte of the plain copy kernel ``k_copy`` of
  the same run;
* with ``--counter NAME`` (repeatable): any other counter of a ``--pmc`` run,
  its mean per dispatch and variant, uncalibrated, beside the same counter's
  mean for ``k_copy`` per KiB copied;
* with ``--rate NAME``: for every kernel of a ``--pmc`` run, the counter's
  mean per dispatch divided by the dispatch's duration, per CU and clock
  (256 CUs at 2.4 GHz), the kernels with the most counts first.

rocprofv3's rocpd databases (``*_results.db``) and its CSV output are both read.

The figures are printed as a Markdown table.
"""

from __future__ import annotations

import argparse
import csv
import glob
import os
import re
import sys

from . import bytes_model as B

ICOSA = {lv: (10 * 4 ** lv + 2, 30 * 4 ** lv, 20 * 4 ** lv) for lv in range(0, 10)}


def _rows(directory: str, suffix: str) -> list[dict]:
    out = []
    for path in sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True)):
        with open(path, newline="") as f:
            out.extend(csv.DictReader(f))
    return out


def _db_rows(directory: str, query: str) -> list[tuple]:
    """Rows of a query over every rocpd database (``*_results.db``, rocprofv3's default output)."""
    import sqlite3

    out = []
    for path in sorted(glob.glob(os.path.join(directory, "**", "*_results.db"), recursive=True)):
        con = sqlite3.connect(f"file:{path}?mode=ro", uri=True)
        try:
            out.extend(con.execute(query).fetchall())
        finally:
            con.close()
    return out


def _matches(kernel: str):
    pat = re.compile(r"(^|[\s:*&])" + re.escape(kernel) + r"(\(|<|$|\s)")
    return lambda name: bool(pat.search(name or ""))


def dispatches(directory: str, kernel: str) -> list[dict]:
    """Kernel-trace rows of ``kernel`` in dispatch order (rocpd database or CSV)."""
    hit = _matches(kernel)
    rows = [{"Kernel_Name": n, "Dispatch_Id": d, "Start_Timestamp": a, "End_Timestamp": b}
            for n, d, a, b in _db_rows(directory, "select name, dispatch_id, start, end from kernels")]
    rows += _rows(directory, "kernel_trace.csv")
    rows = [r for r in rows if hit(r.get("Kernel_Name", ""))]
    rows.sort(key=lambda r: (int(r.get("Dispatch_Id") or 0), int(r.get("Start_Timestamp") or 0)))
    return rows


def counters(directory: str, kernel: str, counter: str) -> list[float]:
    """Per-dispatch values of one counter for ``kernel`` in dispatch order (rocpd database or CSV)."""
    hit = _matches(kernel)
    rows = [{"Kernel_Name": n, "Dispatch_Id": d, "Counter_Name": c, "Counter_Value": v}
            for n, d, c, v in _db_rows(directory, "select name, dispatch_id, counter_name, counter_value "
                                                  "from pmc_events")]
    rows += _rows(directory, "counter_collection.csv")
    rows = [r for r in rows if hit(r.get("Kernel_Name", "")) and r.get("Counter_Name") == counter]
    # a block counter (SQ_*) comes as one record per hardware instance: sum a dispatch's records
    per: dict[int, float] = {}
    for r in rows:
        d = int(r.get("Dispatch_Id") or r.get("Correlation_Id") or 0)
        per[d] = per.get(d, 0.0) + float(r["Counter_Value"])
    return [per[d] for d in sorted(per)]


def copy_calibration(directory: str, counter: str) -> float | None:
    """The counter's reading per byte moved by the plain copy kernel ``k_copy`` of the same run
    (each thread reads and writes one double; copies of at least 1 Mi doubles), or None."""
    grid = {d: g for n, d, g in _db_rows(directory, "select name, dispatch_id, grid_x from kernels")
            if _matches("k_copy")(n)}
    vals: dict[int, float] = {}
    for n, d, c, v in _db_rows(directory, "select name, dispatch_id, counter_name, counter_value from pmc_events"):
        if _matches("k_copy")(n) and c == counter:
            vals[d] = vals.get(d, 0.0) + float(v)
    ratios = sorted(vals[d] * 1024.0 / (grid[d] * 8.0) for d in vals if d in grid and grid[d] >= 1 << 20)
    return ratios[len(ratios) // 2] if ratios else None


def label(n: int, kernel: str) -> list[str]:
    sched = B.launch_schedule(kernel)
    if n % len(sched):
        raise ValueError(f"{n} dispatches of {kernel} are not whole large steps of {len(sched)} launches")
    return sched * (n // len(sched))


def raw_counters(directory: str, kernel: str, names: list[str]) -> str:
    """Mean per dispatch and variant of each named counter (as rocprofv3 reports it)."""
    lines = ["| counter | variant | dispatches | mean per dispatch | k_copy per KiB copied |", "|---|---|---|---|---|"]
    for counter in names:
        vals = counters(directory, kernel, counter)
        if not vals:
            lines.append(f"| {counter} | - | 0 | - | - |")
            continue
        labels = label(len(vals), kernel)
        rate = copy_calibration(directory, counter)
        rs = "-" if rate is None else f"{rate:.4g}"
        for v in sorted(set(labels)):
            vs = [x for x, lv in zip(vals, labels) if lv == v]
            lines.append(f"| {counter} | {v} | {len(vs)} | {sum(vs) / len(vs):.6g} | {rs} |")
        lines.append(f"| {counter} | all | {len(vals)} | {sum(vals) / len(vals):.6g} | {rs} |")
    return "\n".join(lines)


def summarize(directory: str, kernel: str, level: int, levels: int, scalars: int) -> str:
    nC, nE, nV = ICOSA[level]
    dims = (nC, nE, nV, levels, scalars)
    lines = [f"| variant | launches | mean us | algorithmic bytes | GB/s | fraction of 8 TB/s |", "|---|---|---|---|---|---|"]
    rows = dispatches(directory, kernel)
    total_b = total_t = 0.0
    if rows:
        labels = label(len(rows), kernel)
        dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9 for r in rows]
        for v in sorted(set(labels)):
            ts = [d for d, lv in zip(dur, labels) if lv == v]
            b = B.kernel_bytes(kernel, *dims, variant=v)
            mean = sum(ts) / len(ts)
            total_b += b * len(ts)
            total_t += sum(ts)
            lines.append(f"| {v} | {len(ts)} | {mean * 1e6:.1f} | {b:.4e} | {b / mean / 1e9:.1f} | "
                         f"{b / mean / 8e12:.4f} |")
        lines.append(f"| all (launch-weighted) | {len(rows)} | {total_t / len(rows) * 1e6:.1f} | "
                     f"{total_b / len(rows):.4e} | {total_b / total_t / 1e9:.1f} | {total_b / total_t / 8e12:.4f} |")
    out = ["\n".join(lines)]
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        vals = counters(directory, kernel, counter)
        if not vals:
            continue
        labels = label(len(vals), kernel)
        cal = copy_calibration(directory, counter)
        cl = [f"| variant | dispatches | {counter} mean KiB | calibrated bytes (k_copy factor "
              f"{cal if cal is None else round(cal, 4)}) |", "|---|---|---|---|"]
        for v in sorted(set(labels)):
            vs = [x for x, lv in zip(vals, labels) if lv == v]
            mean = sum(vs) / len(vs)
            cb = f"{mean * 1024 / cal:.4e}" if cal else "-"
            cl.append(f"| {v} | {len(vs)} | {mean:.0f} | {cb} |")
        mean = sum(vals) / len(vals)
        cl.append(f"| all | {len(vals)} | {mean:.0f} | {(f'{mean * 1024 / cal:.4e}' if cal else '-')} |")
        out.append("\n".join(cl))
    return "\n\n".join(out)


def rates(directory: str, counter: str, n: int, cus: int = 256, ghz: float = 2.4) -> str:
    """Per kernel: mean count per dispatch, mean duration and count per CU per clock."""
    kern = {d: (name, (end - start) * 1e-9, grid) for name, d, start, end, grid in
            _db_rows(directory, "select name, dispatch_id, start, end, grid_x from kernels")}
    per: dict[int, float] = {}
    for name, d, c, v in _db_rows(directory, "select name, dispatch_id, counter_name, counter_value from pmc_events"):
        if c == counter:
            per[d] = per.get(d, 0.0) + float(v)
    groups: dict[str, list[tuple[float, float]]] = {}
    for d, v in per.items():
        if d not in kern:
            continue
        name, sec, grid = kern[d]
        short = re.sub(r"\(.*", "", name).replace("mpas::", "")
        if short == "k_copy" and grid < 1 << 20:
            continue  # small copies are launch-bound
        groups.setdefault(short, []).append((v, sec))
    lines = [f"| kernel | dispatches | {counter} per dispatch | mean us | per CU per clock |", "|---|---|---|---|---|"]
    order = sorted(groups.items(), key=lambda kv: -sum(v for v, _ in kv[1]))
    for short, vals in order[:n]:
        v = sum(a for a, _ in vals) / len(vals)
        t = sum(b for _, b in vals) / len(vals)
        lines.append(f"| {short} | {len(vals)} | {v:.4g} | {t * 1e6:.1f} | {v / t / cus / (ghz * 1e9):.3f} |")
    return "\n".join(lines)


def top_kernels(directory: str, n: int) -> str:
    """rocprofv3's kernel statistics, the ``n`` largest by total time: the
    ``top_kernels`` view of a rocpd database (microseconds), or the
    ``*kernel_stats.csv`` of ``--output-format csv`` (nanoseconds)."""
    rows = [(name, calls, float(total), float(avg), pct) for name, calls, total, avg, pct in
            _db_rows(directory, "select name, total_calls, total_duration, average, percentage from top_kernels")]
    for path in sorted(glob.glob(os.path.join(directory, "**", "*kernel_stats.csv"), recursive=True)):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                rows.append((r["Name"], r["Calls"], float(r["TotalDurationNs"]) / 1e3, float(r["AverageNs"]) / 1e3,
                             r["Percentage"]))
    rows.sort(key=lambda r: -r[2])
    lines = ["| kernel | calls | total ms | mean us | % |", "|---|---|---|---|---|"]
    for name, calls, total, avg, pct in rows[:n]:
        short = re.sub(r"\(.*", "", name).replace("mpas::", "").replace("adj::", "").replace("tan::", "")
        lines.append(f"| {short} | {calls} | {total / 1e3:.2f} | {avg:.1f} | {float(pct):.1f} |")
    return "\n".join(lines)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("directory")
    ap.add_argument("--kernel", default="k_acoustic_cells")
    ap.add_argument("--level", type=int, default=7)
    ap.add_argument("--levels", type=int, default=55)
    ap.add_argument("--scalars", type=int, default=1)
    ap.add_argument("--top", type=int, default=0, help="also print the N largest kernels by total time")
    ap.add_argument("--counter", action="append", default=[],
                    help="also print this --pmc counter per dispatch and variant (repeatable)")
    ap.add_argument("--rate", default="", help="print this --pmc counter per CU per clock for the top kernels")
    a = ap.parse_args(argv)
    if a.rate:
        print(rates(a.directory, a.rate, a.top or 20))
        print()
    if a.kernel:
        print(summarize(a.directory, a.kernel, a.level, a.levels, a.scalars))
    if a.kernel and a.counter:
        print()
        print(raw_counters(a.directory, a.kernel, a.counter))
    if a.top:
        print()
        print(top_kernels(a.directory, a.top))
    return 0


if __name__ == "__main__":
    sys.exit(main())
