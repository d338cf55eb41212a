"""Summaries of rocprofv3 runs of bench.py (kernel traces and HBM counters).

    python -m violet_amd.rocprof_summary traffic FETCH_DB WRITE_DB OUT.json
    python -m violet_amd.rocprof_summary kernels TRACE_DB OUT.md

`traffic` reads two counter runs of `bench.py --calibration-copy` (one with
`--pmc FETCH_SIZE`, one with `--pmc WRITE_SIZE`; both counters are KiB per
dispatch).  The calibration copy moves a known number of bytes with a plain
device-to-device copy, so the ratio of those bytes to what the counters read
for it is the scale applied to every other kernel's counters.  The output
maps each library kernel to its calibrated HBM bytes read and written per
dispatch; bench.py reports them as the roofline's `traffic`.

`kernels` turns a `--kernel-trace` run into a per-kernel table of calls and
durations.
"""
from __future__ import annotations

import json
import sqlite3
import sys

CALIBRATION_BYTES = 256 << 20  # bench.py --calibration-copy: bytes per copy


def _counter_means(db_path: str, counter: str) -> dict[str, tuple[float, int]]:
    db = sqlite3.connect(db_path)
    rows = db.execute("select kernel_name, avg(value), count(*) from counters_collection where counter_name = ? "
                      "group by kernel_name", (counter,)).fetchall()
    return {name: (float(v), int(n)) for name, v, n in rows}


def _short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "")
    return name.split("(")[0].replace("void ", "")


def traffic(fetch_db: str, write_db: str, out: str) -> dict:
    f = _counter_means(fetch_db, "FETCH_SIZE")
    w = _counter_means(write_db, "WRITE_SIZE")
    copy = [k for k in f if "copy" in k.lower() or "elementwise" in k.lower()]
    if not copy:
        raise SystemExit("no calibration copy kernel found in the FETCH_SIZE run")
    ck = max(copy, key=lambda k: f[k][0])
    fetch_scale = CALIBRATION_BYTES / (f[ck][0] * 1024.0)
    cw = [k for k in w if k == ck] or [k for k in w if "copy" in k.lower() or "elementwise" in k.lower()]
    write_scale = CALIBRATION_BYTES / (w[cw[0]][0] * 1024.0) if cw else 1.0
    res = {"calibration_kernel": _short(ck), "calibration_bytes": CALIBRATION_BYTES,
           "fetch_scale": fetch_scale, "write_scale": write_scale, "kernels": {}}
    for k, (v, n) in f.items():
        if "vgi::" not in k:
            continue
        wb = w.get(k, (0.0, 0))[0]
        res["kernels"][_short(k)] = {"dispatches": n, "fetch_bytes": v * 1024.0 * fetch_scale,
                                     "write_bytes": wb * 1024.0 * write_scale}
    with open(out, "w") as fh:
        json.dump(res, fh, indent=1, sort_keys=True)
    return res


def kernels(trace_db: str, out: str) -> None:
    db = sqlite3.connect(trace_db)
    rows = db.execute("select name, count(*), avg(end-start)/1000.0, min(end-start)/1000.0, max(end-start)/1000.0, "
                      "sum(end-start)/1e6 from kernels group by name order by sum(end-start) desc").fetchall()
    tot = sum(r[5] for r in rows) or 1.0
    lines = ["| kernel | calls | avg us | min us | max us | total ms | share |", "|---|---|---|---|---|---|---|"]
    for r in rows:
        lines.append(f"| `{_short(r[0])}` | {r[1]} | {r[2]:.1f} | {r[3]:.1f} | {r[4]:.1f} | {r[5]:.2f} | {r[5] / tot:.1%} |")
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    if len(sys.argv) == 5 and sys.argv[1] == "traffic":
        r = traffic(sys.argv[2], sys.argv[3], sys.argv[4])
        print(json.dumps({k: v for k, v in r.items() if k != "kernels"}))
    elif len(sys.argv) == 4 and sys.argv[1] == "kernels":
        kernels(sys.argv[2], sys.argv[3])
    else:
        raise SystemExit(__doc__)
