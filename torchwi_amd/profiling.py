"""Reduce rocprofv3 counter dumps to HBM bytes per launch of each step kernel.

    python -m torchwi_amd.profiling FETCH_CSV WRITE_CSV COPY_MIB CELLS_PER_LAUNCH OUT.json WORKLOAD

FETCH_CSV / WRITE_CSV are the ``*_counter_collection.csv`` files of two
``rocprofv3 --pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE`` runs of the same
command (``bench.py ... --copy-probe COPY_MIB``).  The copy probe moves a
known number of bytes through a plain device copy in the same process; its
counter readings give the factor between each counter and real bytes, which
is then applied to the step kernels.  Kernel kinds are recognised from the
step-flag template argument of their names (3 forward, 7 forward with
virtual-source storage, 5 recomputation, 24 adjoint).
"""
import csv
import json
import re
import sys

KIND_BY_FLAGS = {3: "fwd", 7: "fwd_virt", 5: "recomp", 24: "adj"}
ALGORITHMIC = {"fwd": 16, "fwd_virt": 20, "recomp": 20, "adj": 28}
_FLAGS = re.compile(r"k_(?:vstep|pipe)<\d+, (\d+)u")


def _rows(path, counter):
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                out.append((r["Kernel_Name"], float(r["Counter_Value"]), int(r["Grid_Size"])))
    return out


def _per_kind(rows):
    """Mean counter value per dispatch of each step-kernel kind.  Graph nodes
    that lie past the end of the time loop (the tail of a partial block or
    segment) return at once and read almost nothing; dispatches below a tenth
    of the kind's median are those and are left out."""
    vals = {}
    for name, v, grid in rows:
        m = _FLAGS.search(name)
        if not m:
            continue
        kind = KIND_BY_FLAGS.get(int(m.group(1)))
        if kind is not None:
            vals.setdefault(kind, []).append(v)
    out = {}
    for kind, v in vals.items():
        med = sorted(v)[len(v) // 2]
        keep = [x for x in v if x >= 0.1 * med]
        out[kind] = sum(keep) / len(keep)
    return out


def _copy(rows):
    """Counter value (KiB) of the largest copy kernel dispatch."""
    c = [v for name, v, grid in rows if "copyBuffer" in name or "copy" in name.lower()]
    return max(c) if c else None


def summarize(fetch_csv, write_csv, copy_mib, cells_per_launch, workload):
    fr, wr = _rows(fetch_csv, "FETCH_SIZE"), _rows(write_csv, "WRITE_SIZE")
    copy_bytes = copy_mib * (1 << 20)
    fc, wc = _copy(fr), _copy(wr)
    if not fc or not wc:
        raise RuntimeError("no copy-probe dispatch found in the counter dumps")
    ff, wf = copy_bytes / (fc * 1024.0), copy_bytes / (wc * 1024.0)
    fk, wk = _per_kind(fr), _per_kind(wr)
    kernels = {}
    for kind in sorted(set(fk) & set(wk)):
        rb, wb = fk[kind] * 1024.0 * ff, wk[kind] * 1024.0 * wf
        kernels[kind] = {"read_bytes": int(rb), "write_bytes": int(wb), "total_bytes": int(rb + wb),
                         "bytes_per_cell": round((rb + wb) / cells_per_launch, 2),
                         "algorithmic_bytes_per_cell": ALGORITHMIC[kind]}
    return {"workload": workload,
            "calibration": {"copy_bytes_read": copy_bytes, "copy_bytes_written": copy_bytes,
                            "FETCH_SIZE_KiB": fc, "WRITE_SIZE_KiB": wc,
                            "fetch_factor": round(ff, 4), "write_factor": round(wf, 4)},
            "cells_per_launch": cells_per_launch, "kernels": kernels}


def main(argv=None):
    a = sys.argv[1:] if argv is None else argv
    if len(a) != 6:
        print(__doc__.strip().splitlines()[2], file=sys.stderr)
        return 2
    out = summarize(a[0], a[1], int(a[2]), int(a[3]), a[5])
    with open(a[4], "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out["kernels"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
