"""Reduce rocprofv3 counter dumps to HBM bytes per launch of each step kernel.

    python -m torchwi_amd.profiling FETCH_CSV WRITE_CSV COPY_MIB CELLS_PER_LAUNCH OUT.json WORKLOAD [KEY]

FETCH_CSV / WRITE_CSV are the ``*_counter_collection.csv`` files of two
``rocprofv3 --pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE`` runs of the same
command (``bench.py ... --copy-probe COPY_MIB``).  The copy probe moves a
known number of bytes through a plain device copy in the same process; its
counter readings give the factor between each counter and real bytes, which
is then applied to the step kernels.  Kernel kinds are recognised from the
step-flag template argument of their names (3 forward, 7 forward with
virtual-source storage, 5 recomputation, 24 adjoint); the temporally blocked
kernel (``k_pipe<NE, FLAGS, T>``) is compared with its own compulsory bytes
per launch, the one-step kernel (``k_vstep``) with the per-unit bytes of one
cell-update.  With KEY, the kernels go under that key of OUT.json (which may
already hold other keys, e.g. one per storage mode of the same workload).
"""
import csv
import json
import re
import sys

KIND_BY_FLAGS = {3: "fwd", 7: "fwd_virt", 5: "recomp", 24: "adj"}
# one-step kernels: algorithmic bytes per cell-update
ALGORITHMIC = {"fwd": 16, "fwd_virt": 20, "recomp": 20, "adj": 28}
_NAME = re.compile(r"k_(vstep|pipe)<\d+, (\d+)u(?:, (\d+))?")


def pipe_algorithmic(kind, T):
    """Compulsory bytes per cell of one blocked launch of T steps (fields in
    and out once, model arrays, T virtual sources, one gradient
    read-modify-write)."""
    return {"fwd": 20, "fwd_virt": 24 + 4 * T, "recomp": 24 + 4 * T, "adj": 28 + 4 * T}[kind]


def _rows(path, counter):
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                out.append((r["Kernel_Name"], float(r["Counter_Value"]), int(r["Grid_Size"])))
    return out


def _kind(name):
    """(kind, steps per launch or 0 for the one-step kernel) of a step-kernel
    name, or None."""
    m = _NAME.search(name)
    if not m:
        return None
    kind = KIND_BY_FLAGS.get(int(m.group(2)))
    if kind is None:
        return None
    return kind, (int(m.group(3)) if m.group(1) == "pipe" and m.group(3) else 0)


def _per_kind(rows):
    """Mean counter value per dispatch of each step-kernel kind.  Graph nodes
    that lie past the end of the time loop (the tail of a partial block or
    segment) return at once and read almost nothing; dispatches below a tenth
    of the kind's median are those and are left out."""
    vals = {}
    for name, v, grid in rows:
        k = _kind(name)
        if k is not None:
            vals.setdefault(k, []).append(v)
    out = {}
    for k, v in vals.items():
        med = sorted(v)[len(v) // 2]
        keep = [x for x in v if x >= 0.1 * med]
        out[k] = sum(keep) / len(keep)
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
    for key in sorted(set(fk) & set(wk)):
        kind, T = key
        rb, wb = fk[key] * 1024.0 * ff, wk[key] * 1024.0 * wf
        alg = pipe_algorithmic(kind, T) if T else ALGORITHMIC[kind]
        kernels[kind] = {"read_bytes": int(rb), "write_bytes": int(wb), "total_bytes": int(rb + wb),
                         "bytes_per_cell": round((rb + wb) / cells_per_launch, 2),
                         "algorithmic_bytes_per_cell": alg,
                         "steps_per_launch": T or 1}
    return {"workload": workload,
            "calibration": {"copy_bytes_read": copy_bytes, "copy_bytes_written": copy_bytes,
                            "FETCH_SIZE_KiB": fc, "WRITE_SIZE_KiB": wc,
                            "fetch_factor": round(ff, 4), "write_factor": round(wf, 4)},
            "cells_per_launch": cells_per_launch, "kernels": kernels}


def main(argv=None):
    a = sys.argv[1:] if argv is None else argv
    if len(a) not in (6, 7):
        print(__doc__.strip().splitlines()[2], file=sys.stderr)
        return 2
    out = summarize(a[0], a[1], int(a[2]), int(a[3]), a[5])
    if len(a) == 7:
        try:
            with open(a[4]) as f:
                merged = json.load(f)
        except (OSError, ValueError):
            merged = {}
        merged.update({k: v for k, v in out.items() if k != "kernels"})
        merged[a[6]] = out["kernels"]
        out_doc = merged
    else:
        out_doc = out
    with open(a[4], "w") as f:
        json.dump(out_doc, f, indent=1)
    print(json.dumps(out["kernels"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
