"""Helpers to turn rocprofv3 CSV output into the short summaries kept under profiles/.

``python -m witwin_maxwell_amd.profiling stats <dir>`` prints per-kernel call
counts and average durations from a ``--kernel-trace --stats`` run;
``python -m witwin_maxwell_amd.profiling traffic <fetch_dir> <write_dir>`` turns
two ``--pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` runs into HBM bytes per dispatch,
scaled by what the same counters read for torch's plain copy kernel of a known
size (the run must include bench.py's ``--copy-probe``).
"""

from __future__ import annotations

import csv
import glob
import json
import os
import sys
from collections import defaultdict

COPY_PROBE_BYTES = 256 * 1024 * 1024  # bench.py --copy-probe copies this many bytes per dispatch


def _find(directory, suffix):
    hits = sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {directory}")
    return hits


def kernel_stats(directory):
    rows = []
    for path in _find(directory, "kernel_stats.csv"):
        with open(path) as fh:
            rows.extend(csv.DictReader(fh))
    out = {}
    for r in rows:
        name = r.get("Name") or r.get("KernelName") or r.get("Kernel_Name")
        calls = int(float(r.get("Calls", 0)))
        avg = float(r.get("AverageNs", r.get("Average", 0)))
        tot = float(r.get("TotalDurationNs", r.get("TotalDuration", 0)))
        out[name] = {"calls": calls, "avg_us": avg / 1e3, "total_ms": tot / 1e6}
    return out


def _short(name):
    """Kernel family of a rocprofv3 kernel name (template arguments kept for the
    step kernels, whose variants are timed separately)."""
    if "__amd_rocclr_copyBuffer" in name:
        return "copy"
    for key in ("lean_step_kernel", "fused_step_kernel", "lean_batch_kernel", "fused_batch_kernel"):
        if key in name:
            return name[name.index(key):].split("(")[0]
    for key in ("update_e_kernel", "update_h_kernel", "apply_terms_kernel", "accumulate_planes_kernel",
                "accumulate_points_kernel", "accumulate_kernel", "shifted_values_kernel"):
        if key in name:
            return key
    return name[:60]


def counter_per_dispatch(directory, counter):
    """Average counter value per dispatch of each kernel family; for the copy
    probe, the largest dispatch (the probe's 256 MiB copy, not the small
    device-to-device copies a run also makes)."""
    vals = defaultdict(list)
    for path in _find(directory, "counter_collection.csv"):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") != counter:
                    continue
                vals[_short(r.get("Kernel_Name", ""))].append(float(r["Counter_Value"]))
    return {k: (max(v) if k == "copy" else sum(v) / len(v)) for k, v in vals.items()}


def traffic(fetch_dir, write_dir):
    """HBM bytes per dispatch: FETCH_SIZE + WRITE_SIZE (KiB) of each kernel, each
    scaled by bytes / counter reading of the copy probe, which reads and writes
    COPY_PROBE_BYTES."""
    fetch = counter_per_dispatch(fetch_dir, "FETCH_SIZE")
    write = counter_per_dispatch(write_dir, "WRITE_SIZE")
    if not fetch.get("copy") or not write.get("copy"):
        raise RuntimeError("the runs hold no copy-probe dispatch (run bench.py with --copy-probe)")
    f_scale = COPY_PROBE_BYTES / (fetch["copy"] * 1024.0)
    w_scale = COPY_PROBE_BYTES / (write["copy"] * 1024.0)
    out = {"_calibration": {"fetch_scale": f_scale, "write_scale": w_scale}}
    for k in fetch:
        if k == "copy":
            continue
        f = fetch[k] * 1024.0 * f_scale
        w = write.get(k, 0.0) * 1024.0 * w_scale
        out[k] = {"fetch_bytes": f, "write_bytes": w, "total_bytes": f + w}
    return out


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        for k, v in sorted(kernel_stats(sys.argv[2]).items(), key=lambda kv: -kv[1]["total_ms"]):
            print(f"{v['calls']:8d} calls  avg {v['avg_us']:10.2f} us  total {v['total_ms']:10.2f} ms  {k}")
    elif sys.argv[1] == "traffic":
        print(json.dumps(traffic(sys.argv[2], sys.argv[3]), indent=1))
