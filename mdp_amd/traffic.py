"""HBM traffic per kernel launch from rocprofv3 counter runs.

Two subcommands:

``python mdp_amd/traffic.py copy --bytes B --reps R``
    Runs R elementwise copies (dst = 1.0 * src) of B bytes, one kernel each.  Profiled with
    ``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` it calibrates the counters: each launch reads
    and writes exactly B bytes.

``python mdp_amd/traffic.py summarize --kernel REGEX --fetch DIR --write DIR --copy-fetch DIR
--copy-write DIR --copy-bytes B --out FILE``
    Reads the counter CSV files of the four runs, averages FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) over the dispatches of the kernels matching REGEX, scales each by the factor the same
    counter reads for the calibration copy (true bytes / counted bytes), and writes a JSON summary
    whose ``bytes_per_launch`` (read + write) is what bench.py reports as ``roofline.traffic``.
"""
import argparse
import csv
import glob
import json
import os
import re
import sys


def _rows(d):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    for f in files:
        with open(f, newline="") as fh:
            yield from csv.DictReader(fh)


def per_dispatch(d, counter, kernel_re):
    """Mean and count of ``counter`` (summed over its instances within a dispatch) over the
    dispatches whose kernel name matches kernel_re."""
    pat = re.compile(kernel_re)
    by_disp = {}
    for r in _rows(d):
        if r.get("Counter_Name") != counter or not pat.search(r.get("Kernel_Name", "")):
            continue
        key = (r.get("Agent_Id"), r.get("Dispatch_Id"))
        by_disp[key] = by_disp.get(key, 0.0) + float(r["Counter_Value"])
    if not by_disp:
        raise SystemExit(f"no {counter} rows for kernels matching {kernel_re!r} under {d}")
    vals = list(by_disp.values())
    return sum(vals) / len(vals), len(vals)


def cmd_copy(a):
    import torch
    n = a.bytes // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(a.reps):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(f"copied {a.reps} x {a.bytes} bytes", flush=True)


def cmd_summarize(a):
    f_k, nf = per_dispatch(a.fetch, "FETCH_SIZE", a.kernel)
    w_k, nw = per_dispatch(a.write, "WRITE_SIZE", a.kernel)
    cf_k, _ = per_dispatch(a.copy_fetch, "FETCH_SIZE", a.copy_kernel)
    cw_k, _ = per_dispatch(a.copy_write, "WRITE_SIZE", a.copy_kernel)
    sf = a.copy_bytes / (cf_k * 1024.0)
    sw = a.copy_bytes / (cw_k * 1024.0)
    read_b = f_k * 1024.0 * sf
    write_b = w_k * 1024.0 * sw
    out = dict(kernel_regex=a.kernel, dispatches_fetch=nf, dispatches_write=nw,
               fetch_kib_per_dispatch=f_k, write_kib_per_dispatch=w_k,
               copy_bytes=a.copy_bytes, copy_fetch_kib=cf_k, copy_write_kib=cw_k,
               fetch_scale=sf, write_scale=sw, read_bytes_per_launch=read_b, write_bytes_per_launch=write_b,
               bytes_per_launch=read_b + write_b, algorithmic_bytes_per_launch=a.algorithmic)
    if a.algorithmic:
        out["traffic_over_algorithmic"] = (read_b + write_b) / a.algorithmic
    with open(a.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


def main(argv=None):
    ap = argparse.ArgumentParser(description="HBM traffic per launch from rocprofv3 counters")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--bytes", type=int, default=1 << 30)
    c.add_argument("--reps", type=int, default=3)
    s = sub.add_parser("summarize")
    s.add_argument("--kernel", required=True)
    s.add_argument("--fetch", required=True)
    s.add_argument("--write", required=True)
    s.add_argument("--copy-fetch", required=True)
    s.add_argument("--copy-write", required=True)
    s.add_argument("--copy-bytes", type=int, default=1 << 30)
    s.add_argument("--copy-kernel", default="elementwise|Mul|mul")
    s.add_argument("--algorithmic", type=float, default=0.0, help="algorithmic bytes per launch")
    s.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    (cmd_copy if a.cmd == "copy" else cmd_summarize)(a)


if __name__ == "__main__":
    sys.exit(main())
