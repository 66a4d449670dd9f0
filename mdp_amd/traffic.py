"""HBM traffic per kernel launch from rocprofv3 counter runs.

Three subcommands:

``python mdp_amd/traffic.py requests --kernel REGEX --rdreq DIR --wrreq DIR --out FILE``
    Exact bytes from the L2-to-fabric request counts by size, no calibration: read bytes =
    128 TCC_EA0_RDREQ_128B + 64 TCC_EA0_RDREQ_64B + 32 TCC_EA0_RDREQ_32B (the three sizes add up to
    TCC_EA0_RDREQ, which the summary checks), write bytes = 64 TCC_EA0_WRREQ_64B + 32 (TCC_EA0_WRREQ -
    TCC_EA0_WRREQ_64B).  DIR are the output folders of ``rocprofv3 --pmc TCC_EA0_RDREQ_sum
    TCC_EA0_RDREQ_32B_sum TCC_EA0_RDREQ_64B_sum TCC_EA0_RDREQ_128B_sum`` and ``rocprofv3 --pmc
    TCC_EA0_WRREQ_sum TCC_EA0_WRREQ_64B_sum``.  On a 1 GiB copy these give 8,388,608 128-byte reads and
    16,777,216 64-byte writes, exactly 1 GiB each way.  This is the route the committed summaries use.

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
    FETCH_SIZE counts 128-byte reads through TCC_BUBBLE, which reads 0 on this part, so one scale factor
    cannot fit kernels that mix 64- and 128-byte reads; kept for comparison with ``requests``.
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


def per_dispatch_sum(d, counters, kernel_re):
    """{counter: mean over dispatches} for the dispatches whose kernel name matches kernel_re."""
    pat = re.compile(kernel_re)
    by_disp = {}
    for r in _rows(d):
        if r.get("Counter_Name") not in counters or not pat.search(r.get("Kernel_Name", "")):
            continue
        key = (r.get("Agent_Id"), r.get("Dispatch_Id"))
        row = by_disp.setdefault(key, {c: 0.0 for c in counters})
        row[r["Counter_Name"]] += float(r["Counter_Value"])
    if not by_disp:
        raise SystemExit(f"no {counters} rows for kernels matching {kernel_re!r} under {d}")
    n = len(by_disp)
    return {c: sum(v[c] for v in by_disp.values()) / n for c in counters}, n


def cmd_requests(a):
    rd = ["TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum"]
    wr = ["TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum"]
    r, nr = per_dispatch_sum(a.rdreq, rd, a.kernel)
    w, nw = per_dispatch_sum(a.wrreq, wr, a.kernel)
    read_b = 128.0 * r["TCC_EA0_RDREQ_128B_sum"] + 64.0 * r["TCC_EA0_RDREQ_64B_sum"] + 32.0 * r["TCC_EA0_RDREQ_32B_sum"]
    sized = r["TCC_EA0_RDREQ_128B_sum"] + r["TCC_EA0_RDREQ_64B_sum"] + r["TCC_EA0_RDREQ_32B_sum"]
    write_b = 64.0 * w["TCC_EA0_WRREQ_64B_sum"] + 32.0 * (w["TCC_EA0_WRREQ_sum"] - w["TCC_EA0_WRREQ_64B_sum"])
    out = dict(kernel_regex=a.kernel, method="request counts by size", dispatches_read=nr, dispatches_write=nw,
               read_requests=r, write_requests=w, read_sizes_cover_requests=sized / max(r["TCC_EA0_RDREQ_sum"], 1.0),
               read_bytes_per_launch=read_b, write_bytes_per_launch=write_b, bytes_per_launch=read_b + write_b,
               algorithmic_bytes_per_launch=a.algorithmic)
    if a.algorithmic:
        out["traffic_over_algorithmic"] = (read_b + write_b) / a.algorithmic
    if a.ms:
        out["ms_per_launch"] = a.ms
        out["achieved_gb_s"] = (read_b + write_b) / (a.ms * 1e-3) / 1e9
        if a.algorithmic:
            out["algorithmic_gb_s"] = a.algorithmic / (a.ms * 1e-3) / 1e9
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
    q = sub.add_parser("requests")
    q.add_argument("--kernel", required=True)
    q.add_argument("--rdreq", required=True)
    q.add_argument("--wrreq", required=True)
    q.add_argument("--algorithmic", type=float, default=0.0, help="algorithmic bytes per launch")
    q.add_argument("--ms", type=float, default=0.0, help="average launch duration (kernel trace), for GB/s")
    q.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    {"copy": cmd_copy, "summarize": cmd_summarize, "requests": cmd_requests}[a.cmd](a)


if __name__ == "__main__":
    sys.exit(main())
