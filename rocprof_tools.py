"""HBM traffic per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE, calibrated on a plain copy.

`python3 rocprof_tools.py run` launches a streaming copy of COPY_BYTES (read
COPY_BYTES, write COPY_BYTES) and then `ntts` forward NTTs of 2^log_n through
ssip_ntt; run it once under `rocprofv3 --pmc FETCH_SIZE` and once under
`rocprofv3 --pmc WRITE_SIZE` (one counter set per run).

`python3 rocprof_tools.py summarize FETCH.csv WRITE.csv OUT.json OUT.md`
averages each counter per kernel, scales it by (known copy bytes) / (what the
counter reads for the copy), and writes bytes per launch.  A kernel whose
traffic is well above its algorithmic bytes re-reads data.

`python3 rocprof_tools.py valu NTT.csv MSM.csv OUT.json OUT.md` averages
rocprofv3's derived VALUBusy (percent of time the vector ALUs issue) and
VALUUtilization (percent of active lanes) per kernel, from runs of
`rocprofv3 --pmc VALUBusy VALUUtilization` over bench.py's NTT and MSM legs.

`python3 rocprof_tools.py trace DIR TITLE COMMAND OUT.md [KERNEL N]` writes
the per-kernel table of a `rocprofv3 --kernel-trace --stats` run (DIR holds its
run_kernel_stats.csv) as markdown, with the command that made it, and the
average of KERNEL's last N dispatches (bench.py's timed region).

`python3 rocprof_tools.py timeline DIR TITLE COMMAND OUT.md` reads a
`rocprofv3 --kernel-trace --memory-copy-trace` run (DIR holds
run_kernel_trace.csv and run_memory_copy_trace.csv), cuts it after every
batch's window sums leave for the host, and tabulates the pieces that copy host data in and run
kernels (the end-to-end msm symbol): their span, the time the copy engines
and the kernels were busy, and how much of the copy time overlapped kernels;
then lists the events of the last piece with the most parts.

`python3 rocprof_tools.py counter CSV COUNTER TITLE COMMAND OUT.md [LAST]` tabulates
the mean of one counter per kernel and grid size from a `rocprofv3 --pmc` run,
and with LAST the values of the last LAST dispatches in order (also written
as JSON beside OUT.md).

`python3 rocprof_tools.py msmcall DIR WINDOWS CALLS OUT.md` appends to OUT.md
the average time per kernel of the first CALLS device-resident MSM calls
with WINDOWS windows in a kernel trace (calls cut after the copy of their
window sums; WINDOWS read from the reduction tail's grid).
"""
from __future__ import annotations

import csv
import json
import os
import sys
from collections import defaultdict

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "zkpoly-compiler_amd"))

COPY_BYTES = 512 << 20
COPY_KERNEL = "copy_u4_kernel"


def run(log_n: int = 24, ntts: int = 3) -> None:
    import ctypes

    from zkpoly_amd import hip
    from zkpoly_amd.field import mont_limbs, omega, random_elements
    from zkpoly_amd.ntt import gen_powers, load_lib
    from zkpoly_amd.runtime import Libs, PolyPtr

    libs = Libs()
    stream = hip.Stream(0)
    lib = load_lib(libs)
    lib.zkp_copy_u4.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulonglong, ctypes.c_void_p]
    lib.zkp_copy_u4.restype = ctypes.c_int
    src = hip.DeviceBuffer(COPY_BYTES)
    dst = hip.DeviceBuffer(COPY_BYTES)
    src.zero()
    hip.check(lib.zkp_copy_u4(dst.addr, src.addr, COPY_BYTES // 16, stream.raw()), "zkp_copy_u4")
    stream.sync()
    del src, dst
    n = 1 << log_n
    x = hip.DeviceBuffer(n * 32)
    x.upload(random_elements(n, 99))
    w = hip.DeviceBuffer(32)
    w.upload(mont_limbs(omega(log_n)))
    tw = hip.DeviceBuffer(n // 2 * 32)
    gen_powers(libs, tw.addr, n // 2, w.addr, stream)
    ptr = PolyPtr(x.addr, n, 0, 0, n)
    for _ in range(ntts):
        hip.check(lib.ssip_ntt(ptr, tw.addr, log_n, stream.raw(), 8, 8), "ssip_ntt")
    stream.sync()
    print(f"traffic workload done: copy {COPY_BYTES} B, {ntts} NTTs of 2^{log_n}", flush=True)


def _per_kernel(path: str, counter: str) -> dict:
    sums, counts = defaultdict(float), defaultdict(set)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            name = r["Kernel_Name"].split("(")[0].split("::")[-1]
            sums[name] += float(r["Counter_Value"])
            counts[name].add(r["Dispatch_Id"])
    return {k: sums[k] / len(counts[k]) for k in sums}


def summarize(fetch_csv: str, write_csv: str, out_json: str, out_md: str, log_n: int = 24) -> dict:
    fetch = _per_kernel(fetch_csv, "FETCH_SIZE")
    write = _per_kernel(write_csv, "WRITE_SIZE")
    f_fetch = COPY_BYTES / (fetch[COPY_KERNEL] * 1024)
    f_write = COPY_BYTES / (write[COPY_KERNEL] * 1024)
    alg = 64 * (1 << log_n)  # one read and one write of every 32-byte element
    out = {"copy_bytes": COPY_BYTES, "fetch_scale": round(f_fetch, 4), "write_scale": round(f_write, 4)}
    rows = []
    for k in sorted(set(fetch) | set(write)):
        fb = fetch.get(k, 0.0) * 1024 * f_fetch
        wb = write.get(k, 0.0) * 1024 * f_write
        out[k] = int(fb + wb)
        rows.append((k, fetch.get(k, 0.0), write.get(k, 0.0), fb, wb))
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    with open(out_md, "w") as f:
        f.write(f"# HBM traffic per launch (NTT 2^{log_n}, rocprofv3 FETCH_SIZE / WRITE_SIZE)\n\n")
        f.write("Commands (from the repository root on an MI355X; one counter per run):\n\n```\n")
        f.write("cd /tmp && export TMPDIR=/tmp && export PYTHONPATH=$REPO/zkpoly-compiler_amd\n")
        for c, d in (("FETCH_SIZE", "pmc_fetch"), ("WRITE_SIZE", "pmc_write")):
            f.write(f"rocprofv3 --pmc {c} --output-format csv -d $REPO/gpurun_out/{d} -o run -- "
                    "python3 -u rocprof_tools.py run\n")
        f.write("python3 rocprof_tools.py summarize $REPO/gpurun_out/pmc_fetch/run_counter_collection.csv "
                "$REPO/gpurun_out/pmc_write/run_counter_collection.csv profiles/pmc_traffic.json "
                "profiles/pmc_traffic_ntt_2e24.md\n```\n\n")
        f.write(f"Calibration: the copy kernel moves {COPY_BYTES} B each way; the counters read "
                f"{fetch[COPY_KERNEL]:.0f} KiB fetched and {write[COPY_KERNEL]:.0f} KiB written for it, so they "
                f"are scaled by {f_fetch:.4f} (fetch) and {f_write:.4f} (write).\n\n")
        f.write("| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | fetched B (scaled) | written B (scaled) | "
                "total / algorithmic 64n B |\n|---|---|---|---|---|---|\n")
        for k, fk, wk, fb, wb in rows:
            ratio = (fb + wb) / alg if k != COPY_KERNEL else (fb + wb) / (2 * COPY_BYTES)
            f.write(f"| `{k}` | {fk:.0f} | {wk:.0f} | {fb:.4g} | {wb:.4g} | {ratio:.3f} |\n")
    return out


def summarize_valu(ntt_csv: str, msm_csv: str, out_json: str, out_md: str) -> dict:
    out = {}
    lines = []
    for label, path in (("NTT 2^24 (bench.py --quick)", ntt_csv), ("MSM 2^20 (bench.py --msm-only)", msm_csv)):
        busy = _per_kernel(path, "VALUBusy")
        util = _per_kernel(path, "VALUUtilization")
        for k in sorted(busy, key=lambda k: -busy[k]):
            if k.startswith("__amd"):
                continue
            out[k] = {"valu_busy_pct": round(busy[k], 2), "valu_utilization_pct": round(util.get(k, 0.0), 2)}
            lines.append(f"| {label} | `{k}` | {busy[k]:.1f} | {util.get(k, 0.0):.1f} |")
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    with open(out_md, "w") as f:
        f.write("# VALU busy / utilisation per kernel (rocprofv3 derived counters)\n\n")
        f.write("Commands (from the repository root on an MI355X):\n\n```\ncd /tmp && export TMPDIR=/tmp\n")
        f.write("rocprofv3 --pmc VALUBusy VALUUtilization --output-format csv -d $REPO/gpurun_out/pmc_valu_ntt -o run "
                "-- python3 -u $REPO/bench.py --quick --no-cpu --steps 3 --warmup 1\n")
        f.write("rocprofv3 --pmc VALUBusy VALUUtilization --output-format csv -d $REPO/gpurun_out/pmc_valu_msm -o run "
                "-- python3 -u $REPO/bench.py --msm-only --no-cpu --steps 2 --warmup 1\n")
        f.write("python3 rocprof_tools.py valu $REPO/gpurun_out/pmc_valu_ntt/run_counter_collection.csv "
                "$REPO/gpurun_out/pmc_valu_msm/run_counter_collection.csv profiles/pmc_valu.json "
                "profiles/pmc_valu.md\n```\n\n")
        f.write("VALUBusy: percent of the kernel's time the vector ALUs issue instructions; VALUUtilization: percent "
                "of lanes active in those instructions (divergence).  The MSM row averages windows 12..18.\n\n")
        f.write("| workload | kernel | VALUBusy % | VALUUtilization % |\n|---|---|---|---|\n")
        f.write("\n".join(lines) + "\n")
    return out


def summarize_trace(trace_dir: str, title: str, command: str, out_md: str, tail_kernel: str = "",
                    tail_n: int = 0) -> None:
    with open(f"{trace_dir}/run_kernel_stats.csv") as f:
        stats = list(csv.DictReader(f))
    lines = [f"# {title}", "", "Command (from the repository root on an MI355X):", "", "```", command, "```", "",
             "## Kernel stats (all dispatches of the run)", "",
             "| kernel | calls | avg ms | min ms | max ms | % time |", "|---|---|---|---|---|---|"]
    for r in stats:
        name = r["Name"].replace("(anonymous namespace)::", "").split("(")[0]
        lines.append(f"| `{name}` | {r['Calls']} | {float(r['AverageNs']) / 1e6:.4f} | {float(r['MinNs']) / 1e6:.4f} | "
                     f"{float(r['MaxNs']) / 1e6:.4f} | {float(r['Percentage']):.1f} |")
    if tail_kernel and tail_n:
        # the run's last tail_n dispatches of one kernel: bench.py's timed region, after its autotuning
        with open(f"{trace_dir}/run_kernel_trace.csv") as f:
            durs = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6 for r in csv.DictReader(f)
                    if r["Kernel_Name"].split("(")[0].split("::")[-1] == tail_kernel]
        last = durs[-tail_n:]
        if last:
            lines += ["", f"Timed region (the last {len(last)} `{tail_kernel}` dispatches, after the autotuner's "
                      f"trial launches): {sum(last) / len(last):.4f} ms average, "
                      f"{min(last):.4f}-{max(last):.4f} ms."]
    with open(out_md, "w") as f:
        f.write("\n".join(lines) + "\n")


def _union(iv):
    """Total length of a union of [start, end) intervals, and the merged intervals."""
    out = []
    for a, b in sorted(iv):
        if out and a <= out[-1][1]:
            out[-1][1] = max(out[-1][1], b)
        else:
            out.append([a, b])
    return sum(b - a for a, b in out), out


def _overlap(xs, ys):
    tot = 0
    for a, b in xs:
        for c, d in ys:
            lo, hi = max(a, c), min(b, d)
            if hi > lo:
                tot += hi - lo
    return tot


def summarize_timeline(trace_dir: str, title: str, command: str, out_md: str) -> None:
    ev = []
    with open(f"{trace_dir}/run_kernel_trace.csv") as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"].split("(")[0].split("::")[-1]
            # the library's copy kernel (pinned host memory read by the device) is a host-to-device copy
            kind = "copy" if name.startswith("k_pull_host") else "kernel"
            ev.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), kind,
                       "host_to_device (k_pull_host)" if kind == "copy" else name))
    with open(f"{trace_dir}/run_memory_copy_trace.csv") as f:
        for r in csv.DictReader(f):
            kind = r["Direction"].replace("MEMORY_COPY_", "").lower()
            ev.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), "copy", kind))
    ev.sort()
    # a batch ends with the copy of its window sums to the host, after the
    # reduction's tail (a copy engine transfer, or a blit kernel for so few bytes)
    calls, cur, reduced = [], [], False
    for e in ev:
        cur.append(e)
        if e[3].startswith("k_reduce_tail"):
            reduced = True
        elif reduced and (e[3] == "device_to_host" or "copyBuffer" in e[3]):
            calls.append(cur)
            cur, reduced = [], False
    rows = []
    for c in calls:
        h2d = [(a, b) for a, b, k, n in c if k == "copy" and n.startswith("host_to_device")]
        ker = [(a, b) for a, b, k, n in c if k == "kernel"]
        parts = sum(1 for e in c if e[3] == "k_part_count")
        if not h2d or not parts:
            continue
        t0, t1 = min(e[0] for e in c), max(e[1] for e in c)
        cbusy, cu = _union(h2d)
        kbusy, ku = _union(ker)
        rows.append((c, parts, len(h2d), (t1 - t0) / 1e6, cbusy / 1e6, kbusy / 1e6, _overlap(cu, ku) / 1e6))
    lines = [f"# {title}", "", "Command (from the repository root on an MI355X):", "", "```", command, "```", "",
             "## End-to-end batches (host inputs copied in, kernels, window sums back)", "",
             "The trace is cut after every copy of a batch's window sums to the host; a batch that shares a "
             "run with others carries the run's copies in its first piece.  `copy busy`: time at least one "
             "host-to-device copy ran; `kernel busy`: time at least one kernel ran; `overlap`: copy time during "
             "which a kernel also ran.", "",
             "| piece | sorts (parts) | H2D copies | span ms | copy busy ms | kernel busy ms | overlap ms |",
             "|---|---|---|---|---|---|---|"]
    for i, (c, parts, ncopy, span, cb, kb, ov) in enumerate(rows):
        lines.append(f"| {i} | {parts} | {ncopy} | {span:.3f} | {cb:.3f} | {kb:.3f} | {ov:.3f} |")
    if rows:
        best = max(rows, key=lambda r: (r[1], rows.index(r)))
        c = best[0]
        t0 = min(e[0] for e in c)
        lines += ["", f"## Events of the last piece with {best[1]} sorts (us from its first event)", "",
                  "| start | duration | kind | name |", "|---|---|---|---|"]
        for a, b, k, n in c:
            lines.append(f"| {(a - t0) / 1e3:.1f} | {(b - a) / 1e3:.1f} | {k} | `{n}` |")
    with open(out_md, "w") as f:
        f.write("\n".join(lines) + "\n")


def summarize_counter(csv_path: str, counter: str, title: str, command: str, out_md: str, last: int = 0) -> None:
    vals = defaultdict(list)
    seq = []
    with open(csv_path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            name = r["Kernel_Name"].split("(")[0].split("::")[-1]
            vals[(name, int(r["Grid_Size"]))].append(float(r["Counter_Value"]))
            seq.append((int(r["Dispatch_Id"]), name, int(r["Grid_Size"]), float(r["Counter_Value"])))
    lines = [f"# {title}", "", "Command (from the repository root on an MI355X):", "", "```", command, "```", "",
             f"Mean `{counter}` per dispatch, by kernel and grid size (threads).", "",
             "| kernel | grid | dispatches | mean |", "|---|---|---|---|"]
    for (name, grid), v in sorted(vals.items(), key=lambda kv: (-sum(kv[1]) / len(kv[1]), kv[0])):
        lines.append(f"| `{name}` | {grid} | {len(v)} | {sum(v) / len(v):.4g} |")
    if last > 0:
        lines += ["", f"The last {last} dispatches, in order:", "", f"| dispatch | kernel | grid | `{counter}` |",
                  "|---|---|---|---|"]
        for d, name, grid, v in sorted(seq)[-last:]:
            lines.append(f"| {d} | `{name}` | {grid} | {v:.6g} |")
        # the same dispatches for bench.py (beside OUT.md, .json)
        with open(os.path.splitext(out_md)[0] + ".json", "w") as f:
            json.dump({"counter": counter, "last": [[d, name, grid, v] for d, name, grid, v in sorted(seq)[-last:]]},
                      f, indent=1)
    with open(out_md, "w") as f:
        f.write("\n".join(lines) + "\n")


def summarize_msm_call(trace_dir: str, windows: int, calls: int, out_md: str) -> None:
    ev = []
    with open(f"{trace_dir}/run_kernel_trace.csv") as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"].split("(")[0].split("::")[-1]
            ev.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), name, int(r["Grid_Size_X"])))
    h2d = []
    try:
        with open(f"{trace_dir}/run_memory_copy_trace.csv") as f:
            h2d = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in csv.DictReader(f)
                   if r["Direction"].endswith("HOST_TO_DEVICE")]
    except FileNotFoundError:
        pass
    ev.sort()
    pieces, cur, tail_grid = [], [], 0
    for e in ev:
        cur.append(e)
        if e[2].startswith("k_reduce_tail"):
            tail_grid = e[3]
        elif tail_grid and "copyBuffer" in e[2]:
            pieces.append((cur, tail_grid))
            cur, tail_grid = [], 0
    want = []
    for p, grid in pieces:
        t0, t1 = p[0][0], p[-1][1]
        if grid // 512 != windows or any(a < t1 and b > t0 for a, b in h2d) or p[0][2] == "k_mul_generator":
            continue
        want.append(p)
        if len(want) == calls:
            break
    per = defaultdict(float)
    order = []
    for p in want:
        for e in p:
            if e[2] not in per:
                order.append(e[2])
            per[e[2]] += (e[1] - e[0]) / 1e6 / len(want)
    span = sum((p[-1][1] - p[0][0]) / 1e6 for p in want) / max(1, len(want))
    lines = ["", f"## One device-resident call with {windows} windows (average of {len(want)} calls)", "",
             f"First kernel start to the window sums' copy: {span:.4f} ms; kernels in order of first launch, "
             "time summed over the launches of one call.", "", "| kernel | ms per call |", "|---|---|"]
    lines += [f"| `{n}` | {per[n]:.4f} |" for n in order]
    lines.append(f"| all kernels | {sum(per.values()):.4f} |")
    with open(out_md, "a") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    if len(sys.argv) >= 2 and sys.argv[1] == "run":
        run()
    elif len(sys.argv) >= 6 and sys.argv[1] == "summarize":
        print(json.dumps(summarize(*sys.argv[2:6]), indent=1))
    elif len(sys.argv) >= 6 and sys.argv[1] == "valu":
        print(json.dumps(summarize_valu(*sys.argv[2:6]), indent=1))
    elif len(sys.argv) >= 6 and sys.argv[1] == "trace":
        summarize_trace(*sys.argv[2:6], *(sys.argv[6:7] + [int(a) for a in sys.argv[7:8]]))
    elif len(sys.argv) >= 7 and sys.argv[1] == "counter":
        summarize_counter(*sys.argv[2:7], *[int(a) for a in sys.argv[7:8]])
    elif len(sys.argv) >= 6 and sys.argv[1] == "msmcall":
        summarize_msm_call(sys.argv[2], int(sys.argv[3]), int(sys.argv[4]), sys.argv[5])
    elif len(sys.argv) >= 6 and sys.argv[1] == "timeline":
        summarize_timeline(*sys.argv[2:6])
    else:
        print(__doc__)
        sys.exit(2)
