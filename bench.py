#!/usr/bin/env python3
"""Benchmark: BN254 NTT 2^24 Gelem/s on one MI355X (plus the NTT suite, MSM
2^20 and the batched form when built), bit-exact against the CPU oracle.

A step is one in-place forward NTT of 2^24 BN254 Fr elements (synthetic,
uniform, resident in HBM) through the reference's ssip_ntt C ABI.  The timed
region is K steps between two events on the launch stream; per-kernel times
come from events recorded between the kernels of those same launches.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]
"""
import argparse
import ctypes
import json
import os
import sys
import time

# HIP gives a process four hardware queues by default and maps further streams
# onto them, so two streams that share a queue run one after the other.  The
# legs below create a dozen streams (the MSM's compute and two copy streams,
# the serving slots, the pipelined kernel's streams); with eight queues the
# MSM's copies still overlap its sort and accumulation after the other legs
# have made theirs.  Set before anything loads the HIP runtime.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "zkpoly-compiler_amd")
for _p in (PKG, ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# 256 CUs x 64 lanes x 2.4 GHz: one v_mad_u64_u32 per lane per clock
INT_MUL_PEAK = 256 * 64 * 2.4e9
MADS_PER_MODMUL = 136


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cores():
    """Cores this process may run on (its affinity set), and what the OS reports."""
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return usable, os.cpu_count() or usable


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log-n", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--quick", action="store_true", help="headline NTT only (no suite / MSM / batched)")
    ap.add_argument("--msm-only", action="store_true", help="MSM 2^20 leg only (for profiling)")
    ap.add_argument("--batched-only", action="store_true", help="config-4 batched leg only")
    ap.add_argument("--serving-only", action="store_true", help="config-5 serving leg only")
    ap.add_argument("--serve-count", type=int, default=200, help="requests per rate in the serving leg")
    ap.add_argument("--msm-log-n", type=int, default=20)
    ap.add_argument("--poly-only", action="store_true", help="libpoly.so leg only")
    ap.add_argument("--pipeline-only", action="store_true", help="pipelined fused-kernel leg only")
    return ap.parse_args()


class Ntt:
    """Device-resident NTT workload of size 2^k through libntt.so."""

    def __init__(self, libs, stream, k, seed):
        from zkpoly_amd import hip
        from zkpoly_amd.field import mont_limbs, omega, random_elements
        from zkpoly_amd.ntt import gen_powers, load_lib
        from zkpoly_amd.runtime import PolyPtr, ScalarArray

        self.k, self.n = k, 1 << k
        self.lib = load_lib(libs)
        self.stream = stream
        self.host = random_elements(self.n, seed)
        self.buf = hip.DeviceBuffer(self.n * 32)
        self.buf.upload(self.host)
        self.x = ScalarArray(self.n, self.buf.addr)
        self.ptr = PolyPtr.from_array(self.x)
        self.tw = {}
        for inv in (False, True):
            wb = hip.DeviceBuffer(32)
            wb.upload(mont_limbs(omega(k, inv)))
            t = hip.DeviceBuffer(self.n // 2 * 32)
            gen_powers(libs, t.addr, self.n // 2, wb.addr, stream)
            self.tw[inv] = t
        nb = hip.DeviceBuffer(32)
        from zkpoly_amd.field import FR_MODULUS

        nb.upload(mont_limbs(pow(self.n, -1, FR_MODULUS)))
        self.ninv = nb
        stream.sync()

    def reset(self):
        self.buf.upload(self.host)

    def ntt(self, inverse=False):
        rc = self.lib.ssip_ntt(self.ptr, self.tw[inverse].addr, self.k, self.stream.raw(), 8, 8)
        if rc:
            raise RuntimeError(f"ssip_ntt rc={rc}")
        if inverse:
            rc = self.lib.distribute_powers(self.ptr, self.ninv.addr, 1, self.stream.raw())
            if rc:
                raise RuntimeError(f"distribute_powers rc={rc}")

    def download(self):
        return self.buf.download(np.uint64, self.n * 4).reshape(self.n, 4)


def time_loop(stream, fn, steps, warmup):
    from zkpoly_amd import hip

    for _ in range(warmup):
        fn()
    stream.sync()
    a, b = hip.Event(), hip.Event()
    a.record(stream)
    for _ in range(steps):
        fn()
    b.record(stream)
    b.sync()
    return a.elapsed_ms(b) / steps


def headline(libs, stream, args):
    """K timed forward NTTs of 2^log_n with per-kernel events."""
    from zkpoly_amd import autotune, hip
    from zkpoly_amd.runtime import PolyPtr  # noqa: F401

    # the project's autotuner picks the pass depth for this size (untimed, before the run)
    tuning = autotune.autotune(libs, stream, ntt_sizes=[args.log_n])
    w = Ntt(libs, stream, args.log_n, seed=2024)
    w.tuning = tuning
    lib = w.lib
    lib.zkp_ntt_with_events.argtypes = [
        type(w.ptr),
        ctypes.c_void_p,
        ctypes.c_uint,
        ctypes.c_void_p,
        ctypes.POINTER(ctypes.c_void_p),
        ctypes.c_int,
        ctypes.POINTER(ctypes.c_int),
        ctypes.POINTER(ctypes.c_int),
    ]
    lib.zkp_ntt_with_events.restype = ctypes.c_int
    max_ev = 16
    # bit-exact check of one step against the oracle (and CPU baseline timing)
    w.ntt()
    stream.sync()
    gpu_out = w.download()
    w.reset()
    for _ in range(args.warmup):
        w.ntt()
    stream.sync()
    events = [[hip.Event() for _ in range(max_ev)] for _ in range(args.steps)]
    ev_arrays = [(ctypes.c_void_p * max_ev)(*[e.handle.value for e in evs]) for evs in events]
    kinds = (ctypes.c_int * max_ev)()
    nk = ctypes.c_int(0)
    t0, t1 = hip.Event(), hip.Event()
    t0.record(stream)
    for s in range(args.steps):
        rc = lib.zkp_ntt_with_events(w.ptr, w.tw[False].addr, w.k, stream.raw(), ev_arrays[s], max_ev, kinds, ctypes.byref(nk))
        if rc:
            raise RuntimeError(f"zkp_ntt_with_events rc={rc}")
    t1.record(stream)
    t1.sync()
    total_ms = t0.elapsed_ms(t1)
    nkern = nk.value
    per_kind = {0: [], 1: []}
    for s in range(args.steps):
        for i in range(nkern):
            per_kind[kinds[i]].append(events[s][i].elapsed_ms(events[s][i + 1]))
    return w, gpu_out, total_ms / args.steps, per_kind, nkern


def suite(libs, stream, steps):
    """Config 2: forward / inverse / coset at 2^20, 2^22, 2^24 and the 4x extended coset of 2^22."""
    from zkpoly_amd import hip
    from zkpoly_amd.field import mont_limbs, omega
    from zkpoly_amd.ntt import GenPqOmegas, gen_powers
    from zkpoly_amd.runtime import ScalarArray

    out = {}
    for k in (20, 22, 24):
        w = Ntt(libs, stream, k, seed=3000 + k)
        n = w.n
        gb = hip.DeviceBuffer(32)
        gb.upload(mont_limbs(7))
        pw = hip.DeviceBuffer(n * 32)
        gen_powers(libs, pw.addr, n, gb.addr, stream)
        stream.sync()

        def coset():
            w.lib.distribute_powers(w.ptr, pw.addr, n, stream.raw())
            w.ntt()

        # recompute_ntt: twiddles formed from the pq table and the omega ladder
        # (no n/2 table), pq sized as the reference does (GenPqOmegas::get_pq_len)
        pq_len = GenPqOmegas.get_pq_len(k)
        pq = np.zeros((pq_len, 4), dtype=np.uint64)
        om = np.zeros((32, 4), dtype=np.uint64)
        GenPqOmegas(libs).get_fn()(ScalarArray(pq_len, pq.ctypes.data, device="cpu"),
                                   ScalarArray(32, om.ctypes.data, device="cpu"), n, mont_limbs(omega(k)))
        pqb, omb = hip.DeviceBuffer(pq.nbytes), hip.DeviceBuffer(om.nbytes)
        pqb.upload(pq)
        omb.upload(om)
        pq_deg = pq_len.bit_length()

        def recompute():
            hip.check(w.lib.recompute_ntt(w.ptr, pqb.addr, pq_deg, omb.addr, k, stream.raw(), 8, 8), "recompute_ntt")

        for name, fn in (("ntt", lambda: w.ntt()), ("intt", lambda: w.ntt(True)), ("coset_ntt", coset),
                         ("recompute_ntt", recompute)):
            ms = time_loop(stream, fn, steps, 2)
            out[f"{name}_2^{k}"] = suite_entry(n, k, ms, name)
        del w, pw, pqb, omb
    # 4x extended coset of a 2^22 polynomial: zero-pad to 2^24, coset g = 7, NTT on 2^24
    w = Ntt(libs, stream, 24, seed=4422)
    n = w.n
    w.host[1 << 22 :] = 0  # coefficients of a 2^22 polynomial, zero-padded to 2^24
    w.reset()
    gb = hip.DeviceBuffer(32)
    gb.upload(mont_limbs(7))
    pw = hip.DeviceBuffer(n * 32)
    gen_powers(libs, pw.addr, n, gb.addr, stream)
    stream.sync()

    def ext():
        w.lib.distribute_powers(w.ptr, pw.addr, n, stream.raw())
        w.ntt()

    ms = time_loop(stream, ext, steps, 2)
    e = suite_entry(n, 24, ms, "coset_ntt")
    # algorithmic bytes of the extended transform: the 2^22 coefficients in, 2^24 values out
    e["hbm_gbs"] = round(32 * ((1 << 22) + n) / (ms * 1e-3) / 1e9, 1)
    e["hbm_frac"] = round(e["hbm_gbs"] / HBM_PEAK_GBS, 4)
    e["gelem_s_out"] = e.pop("gelem_s")
    out["extended_coset_4x_2^22"] = e
    return out


def suite_entry(n, k, ms, name):
    """ms, Gelem/s, algorithmic HBM rate (one read and one write of every
    element) and integer-multiply use (136 v_mad_u64_u32 per product: n/2
    products per stage, plus one per element for the coset's powers or the
    inverse's n^-1 scaling) of one transform."""
    mads = MADS_PER_MODMUL * (n // 2) * k
    if name in ("intt", "coset_ntt"):
        mads += MADS_PER_MODMUL * n
    gbs = 64 * n / (ms * 1e-3) / 1e9
    return {"ms": round(ms, 4), "gelem_s": round(n / ms / 1e6, 3), "hbm_gbs": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "int_frac": round(mads / (ms * 1e-3) / INT_MUL_PEAK, 4)}


def msm_leg(libs, stream, log_n, steps, with_cpu):
    """Config 3: MSM of 2^log_n points P_i = k_i G, window sweep 12..18, uniform and 10 % / 10 % scalars."""
    from zkpoly_amd import hip
    from zkpoly_amd.field import mont_limbs, random_elements
    from zkpoly_amd.msm import msm_device, mul_generator, workspace_size

    n = 1 << log_n
    ks = random_elements(n, 3000)
    ss = random_elements(n, 3001)
    rng = np.random.Generator(np.random.PCG64(3002))
    perm = rng.permutation(n)
    ss01 = ss.copy()
    ss01[perm[: n // 10]] = 0
    ss01[perm[n // 10 : n // 5]] = mont_limbs(1)
    kb = hip.DeviceBuffer(n * 32)
    kb.upload(ks)
    pb = hip.DeviceBuffer(n * 64)
    mul_generator(libs, kb.addr, n, pb.addr, stream)
    sb = hip.DeviceBuffer(n * 32)
    sb.upload(ss)
    sb01 = hip.DeviceBuffer(n * 32)
    sb01.upload(ss01)
    stream.sync()
    out = {"log_n": log_n, "windows": {}}
    results = {}
    for window in range(12, 19):
        target = -(-254 // window)
        ws = hip.DeviceBuffer(workspace_size(libs, n, window, target))
        for scal, tag in ((sb, "uniform"), (sb01, "zeros_ones")):
            if tag == "zeros_ones" and window != 16:
                continue
            r = msm_device(libs, pb.addr, scal.addr, n, window, target, ws, stream)
            times = []
            for _ in range(steps):
                t = time.perf_counter()
                r = msm_device(libs, pb.addr, scal.addr, n, window, target, ws, stream)
                times.append((time.perf_counter() - t) * 1e3)
            times.sort()
            key = f"w{window}" if tag == "uniform" else f"w{window}_zeros_ones"
            out["windows"][key] = {"ms_median": round(times[len(times) // 2], 4), "ms_min": round(times[0], 4)}
            results[(window, tag)] = r
        del ws
    best = min((v["ms_median"], k) for k, v in out["windows"].items() if "zeros" not in k)
    out["best_window"] = best[1]
    out["ms"] = best[0]
    log("[bench] msm symbol end to end (host points and scalars in pinned memory)")
    out["reference_abi"] = msm_abi_leg(libs, pb, ss, n, steps, results[(16, "uniform")])
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        from zkpoly_amd.field import from_mont

        want = {}
        for tag, arr in (("uniform", ss), ("zeros_ones", ss01)):
            total = from_mont(orc.to_int(orc.inner_product(ks, arr)))
            want[tag] = list(orc.g1_mul_canonical(orc.generator(), total))
        out["bit_exact"] = all(list(r) == want[tag] for (w, tag), r in results.items())
        threads, reported = host_cores()
        bases = pb.download(np.uint64, n * 8).reshape(n, 8)
        log(f"[bench] cpu baseline: oracle best_multiexp 2^{log_n} on {threads} threads")
        t = time.perf_counter()
        cpu = orc.best_multiexp(ss, bases, threads)
        cpu_s = time.perf_counter() - t
        out["cpu_baseline"] = {"ms": round(cpu_s * 1e3, 2), "cores": threads, "kind": "port", "os_cpu_count": reported,
                               "sample": f"best_multiexp restatement on the same 2^{log_n} uniform instance, "
                                         f"all {threads} cores of the process's affinity set"}
        m1 = min(n, 1 << 17)
        t = time.perf_counter()
        orc.best_multiexp(ss[:m1], bases[:m1], 1)
        cpu1_s = time.perf_counter() - t
        out["cpu_baseline_1core"] = {"ms": round(cpu1_s * 1e3, 2), "cores": 1, "kind": "port", "points": m1,
                                     "sample": f"best_multiexp on the first 2^{m1.bit_length() - 1} points, one core"}
        out["bit_exact"] = out["bit_exact"] and list(cpu) == want["uniform"]
    return out


def msm_abi_leg(libs, pb, ss, n, steps, expected):
    """The drop-in msm symbol end to end: points and scalars in host memory
    from the runtime's pinned CPU pool, copied in by the library on every call
    (as the reference's msm does), the affine result back on the host.  One
    batch with the whole range in one part, the same in four staged parts
    (copies overlapping the accumulation), the planner's choice, and four
    batches sharing one pass over the points."""
    from zkpoly_amd import hip
    from zkpoly_amd.msm import MSM, MsmConfig
    from zkpoly_amd.msm_plan import plan_msm
    from zkpoly_amd.pool import CpuMemoryPool
    from zkpoly_amd.runtime import GpuBuffer, Point, PointArray, ScalarArray

    pool = CpuMemoryPool(22, 64, libs)  # 256 MiB pinned slabs
    pts = pool.allocate(n * 64)
    pool.view(pts, n * 64)[:] = pb.download(np.uint64, n * 8)
    sc = pool.allocate(n * 32)
    pool.view(sc, n * 32)[:] = ss.reshape(-1)
    planned = plan_msm(n, 1, 288 << 30, libs)
    cases = {
        "w16_one_part": (MsmConfig(16, 16), 1),
        "w16_four_parts_staged": (MsmConfig(16, 16, parts=4, stage_scalars=2, stage_points=2), 1),
        "planned": (planned, 1),
        "w16_four_batches_one_run": (MsmConfig(16, 16, batch_per_run=4), 4),
    }
    # the host link's rate on this machine, from the same pinned points and the way the library copies
    # (hipMemcpyAsync on a stream): every case moves 96 B per point
    h = hip._lib()
    link = hip.DeviceBuffer(n * 64)
    cs = hip.Stream(0)
    for k in range(5):
        if k == 1:
            t = time.perf_counter()
        hip.check(h.hipMemcpyAsync(ctypes.c_void_p(link.addr), ctypes.c_void_p(pts), n * 64,
                                   hip.hipMemcpyHostToDevice, cs.raw()), "link copy")
        cs.sync()
    out = {"link_h2d_gbs": round(4 * n * 64 / (time.perf_counter() - t) / 1e9, 2)}
    cs.destroy()
    del link
    for name, (cfg, batches) in cases.items():
        op = MSM(libs, cfg)
        size = op.get_buffer_size(n)[0]
        buf = hip.DeviceBuffer(size)
        res = [Point() for _ in range(batches)]
        f = op.get_fn().f
        args = ([GpuBuffer(buf.addr, size)] + res, [PointArray(pts, n)] + [ScalarArray(n, sc, "cpu")] * batches)
        f(*args, lambda c: c)
        times = []
        for _ in range(steps):
            t = time.perf_counter()
            f(*args, lambda c: c)
            times.append((time.perf_counter() - t) * 1e3)
        times.sort()
        out[name] = {"window": cfg.window_size, "target_window": cfg.target_window, "parts": cfg.parts,
                     "batch_per_run": cfg.batch_per_run, "stages": [cfg.stage_scalars, cfg.stage_points],
                     "batches": batches, "ms_median": round(times[len(times) // 2], 3),
                     "ms_per_batch": round(times[len(times) // 2] / batches, 3),
                     "device_bytes": size, "bit_exact_vs_device": all(list(r.data) == list(expected) for r in res)}
        del buf
    pool.destroy()
    return out


def batched_leg(libs, stream, steps, with_cpu):
    """Config 4: 64 x (NTT of log size U{10..22} + fused a*b + gamma*c - z), one hipGraph
    launch vs the same 64 x (ssip_ntt + fused call) issued one by one."""
    from zkpoly_amd import autotune, fused
    from zkpoly_amd.batch import Config4Workload
    from zkpoly_amd.ntt import kernel_count

    kernel = fused.FusedKernel(libs, fused.FusedKernelMeta("fused_arith_abgcz", 5, 1, fused.DEMO_DIR))
    log(f"[bench] autotuning NTT pass depth for log sizes 10..22")
    tuning = autotune.autotune(libs, stream, ntt_sizes=range(10, 23))
    w = Config4Workload(libs, stream, kernel, count=64, lo=10, hi=22, seed=4, tuning=tuning)
    graph = w.capture()
    plain_ms = time_loop(stream, w.run_plain, steps, 2)
    batched_ms = time_loop(stream, lambda: graph.launch(stream), steps, 2)
    out = {
        "instances": 64,
        "log_sizes": w.log_lens,
        "elements": w.total_elements,
        "plain_ms": round(plain_ms, 4),
        "batched_ms": round(batched_ms, 4),
        "plain_gelem_s": round(w.total_elements / plain_ms / 1e6, 4),
        "batched_gelem_s": round(w.total_elements / batched_ms / 1e6, 4),
        "speedup": round(plain_ms / batched_ms, 4),
        "launches_plain": sum(kernel_count(libs, k) + 1 for k in w.log_lens),
        "graph_launches_batched": 1,
        "autotune_ntt_min_col_log": {str(k): v for k, v in sorted(tuning.ntt.items())},
    }
    # the same 64 transforms and fused kernels plus 16 MSMs of 2^12..2^20 points in the one graph;
    # plain: every MSM its own synchronous call (device work, window sums out, host combine)
    log("[bench] batched form with MSMs (64 NTT + fused + 16 MSMs)")
    wm = Config4Workload(libs, stream, kernel, count=64, lo=10, hi=22, seed=4, tuning=tuning, msm_count=16,
                         msm_lo=12, msm_hi=20)
    gm = wm.capture()
    plain_m = time_loop(stream, wm.run_plain, steps, 2)

    def batched_with_msms():
        gm.launch(stream)
        stream.sync()
        gm.msm_results()

    batched_m = time_loop(stream, batched_with_msms, steps, 2)
    # the same graph with every MSM after the fused kernel on the capturing stream (no parallel branches)
    gs = wm.capture(branches=0)

    def batched_serial():
        gs.launch(stream)
        stream.sync()
        gs.msm_results()

    serial_m = time_loop(stream, batched_serial, steps, 2)
    work = wm.total_elements + wm.total_msm_points
    res_plain = wm.run_plain()
    stream.sync()
    gm.launch(stream)
    stream.sync()
    out["with_msms"] = {
        "msm_log_sizes": wm.msm_log_lens,
        "msm_windows": [m.window for m in wm.msms],
        "ntt_elements": wm.total_elements,
        "msm_points": wm.total_msm_points,
        "plain_ms": round(plain_m, 4),
        "batched_ms": round(batched_m, 4),
        "plain_gelem_s": round(work / plain_m / 1e6, 4),
        "batched_gelem_s": round(work / batched_m / 1e6, 4),
        "speedup": round(plain_m / batched_m, 4),
        "msm_graph_branches": gm.branches,
        "batched_one_branch_ms": round(serial_m, 4),
        "msm_results_equal_plain": [list(a) for a in res_plain] == [list(b) for b in gm.msm_results()],
    }
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc

        # bit-exact spot check of the smallest and largest instances after one fresh graph launch
        w.reset_inputs()
        graph.launch(stream)
        stream.sync()
        order = sorted(range(64), key=lambda i: w.lens[i])
        ok = True
        for i in (order[0], order[1], order[-1]):
            x, o = w.results(i)
            xh, a, b, c = w.host[i]
            n = w.lens[i]
            g = np.broadcast_to(w.gamma_z[0], (n, 4)).copy()
            z = np.broadcast_to(w.gamma_z[1], (n, 4)).copy()
            want = orc.vec("orc_vec_sub", orc.FR, orc.vec("orc_vec_add", orc.FR, orc.vec("orc_vec_mul", orc.FR, a, b),
                                                          orc.vec("orc_vec_mul", orc.FR, g, c)), z)
            ok = ok and np.array_equal(x, orc.ntt(xh, w.log_lens[i], host_cores()[0]))
            ok = ok and np.array_equal(o, want)
        out["bit_exact_spot_check"] = bool(ok)
    return out


def serving_leg(libs, stream, rates, count, with_cpu):
    """Config 5: Poisson streams (70 % NTT 2^14..2^22, 30 % MSM 2^16..2^21) through the
    continuous-batching engine; p50 / p99 latency and aggregate elements/s per rate."""
    from zkpoly_amd.serve import Server, Trace, latency_stats, make_resident

    res = make_resident(libs, stream, 14, 22, 21, seed=55)
    srv = Server(libs, res, slots=4, hbm_budget=16 << 30, msm_window=16)
    out = {"slots": 4, "hbm_budget_gib": 16, "msm_window": 16, "rates": {}}
    all_ok = True
    want = {}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        from zkpoly_amd.field import from_mont
    for rate in rates:
        trace = Trace(rate=float(rate), count=count, seed=int(rate))
        reqs = trace.requests()
        srv.run(reqs)
        out["rates"][str(rate)] = latency_stats(reqs, trace.jobs)
        if with_cpu:
            threads = host_cores()[0]
            for i, j in enumerate(trace.jobs):
                key = (j.kind, j.log_n, j.input_id, rate)
                if key not in want:
                    x = trace.input(key[:3])
                    if j.kind == 0:
                        want[key] = orc.ntt(x, j.log_n, threads)
                    else:
                        n = 1 << j.log_n
                        total = from_mont(orc.to_int(orc.inner_product(res.base_scalars[:n], x)))
                        want[key] = orc.g1_mul_canonical(orc.generator(), total)
                all_ok = all_ok and reqs[i].status == 0 and np.array_equal(trace.output(i), want[key])
            want.clear()
    if with_cpu:
        out["bit_exact_all_requests"] = bool(all_ok)
    return out


# Per element of each libpoly entry point, the algorithmic minimum: HBM bytes
# (each operand read once, each result written once) and Montgomery products
# (the sequential recurrences: Horner 1, running product 1, Montgomery's batch
# inversion trick 3; 136 multiply-adds each).  The parallel scans do more
# (a tile fold and a rescan), which shows as a lower fraction.
POLY_COSTS = {
    "poly_add": (96, 0),
    "poly_mul": (96, 1),
    "poly_eval": (32, 1),
    "kate_division": (64, 1),
    "scan_mul": (64, 1),
    "batched_invert": (64, 3),
    "permute": (128, 0),
}


def poly_leg(libs, stream, log_n, steps, with_cpu):
    """libpoly.so at 2^log_n resident elements: ms, Gelem/s, HBM and multiply fractions."""
    from zkpoly_amd import hip
    from zkpoly_amd.field import random_elements
    from zkpoly_amd.poly import load_lib
    from zkpoly_amd.runtime import PolyPtr

    lib = load_lib(libs)
    n = 1 << log_n
    a, b, r = hip.DeviceBuffer(n * 32), hip.DeviceBuffer(n * 32), hip.DeviceBuffer(n * 32)
    ha = random_elements(n, 70 + log_n)
    a.upload(ha)
    b.upload(random_elements(n, 80 + log_n))
    x = hip.DeviceBuffer(32)
    x.upload(random_elements(1, 90))
    res = hip.DeviceBuffer(32)
    pa, pb, pr = (PolyPtr(buf.addr, n, 0, 0, n) for buf in (a, b, r))
    size = ctypes.c_ulong(0)

    def temp(fn, *args):
        hip.check(fn(None, ctypes.byref(size), *args), "size query")
        return hip.DeviceBuffer(max(1, size.value))

    t_eval = temp(lib.poly_eval, pa, None, None, None)
    t_kate = temp(lib.kate_division, log_n, pa, None, pr, None)
    t_scan = temp(lib.scan_mul, pr, None)
    t_inv = temp(lib.batched_invert, pr, None)
    # lookup columns: a table of n/2 distinct random elements with repeats, inputs drawn from it
    rng = np.random.default_rng(91)
    pool = random_elements(n // 2, 92)
    tab_h = pool[rng.integers(0, n // 2, n)]
    inp_h = tab_h[rng.integers(0, n, n)]
    ti, tt, ri2 = hip.DeviceBuffer(n * 32), hip.DeviceBuffer(n * 32), hip.DeviceBuffer(n * 32)
    ti.upload(inp_h)
    tt.upload(tab_h)
    del pool, tab_h, inp_h
    pti, ptt, pri2 = (PolyPtr(buf.addr, n, 0, 0, n) for buf in (ti, tt, ri2))
    t_perm = temp(lambda t_, s_: lib.permute(t_, s_, n, pti, ptt, pr, pri2, None))
    s = stream.raw()
    ops = {
        "poly_add": lambda: hip.check(lib.poly_add(pr, pa, pb, s), "poly_add"),
        "poly_mul": lambda: hip.check(lib.poly_mul(pr, pa, pb, s), "poly_mul"),
        "poly_eval": lambda: hip.check(lib.poly_eval(t_eval.addr, None, pa, res.addr, x.addr, s), "poly_eval"),
        "kate_division": lambda: hip.check(lib.kate_division(t_kate.addr, None, log_n, pa, x.addr, pr, s), "kate"),
        "scan_mul": lambda: hip.check(lib.scan_mul(t_scan.addr, None, pr, s), "scan_mul"),
        "batched_invert": lambda: hip.check(lib.batched_invert(t_inv.addr, None, pr, s), "batched_invert"),
        "permute": lambda: hip.check(lib.permute(t_perm.addr, None, n, pti, ptt, pr, pri2, s), "permute"),
    }
    out = {"log_n": log_n, "ops": {}}
    for name, fn in ops.items():
        if name in ("scan_mul", "batched_invert"):
            hip.check(lib.poly_add(pr, pa, PolyPtr(None, 0, 0, 0, 0), s), "copy")  # fresh non-zero input
        ms = time_loop(stream, fn, steps, 2)
        nbytes, products = POLY_COSTS[name]
        gbs = nbytes * n / (ms * 1e-3) / 1e9
        mads = products * MADS_PER_MODMUL * n / (ms * 1e-3)
        out["ops"][name] = {
            "ms": round(ms, 4),
            "gelem_s": round(n / ms / 1e6, 3),
            "hbm_gbs": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "int_mul_frac": round(mads / INT_MUL_PEAK, 4),
        }
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc

        m = min(n, 1 << 22)
        t = time.perf_counter()
        orc.kate_division(ha[:m], random_elements(1, 90)[0])
        cpu_s = time.perf_counter() - t
        out["cpu_baseline_kate_division"] = {"value": round(m / cpu_s / 1e9, 5), "unit": "Gelem/s", "cores": 1,
                                             "kind": "port", "sample": f"kate_division of 2^{m.bit_length() - 1}"}
        mp = 1 << 20
        pool = random_elements(mp // 2, 93)
        rng = np.random.default_rng(94)
        tab = pool[rng.integers(0, mp // 2, mp)]
        inp = tab[rng.integers(0, mp, mp)]
        t = time.perf_counter()
        orc.permute(inp, tab, mp)
        cpu_s = time.perf_counter() - t
        out["cpu_baseline_permute"] = {"value": round(mp / cpu_s / 1e9, 5), "unit": "Gelem/s", "cores": 1,
                                       "kind": "port", "sample": "permute of 2^20 rows (same value distribution)"}
    return out


def pipeline_host(log_n):
    """The pinned host arrays of the pipelined leg (a, b, c, out; gamma, z).
    A runtime pins its host pool once at start-up (the reference's CPU memory
    pool does), so main() allocates these before the other legs run."""
    from zkpoly_amd import hip
    from zkpoly_amd.field import random_elements

    n = 1 << log_n
    host = [hip.PinnedBuffer(n * 32) for _ in range(4)]
    for i in range(3):
        host[i].view()[:] = random_elements(n, 500 + i).reshape(-1)
    sc = [hip.PinnedBuffer(32) for _ in range(2)]
    for i in range(2):
        sc[i].view()[:] = random_elements(1, 510 + i)[0]
    return host, sc


def pipeline_leg(libs, stream, log_n, parts, steps, with_cpu=False, pinned=None):
    """PipelinedFusedKernel: out = a*b + gamma*c - z over 2^log_n host-resident
    (pinned) elements in `parts` chunks, against one whole copy-in, kernel and
    copy-out on a single stream, and the plain pinned copy rates of the link."""
    from zkpoly_amd import fused, hip
    from zkpoly_amd.runtime import GpuBuffer, PolyPtr, Scalar, ScalarArray

    n = 1 << log_n
    nbytes = n * 32
    host, sc = pinned if pinned is not None else pipeline_host(log_n)
    meta = fused.FusedKernelMeta("fused_arith_pipe_abgcz", 5, 1, fused.DEMO_DIR, fused.PipelinedMeta(parts, 2, 0))
    k = fused.PipelinedFusedKernel(libs, meta)
    chunk = nbytes // parts
    keep = []

    def dev(nb):
        b = hip.DeviceBuffer(nb)
        keep.append(b)
        return GpuBuffer(b.addr, nb)

    arg, loc = dev(fused.pipelined_arg_buffer_bytes(meta)), dev(256)
    mut_var = [arg, loc, ScalarArray(n, host[3].addr, device="cpu")] + [dev(chunk) for _ in range(3)]
    var = ([Scalar(sc[0].addr, device="cpu"), Scalar(sc[1].addr, device="cpu")]
           + [ScalarArray(n, host[i].addr, device="cpu") for i in range(3)]
           + [dev(32), dev(32)] + [dev(chunk) for _ in range(6)])
    f = k.get_fn().f
    f(mut_var, var, lambda i: i)  # warm
    t = time.perf_counter()
    for _ in range(steps):
        f(mut_var, var, lambda i: i)
    piped_ms = (time.perf_counter() - t) * 1e3 / steps

    # whole-array reference point: copy in, one kernel, copy out, one stream
    whole = [hip.DeviceBuffer(nbytes) for _ in range(4)]
    dsc = [hip.DeviceBuffer(32) for _ in range(2)]
    plain = fused.FusedKernel(libs, fused.FusedKernelMeta("fused_arith_pipe_abgcz", 5, 1, fused.DEMO_DIR))
    args = (PolyPtr * 6)(PolyPtr(dsc[0].addr, 1, 0, 0, 1), PolyPtr(dsc[1].addr, 1, 0, 0, 1),
                         *[PolyPtr(whole[i].addr, n, 0, 0, n) for i in range(4)])
    dargs = hip.DeviceBuffer(ctypes.sizeof(args))
    h = hip._lib()
    hip.check(h.hipMemcpy(ctypes.c_void_p(dargs.addr), ctypes.addressof(args), ctypes.sizeof(args),
                          hip.hipMemcpyHostToDevice), "args")
    s = stream.raw()

    def serial():
        for i in range(2):
            hip.check(h.hipMemcpyAsync(ctypes.c_void_p(dsc[i].addr), ctypes.c_void_p(sc[i].addr), 32,
                                       hip.hipMemcpyHostToDevice, s), "H2D")
        for i in range(3):
            hip.check(h.hipMemcpyAsync(ctypes.c_void_p(whole[i].addr), ctypes.c_void_p(host[i].addr), nbytes,
                                       hip.hipMemcpyHostToDevice, s), "H2D")
        hip.check(plain.c_func(dargs.addr, dargs.addr + 5 * 40, n, True, None, s), "kernel")
        hip.check(h.hipMemcpyAsync(ctypes.c_void_p(host[3].addr), ctypes.c_void_p(whole[3].addr), nbytes,
                                   hip.hipMemcpyDeviceToHost, s), "D2H")
        stream.sync()

    serial()
    t = time.perf_counter()
    for _ in range(steps):
        serial()
    serial_ms = (time.perf_counter() - t) * 1e3 / steps

    def copy_rate(kind, dst, src):
        hip.check(h.hipMemcpyAsync(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, kind, s), "copy")
        stream.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            hip.check(h.hipMemcpyAsync(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, kind, s), "copy")
        stream.sync()
        return nbytes * steps / (time.perf_counter() - t0) / 1e9

    h2d = copy_rate(hip.hipMemcpyHostToDevice, whole[0].addr, host[0].addr)
    d2h = copy_rate(hip.hipMemcpyDeviceToHost, host[3].addr, whole[3].addr)
    moved = 4 * nbytes  # three polynomials in, one out
    # the link bound: the in-copies (3/4 of the bytes) at the H2D rate overlapping the out-copy at the D2H rate
    bound_ms = max(3 * nbytes / h2d, nbytes / d2h) / 1e6
    cpu = None
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc

        m = 1 << 20  # the same expression on the host, one core, over a 2^20 sample of the inputs
        a, b, c = (host[i].view()[: 4 * m].reshape(m, 4) for i in range(3))
        g = np.broadcast_to(sc[0].view(), (m, 4)).copy()
        z = np.broadcast_to(sc[1].view(), (m, 4)).copy()
        t = time.perf_counter()
        orc.vec("orc_vec_sub", orc.FR, orc.vec("orc_vec_add", orc.FR, orc.vec("orc_vec_mul", orc.FR, a, b),
                                                orc.vec("orc_vec_mul", orc.FR, g, c)), z)
        cpu_s = time.perf_counter() - t
        cpu = {"value": round(m / cpu_s / 1e9, 5), "unit": "Gelem/s", "cores": 1, "kind": "port",
               "sample": "a*b + gamma*c - z over 2^20 host elements"}
    return {
        "cpu_baseline": cpu,
        "log_n": log_n,
        "divide_parts": parts,
        "pipelined_ms": round(piped_ms, 3),
        "serial_ms": round(serial_ms, 3),
        "speedup": round(serial_ms / piped_ms, 3),
        "pipelined_gbs": round(moved / (piped_ms * 1e-3) / 1e9, 2),
        "link_h2d_gbs": round(h2d, 2),
        "link_d2h_gbs": round(d2h, 2),
        "link_bound_ms": round(bound_ms, 3),
        "frac_of_link_bound": round(bound_ms / piped_ms, 4),
    }


def fused_device_leg(libs, stream, log_n, steps):
    """The generated fused kernel a*b + gamma*c - z over 2^log_n device-resident
    elements: three polynomials read and one written per element (128 B), two
    scalars; HBM rate against the 8 TB/s peak."""
    from zkpoly_amd import fused, hip
    from zkpoly_amd.field import random_elements
    from zkpoly_amd.runtime import PolyPtr

    n = 1 << log_n
    bufs = [hip.DeviceBuffer(n * 32) for _ in range(4)]
    for i in range(3):
        bufs[i].upload(random_elements(n, 700 + i))
    dsc = [hip.DeviceBuffer(32) for _ in range(2)]
    for i in range(2):
        dsc[i].upload(random_elements(1, 710 + i))
    plain = fused.FusedKernel(libs, fused.FusedKernelMeta("fused_arith_pipe_abgcz", 5, 1, fused.DEMO_DIR))
    args = (PolyPtr * 6)(PolyPtr(dsc[0].addr, 1, 0, 0, 1), PolyPtr(dsc[1].addr, 1, 0, 0, 1),
                         *[PolyPtr(bufs[i].addr, n, 0, 0, n) for i in range(4)])
    dargs = hip.DeviceBuffer(ctypes.sizeof(args))
    h = hip._lib()
    hip.check(h.hipMemcpy(ctypes.c_void_p(dargs.addr), ctypes.addressof(args), ctypes.sizeof(args),
                          hip.hipMemcpyHostToDevice), "args")

    def run():
        hip.check(plain.c_func(dargs.addr, dargs.addr + 5 * 40, n, True, None, stream.raw()), "fused kernel")

    ms = time_loop(stream, run, steps, 3)
    gbs = 128 * n / (ms * 1e-3) / 1e9
    return {"log_n": log_n, "kernel": "fused_arith_pipe_abgcz", "ms": round(ms, 4), "bytes_alg": 128 * n,
            "hbm_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "int_frac": round(2 * MADS_PER_MODMUL * n / (ms * 1e-3) / INT_MUL_PEAK, 4)}


def traffic_from_profiles(kernel_key):
    """HBM bytes per launch from the committed PMC summary, if one exists."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel_key)
    except (OSError, ValueError):
        return None


# wave64 VALU instructions the chip can issue per second: 256 CUs x 4 SIMDs, each
# taking 4 cycles for a 64-lane instruction on its 16 lanes, at 2.4 GHz
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4


def valu_issue(ms_step, dit_avg_ms):
    """VALU issue rate of the NTT step and of `dit_pass_r4` against VALU_ISSUE_PEAK: wave-level
    VALU instructions per launch from the committed SQ_INSTS_VALU summary of the same transform
    (its last step: head, pass, pass), divided by this run's launch times."""
    path = os.path.join(ROOT, "profiles", "pmc_insts_ntt_2e24.json")
    try:
        with open(path) as f:
            last = json.load(f)["last"][-3:]
    except (OSError, ValueError, KeyError):
        return None
    if [e[1] for e in last] != ["bitrev_head_stream", "dit_pass_r4", "dit_pass_r4"]:
        return None
    step = sum(e[3] for e in last)
    dit = (last[1][3] + last[2][3]) / 2
    return {
        "bound": "valu",
        "unit": "wave64 VALU instructions/s",
        "peak": VALU_ISSUE_PEAK,
        "step_insts": step,
        "step_frac": round(step / (ms_step * 1e-3) / VALU_ISSUE_PEAK, 4),
        "dit_pass_r4_insts_per_launch": dit,
        "dit_pass_r4_frac": round(dit / (dit_avg_ms * 1e-3) / VALU_ISSUE_PEAK, 4),
        "source": "profiles/pmc_insts_ntt_2e24.json (SQ_INSTS_VALU), launch times of this run",
    }


def valu_from_profiles(kernels):
    """VALUBusy / VALUUtilization per kernel from the committed rocprofv3 summary, if present."""
    path = os.path.join(ROOT, "profiles", "pmc_valu.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    return {k: d[k] for k in kernels if k in d}


def main():
    args = parse()
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
            "(multi-GPU sharding is not implemented)",
            file=sys.stderr,
        )
        sys.exit(2)
    from zkpoly_amd import hip
    from zkpoly_amd.runtime import Libs

    if hip.device_count() < 1:
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    libs = Libs()
    stream = hip.Stream(0)
    if args.serving_only:
        log("[bench] serving form only")
        line = {"serving": serving_leg(libs, stream, (50, 100, 200, 500, 1000, 2000), args.serve_count,
                                       not args.no_cpu)}
        print(json.dumps(line), flush=True)
        return
    if args.batched_only:
        log("[bench] batched form only")
        print(json.dumps({"batched": batched_leg(libs, stream, args.steps, not args.no_cpu)}), flush=True)
        return
    if args.pipeline_only:
        log("[bench] pipelined fused kernel only")
        pinned = pipeline_host(args.log_n)
        out = {f"parts_{p}": pipeline_leg(libs, stream, args.log_n, p, args.steps, pinned=pinned)
               for p in (4, 8, 16, 32)}
        print(json.dumps({"pipelined_fused": out}), flush=True)
        return
    if args.poly_only:
        log(f"[bench] libpoly.so 2^{args.log_n} only")
        print(json.dumps({"poly": poly_leg(libs, stream, args.log_n, args.steps, not args.no_cpu)}), flush=True)
        return
    if args.msm_only:
        log(f"[bench] MSM 2^{args.msm_log_n} only")
        print(json.dumps({"msm": msm_leg(libs, stream, args.msm_log_n, args.steps, not args.no_cpu)}), flush=True)
        return
    k = args.log_n
    n = 1 << k
    # the pipelined leg's host pool, pinned before anything else runs
    pinned = None if args.quick else pipeline_host(24)
    log(f"[bench] headline: forward NTT 2^{k}, warmup {args.warmup}, steps {args.steps}")
    w, gpu_out, ms_step, per_kind, nkern = headline(libs, stream, args)
    dit = per_kind[1]
    dit_avg_ms = sum(dit) / len(dit)
    bytes_alg = 64 * n
    achieved = bytes_alg / (dit_avg_ms * 1e-3) / 1e9
    mads_total = MADS_PER_MODMUL * (n // 2) * k
    int_rate = mads_total / (ms_step * 1e-3)  # all k stages over the whole step
    line = {
        "metric": "BN254 NTT 2^24 Gelem/s; MSM 2^20 ms; batched vs plain speedup, bit-exact",
        "value": round(n / ms_step / 1e6, 4),
        "unit": "Gelem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32x8 Montgomery (BN254 Fr)",
        "data": "synthetic uniform Fr elements, PCG64 seed 2024, resident in HBM",
        "config": {
            "workload": f"in-place forward NTT of 2^{k} BN254 Fr elements (natural order), ssip_ntt C ABI",
            "log_n": k,
            "kernels_per_step": nkern,
            "autotuned_min_col_log": w.tuning.ntt.get(k),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "dit_pass_r4",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles("dit_pass_r4"),
            "bytes_alg_per_launch": bytes_alg,
            "avg_launch_ms": round(dit_avg_ms, 5),
            "launches": len(dit),
        },
        "int_mul": {
            "note": "v_mad_u64_u32 rate over the whole step: 136 per Montgomery product, n/2 products per "
                    "stage; they are about half of the product's VALU instructions, so the binding rate is "
                    "the VALU issue in valu_issue",
            "achieved_mads_per_s": float(f"{int_rate:.4e}"),
            "peak_mads_per_s": INT_MUL_PEAK,
            "frac": round(int_rate / INT_MUL_PEAK, 4),
        },
        "valu_issue": valu_issue(ms_step, dit_avg_ms) if k == 24 else None,
        "bitrev_head_avg_ms": round(sum(per_kind[0]) / max(1, len(per_kind[0])), 5),
        "valu_rocprof": valu_from_profiles(["dit_pass_r4", "bitrev_head_stream", "k_accumulate0", "k_close_runs",
                                            "k_reduce_level", "k_reduce_tail", "k_seg_scatter"]),
    }
    if not args.quick:
        # GPU legs first: in this process, after the 16-thread CPU NTT baseline of
        # 2^24 the pipelined leg's copies stopped overlapping (39 vs 30 ms), so the
        # multi-threaded CPU baselines come after it
        host_in = w.host
        del w
        log("[bench] config-2 suite")
        line["suite"] = suite(libs, stream, max(10, args.steps))
        log("[bench] fused a*b + gamma*c - z, device-resident 2^24")
        line["fused_device"] = fused_device_leg(libs, stream, 24, max(10, args.steps))
        log("[bench] libpoly.so at 2^24")
        line["poly"] = poly_leg(libs, stream, 24, max(3, args.steps // 4), not args.no_cpu)
        log("[bench] pipelined fused kernel over host memory (2^24, 8 chunks)")
        line["pipelined_fused"] = pipeline_leg(libs, stream, 24, 8, 3, not args.no_cpu, pinned)
        del pinned
        log("[bench] batched form (64 NTT + fused, one hipGraph launch)")
        line["batched"] = batched_leg(libs, stream, max(5, args.steps // 2), not args.no_cpu)
        log("[bench] serving form (Poisson 50..2000 req/s)")
        line["serving"] = serving_leg(libs, stream, (50, 100, 200, 500, 1000, 2000), args.serve_count,
                                      not args.no_cpu)
        log("[bench] MSM 2^20 window sweep")
        line["msm"] = msm_leg(libs, stream, args.msm_log_n, max(3, args.steps // 4), not args.no_cpu)
    else:
        host_in = w.host
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc

        usable, reported = host_cores()
        log(f"[bench] cpu baseline: oracle best_fft 2^{k} on {usable} threads")
        t = time.perf_counter()
        want = orc.ntt(host_in, k, usable)
        cpu_s = time.perf_counter() - t
        line["cpu_baseline"] = {
            "value": round(n / cpu_s / 1e9, 5),
            "unit": "Gelem/s",
            "cores": usable,
            "kind": "port",
            "sample": f"one forward NTT of the same 2^{k} input (best_fft restatement, 4x64-bit Montgomery), "
                      f"all {usable} cores of the process's affinity set (os.cpu_count() = {reported})",
            "seconds": round(cpu_s, 3),
            "os_cpu_count": reported,
        }
        k1 = min(k, 21)
        log(f"[bench] cpu baseline: oracle best_fft 2^{k1} on 1 thread")
        t = time.perf_counter()
        orc.ntt(host_in[: 1 << k1], k1, 1)
        cpu1_s = time.perf_counter() - t
        line["cpu_baseline_1core"] = {
            "value": round((1 << k1) / cpu1_s / 1e9, 5), "unit": "Gelem/s", "cores": 1, "kind": "port",
            "sample": f"one forward NTT of 2^{k1} elements (the first 2^{k1} of the same input) on one core",
            "seconds": round(cpu1_s, 3),
        }
        line["bit_exact"] = bool(np.array_equal(gpu_out, want))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
