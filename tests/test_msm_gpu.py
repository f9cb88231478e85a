"""GPU parity of the MSM libraries against the oracle (affine, bit-exact).

Mirrors core/tests/msm.rs:122-250 (bases and scalars from a seed, several
batches, result compared with best_multiexp) and adds the workload: 2^20
points P_i = k_i G, uniform scalars and the 10 % zero / 10 % one variant,
window sweep 12..18, checked through the closed form (sum k_i s_i) G.  Edge
cases: n = 1, all-zero scalars (identity (0, 0)), repeated and opposite bases
(the equal-point / inverse-point branches), precomputed tables.
"""
import ctypes

import numpy as np
import pytest

import oracle as orc
from zkpoly_amd import hip
from zkpoly_amd.field import FR_MODULUS, array_to_ints, from_mont, mont_limbs, random_elements
from zkpoly_amd.msm import (
    MSM,
    MSMPrecompute,
    MsmConfig,
    msm_device,
    mul_generator,
    workspace_size,
)
from zkpoly_amd.runtime import GpuBuffer, Point, PointArray, ScalarArray

pytestmark = pytest.mark.gpu


def gpu_bases(libs, stream, ks):
    n = ks.shape[0]
    kb = hip.DeviceBuffer(n * 32)
    kb.upload(ks)
    pb = hip.DeviceBuffer(n * 64)
    mul_generator(libs, kb.addr, n, pb.addr, stream)
    stream.sync()
    return pb


def closed_form(ks, ss):
    """(sum k_i s_i mod r) G through the oracle (inner product in Montgomery form)."""
    ip = orc.inner_product(ks, ss)  # Montgomery(k.s) ... both inputs Montgomery -> product Montgomery
    total = from_mont(orc.to_int(ip))
    return orc.g1_mul_canonical(orc.generator(), total)


def run_device(libs, stream, pb, ss, window, target=None):
    n = ss.shape[0]
    target = target or -(-254 // window)
    sb = hip.DeviceBuffer(max(1, n) * 32)
    if n:
        sb.upload(ss)
    ws = hip.DeviceBuffer(workspace_size(libs, n, window, target))
    return msm_device(libs, pb.addr if pb else 0, sb.addr, n, window, target, ws, stream)


def test_generator_multiples_match_oracle(libs, stream):
    ks = random_elements(64, 1)
    pb = gpu_bases(libs, stream, ks)
    got = pb.download(np.uint64, 64 * 8).reshape(64, 8)
    for i in range(64):
        want = orc.g1_mul_canonical(orc.generator(), from_mont(array_to_ints(ks[i : i + 1])[0]))
        assert list(got[i]) == list(want)
        assert orc.g1_on_curve(got[i])


@pytest.mark.parametrize("n", [1, 2, 3, 31, 100, 1000])
def test_small_against_best_multiexp(libs, stream, n):
    ks = random_elements(n, 10 + n)
    ss = random_elements(n, 20 + n)
    pb = gpu_bases(libs, stream, ks)
    bases = pb.download(np.uint64, n * 8).reshape(n, 8)
    got = run_device(libs, stream, pb, ss, 12)
    assert list(got) == list(orc.best_multiexp(ss, bases))


def test_zero_scalars_give_identity(libs, stream):
    n = 500
    pb = gpu_bases(libs, stream, random_elements(n, 5))
    got = run_device(libs, stream, pb, np.zeros((n, 4), dtype=np.uint64), 13)
    assert list(got) == [0] * 8


def test_repeated_and_opposite_bases(libs, stream):
    """Equal points hit the doubling branch of the mixed addition; P + (-P) the identity."""
    n = 4096
    ks = np.zeros((n, 4), dtype=np.uint64)
    k = mont_limbs(123456789)
    km = mont_limbs(FR_MODULUS - 123456789)
    ks[0::2] = k
    ks[1::2] = km
    ks[: n // 4] = k  # a run of identical bases
    pb = gpu_bases(libs, stream, ks)
    ss = random_elements(n, 77)
    ss[n // 2 :: 3] = mont_limbs(1)  # many equal (point, digit) pairs
    bases = pb.download(np.uint64, n * 8).reshape(n, 8)
    want = orc.best_multiexp(ss, bases)
    assert list(want) == list(closed_form(ks, ss))
    for w in (12, 16):
        assert list(run_device(libs, stream, pb, ss, w)) == list(want)


@pytest.mark.parametrize("window", [12, 13, 14, 15, 16, 17, 18])
def test_reference_abi_window_sweep(libs, stream, window):
    """The reference's RegisteredFunction path: host points / scalars, several batches."""
    n, batches = 1 << 14, 3
    cfg = MsmConfig(window_size=window, target_window=-(-254 // window))
    ks = random_elements(n, 1000 + window)
    pb = gpu_bases(libs, stream, ks)
    host_pts = np.ascontiguousarray(pb.download(np.uint64, n * 8).reshape(n, 8))
    scal = [random_elements(n, 2000 + 10 * window + b) for b in range(batches)]
    op = MSM(libs, cfg)
    size = op.get_buffer_size(n)[0]
    buf = hip.DeviceBuffer(size)
    results = [Point() for _ in range(batches)]
    var = [PointArray(host_pts.ctypes.data, n)] + [ScalarArray(n, s.ctypes.data, "cpu") for s in scal]
    op.get_fn().f([GpuBuffer(buf.addr, size)] + results, var, lambda c: c)
    for b in range(batches):
        assert list(results[b].data) == list(closed_form(ks, scal[b])), f"batch {b}"


def test_reference_abi_parts_and_precompute(libs, stream):
    """parts > 1 streams points through the buffer; target_window < windows uses precomputed tables."""
    n = 10000
    ks = random_elements(n, 31337)
    pb = gpu_bases(libs, stream, ks)
    t0 = np.ascontiguousarray(pb.download(np.uint64, n * 8).reshape(n, 8))
    ss = random_elements(n, 4242)
    want = closed_form(ks, ss)
    for window, target, parts in ((16, 16, 3), (16, 8, 1), (13, 5, 2)):
        cfg = MsmConfig(window_size=window, target_window=target, parts=parts)
        n_pre = cfg.get_precompute()
        tables = [t0] + [np.zeros_like(t0) for _ in range(n_pre - 1)]
        arrays = [PointArray(t.ctypes.data, n) for t in tables]
        MSMPrecompute(libs, cfg).get_fn()(arrays, n, 1)
        if n_pre > 1:
            # table j = 2^(j * n_windows * s) P
            shift = min(-(-254 // window), target) * window
            assert list(tables[1][7]) == list(orc.g1_mul_canonical(t0[7], 1 << shift))
        op = MSM(libs, cfg)
        size = op.get_buffer_size(n)[0]
        buf = hip.DeviceBuffer(size)
        res = Point()
        op.get_fn().f([GpuBuffer(buf.addr, size), res], arrays + [ScalarArray(n, ss.ctypes.data, "cpu")])
        assert list(res.data) == list(want), (window, target, parts)


@pytest.mark.parametrize("batches,bpr,parts,ss,sp", [(5, 2, 3, 2, 2), (4, 3, 1, 2, 1), (3, 1, 4, 1, 1), (6, 4, 2, 3, 3)])
def test_reference_abi_runs_and_stages(libs, stream, batches, bpr, parts, ss, sp):
    """batch_per_run batches share one pass over the parts (each run reverses
    the previous run's order, so its first part is still resident), with
    stage_scalars / stage_points rotating buffers on copy streams; host arrays
    in pinned memory from the CPU pool.  Every batch against the closed form."""
    from zkpoly_amd.pool import CpuMemoryPool

    n = 10007
    ks = random_elements(n, 8500 + batches)
    pb = gpu_bases(libs, stream, ks)
    pool = CpuMemoryPool(20, 64, libs)  # 64 MiB pinned slabs
    pts_ptr = pool.allocate(n * 64)
    pool.view(pts_ptr, n * 64)[:] = pb.download(np.uint64, n * 8)
    scal, arrays = [], []
    for b in range(batches):
        s_host = random_elements(n, 8600 + 10 * batches + b)
        if b == 1:
            s_host[::9] = 0
        ptr = pool.allocate(n * 32)
        pool.view(ptr, n * 32)[:] = s_host.reshape(-1)
        scal.append(s_host)
        arrays.append(ScalarArray(n, ptr, "cpu"))
    cfg = MsmConfig(window_size=13, target_window=20, batch_per_run=bpr, parts=parts, stage_scalars=ss,
                    stage_points=sp)
    op = MSM(libs, cfg)
    size = op.get_buffer_size(n)[0]
    buf = hip.DeviceBuffer(size)
    results = [Point() for _ in range(batches)]
    print(f"msm runs batches={batches} bpr={bpr} parts={parts} stages={ss}/{sp}", flush=True)
    op.get_fn().f([GpuBuffer(buf.addr, size)] + results, [PointArray(pts_ptr, n)] + arrays, lambda c: c)
    for b in range(batches):
        assert list(results[b].data) == list(closed_form(ks, scal[b])), f"batch {b}"
    pool.destroy()


@pytest.mark.parametrize("pt_shift,sc_shift,parts", [(16, 32, 1), (8, 4, 1), (16, 32, 2)])
def test_reference_abi_pinned_offsets(libs, stream, pt_shift, sc_shift, parts):
    """Pinned host arrays that start pt_shift / sc_shift bytes into a pool
    block, two batches in one run: in one part, 16-byte aligned ones are
    pulled into HBM by the library's copy kernel and others go through
    hipMemcpyAsync; in two parts every copy is a DMA copy.  Against the closed
    form."""
    from zkpoly_amd.pool import CpuMemoryPool

    n = 6001
    ks = random_elements(n, 8700 + pt_shift)
    pb = gpu_bases(libs, stream, ks)
    pool = CpuMemoryPool(20, 64, libs)
    pts_ptr = pool.allocate(n * 64 + 64) + pt_shift
    pv = (ctypes.c_uint8 * (n * 64)).from_address(pts_ptr)
    np.frombuffer(pv, dtype=np.uint8)[:] = pb.download(np.uint64, n * 8).view(np.uint8)
    scal, arrays = [], []
    for b in range(2):
        ss = random_elements(n, 8800 + sc_shift + b)
        sc_ptr = pool.allocate(n * 32 + 64) + sc_shift
        sv = (ctypes.c_uint8 * (n * 32)).from_address(sc_ptr)
        np.frombuffer(sv, dtype=np.uint8)[:] = ss.reshape(-1).view(np.uint8)
        scal.append(ss)
        arrays.append(ScalarArray(n, sc_ptr, "cpu"))
    cfg = MsmConfig(window_size=12, target_window=22, parts=parts, stage_scalars=2, stage_points=2, batch_per_run=2)
    op = MSM(libs, cfg)
    size = op.get_buffer_size(n)[0]
    buf = hip.DeviceBuffer(size)
    res = [Point(), Point()]
    print(f"msm pinned offsets points +{pt_shift} scalars +{sc_shift} parts {parts}", flush=True)
    op.get_fn().f([GpuBuffer(buf.addr, size)] + res, [PointArray(pts_ptr, n)] + arrays, lambda c: c)
    for b in range(2):
        assert list(res[b].data) == list(closed_form(ks, scal[b])), f"batch {b}"
    pool.destroy()


def test_parts_with_a_heavy_bucket(libs, stream):
    """Window 16 in two parts with 30 % of the scalars equal to 1: the first
    part stores its buckets and the second adds to them, while the scalar-1
    bucket is a run far longer than the direct run pass closes, so the
    chunked levels finish it in both parts."""
    n = 1 << 14
    ks = random_elements(n, 8100)
    ss = random_elements(n, 8101)
    rng = np.random.Generator(np.random.PCG64(8102))
    ss[rng.permutation(n)[: (3 * n) // 10]] = mont_limbs(1)
    pb = gpu_bases(libs, stream, ks)
    t0 = np.ascontiguousarray(pb.download(np.uint64, n * 8).reshape(n, 8))
    cfg = MsmConfig(window_size=16, target_window=16, parts=2)
    op = MSM(libs, cfg)
    size = op.get_buffer_size(n)[0]
    buf = hip.DeviceBuffer(size)
    res = Point()
    op.get_fn().f([GpuBuffer(buf.addr, size), res], [PointArray(t0.ctypes.data, n), ScalarArray(n, ss.ctypes.data, "cpu")])
    assert list(res.data) == list(closed_form(ks, ss))


@pytest.mark.parametrize("variant", ["uniform", "zeros_ones"])
def test_workload_2_20(libs, stream, variant):
    """Config 3: 2^20 points, window 16, closed-form check; the variant has 10 % zeros, 10 % ones."""
    n = 1 << 20
    ks = random_elements(n, 3000)
    ss = random_elements(n, 3001)
    if variant == "zeros_ones":
        rng = np.random.Generator(np.random.PCG64(3002))
        perm = rng.permutation(n)
        ss[perm[: n // 10]] = 0
        ss[perm[n // 10 : n // 5]] = mont_limbs(1)
    pb = gpu_bases(libs, stream, ks)
    want = closed_form(ks, ss)
    for window in (16, 12, 18):
        assert list(run_device(libs, stream, pb, ss, window)) == list(want), window


@pytest.mark.parametrize("window", [13, 14, 15, 17])
def test_workload_2_20_other_windows(libs, stream, window):
    """The rest of config 3's sweep (12, 16 and 18 run above) at 2^20 points."""
    n = 1 << 20
    ks = random_elements(n, 3000)
    ss = random_elements(n, 3001)
    pb = gpu_bases(libs, stream, ks)
    assert list(run_device(libs, stream, pb, ss, window)) == list(closed_form(ks, ss)), window


@pytest.mark.parametrize("n", [1, 2, 3, 17])
def test_reference_abi_tiny(libs, stream, n):
    """The drop-in msm on a handful of points, in one part and in up to four
    (more parts than points leaves some empty), two batches in one run, the
    second all zeros: against best_multiexp, the zero batch the identity."""
    ks = random_elements(n, 9600 + n)
    pb = gpu_bases(libs, stream, ks)
    host_pts = np.ascontiguousarray(pb.download(np.uint64, n * 8).reshape(n, 8))
    scal = [random_elements(n, 9700 + n), np.zeros((n, 4), dtype=np.uint64)]
    for parts in (1, min(4, n + 1)):
        cfg = MsmConfig(window_size=12, target_window=22, parts=parts, batch_per_run=2)
        op = MSM(libs, cfg)
        size = op.get_buffer_size(n)[0]
        buf = hip.DeviceBuffer(size)
        res = [Point(), Point()]
        var = [PointArray(host_pts.ctypes.data, n)] + [ScalarArray(n, s.ctypes.data, "cpu") for s in scal]
        print(f"msm tiny n={n} parts={parts}", flush=True)
        op.get_fn().f([GpuBuffer(buf.addr, size)] + res, var, lambda c: c)
        assert list(res[0].data) == list(orc.best_multiexp(scal[0], host_pts)), parts
        assert list(res[1].data) == [0] * 8, parts


def test_reference_abi_concurrent_calls(libs, stream):
    """Three threads call the drop-in msm at once (ctypes drops the GIL for the
    call), each with its own configuration, device buffer and host arrays:
    one part from pinned pool memory (the copy kernel), one part from pageable
    memory, three parts.  Each call takes its own set of streams, events and
    pinned sums from the library's free list; every result against the
    closed form."""
    import threading

    from zkpoly_amd.pool import CpuMemoryPool

    n = (1 << 16) + 5
    ks = random_elements(n, 9400)
    pb = gpu_bases(libs, stream, ks)
    host_pts = np.ascontiguousarray(pb.download(np.uint64, n * 8).reshape(n, 8))
    pool = CpuMemoryPool(22, 64, libs)
    pin_pts = pool.allocate(n * 64)
    pool.view(pin_pts, n * 64)[:] = host_pts.reshape(-1)
    jobs = []
    for j, (window, parts, pinned) in enumerate(((16, 1, True), (13, 1, False), (14, 3, False))):
        ss = random_elements(n, 9500 + j)
        if pinned:
            sp = pool.allocate(n * 32)
            pool.view(sp, n * 32)[:] = ss.reshape(-1)
            arrays = [PointArray(pin_pts, n), ScalarArray(n, sp, "cpu")]
        else:
            arrays = [PointArray(host_pts.ctypes.data, n), ScalarArray(n, ss.ctypes.data, "cpu")]
        op = MSM(libs, MsmConfig(window_size=window, target_window=-(-254 // window), parts=parts))
        size = op.get_buffer_size(n)[0]
        jobs.append((op, hip.DeviceBuffer(size), size, Point(), arrays, ss))
    errors = []

    def run(job):
        op, buf, size, res, arrays, _ = job
        try:
            for _ in range(3):
                op.get_fn().f([GpuBuffer(buf.addr, size), res], arrays, lambda c: c)
        except Exception as e:  # reported below, on the main thread
            errors.append(e)

    threads = [threading.Thread(target=run, args=(job,)) for job in jobs]
    print("msm concurrent calls: 3 threads", flush=True)
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for j, (_, _, _, res, _, ss) in enumerate(jobs):
        assert list(res.data) == list(closed_form(ks, ss)), j
    pool.destroy()


@pytest.mark.parametrize("window,target,parts", [(16, 16, 1), (16, 16, 2), (13, 10, 3)])
def test_reference_abi_2_20(libs, stream, window, target, parts):
    """The drop-in msm symbol at 2^20: host points and two batches of host
    scalars (the second with 10 % zeros and 10 % ones), the points streamed in
    `parts` pieces; with target < windows the precomputed table of
    msm_precompute is used too."""
    n, batches = 1 << 20, 2
    ks = random_elements(n, 9100)
    pb = gpu_bases(libs, stream, ks)
    t0 = np.ascontiguousarray(pb.download(np.uint64, n * 8).reshape(n, 8))
    del pb
    scal = [random_elements(n, 9200 + b) for b in range(batches)]
    rng = np.random.Generator(np.random.PCG64(9300))
    perm = rng.permutation(n)
    scal[1][perm[: n // 10]] = 0
    scal[1][perm[n // 10 : n // 5]] = mont_limbs(1)
    cfg = MsmConfig(window_size=window, target_window=target, parts=parts)
    n_pre = cfg.get_precompute()
    tables = [t0] + [np.zeros_like(t0) for _ in range(n_pre - 1)]
    arrays = [PointArray(t.ctypes.data, n) for t in tables]
    if n_pre > 1:
        MSMPrecompute(libs, cfg).get_fn()(arrays, n, 1)
    op = MSM(libs, cfg)
    size = op.get_buffer_size(n)[0]
    buf = hip.DeviceBuffer(size)
    results = [Point() for _ in range(batches)]
    var = arrays + [ScalarArray(n, s.ctypes.data, "cpu") for s in scal]
    print(f"msm 2^20 window={window} target={target} parts={parts}", flush=True)
    op.get_fn().f([GpuBuffer(buf.addr, size)] + results, var, lambda c: c)
    for b in range(batches):
        assert list(results[b].data) == list(closed_form(ks, scal[b])), f"batch {b}"


def test_workload_2_20_matches_best_multiexp(libs, stream):
    """One 2^20 instance also against the restated best_multiexp itself."""
    n = 1 << 20
    ks = random_elements(n, 5000)
    ss = random_elements(n, 5001)
    pb = gpu_bases(libs, stream, ks)
    bases = pb.download(np.uint64, n * 8).reshape(n, 8)
    assert list(run_device(libs, stream, pb, ss, 16)) == list(orc.best_multiexp(ss, bases))


@pytest.mark.parametrize("window", [2, 5, 8, 20, 21])
def test_narrow_and_wide_windows(libs, stream, window):
    """Windows at both ends of the partition sort (2: tiles of 64 points and 128
    windows; 20: 4096 low bins) and past it (21: the atomic counting sort)."""
    n = 3000
    ks = random_elements(n, 6000 + window)
    ss = random_elements(n, 6100 + window)
    pb = gpu_bases(libs, stream, ks)
    assert list(run_device(libs, stream, pb, ss, window)) == list(closed_form(ks, ss))


@pytest.mark.parametrize("window", [2, 3, 13, 16, 21])
def test_carries_through_every_window(libs, stream, window):
    """Every fifth scalar r - 1, whose signed digits carry through every
    window into the top one, among uniform scalars, at 2^16 - 1 and 2^16
    points: odd window counts (13: 20 windows, 21: 13) and the extra top
    window of 2 and 3 bits, against the closed form."""
    for n in ((1 << 16) - 1, 1 << 16):
        ks = random_elements(n, 6200 + window)
        ss = random_elements(n, 6300 + window)
        ss[::5] = mont_limbs(FR_MODULUS - 1)
        pb = gpu_bases(libs, stream, ks)
        print(f"msm carries n={n} window={window}", flush=True)
        assert list(run_device(libs, stream, pb, ss, window)) == list(closed_form(ks, ss)), n


@pytest.mark.parametrize("value", [1, FR_MODULUS - 1, 12345])
def test_all_scalars_equal(libs, stream, value):
    """Every scalar the same: each window's entries land in one bucket, one sort
    segment holds them all, and the result is value * sum P_i."""
    n = 1 << 15
    ks = random_elements(n, 7000)
    ss = np.broadcast_to(mont_limbs(value), (n, 4)).copy()
    pb = gpu_bases(libs, stream, ks)
    for window in (16, 13):
        assert list(run_device(libs, stream, pb, ss, window)) == list(closed_form(ks, ss)), window
