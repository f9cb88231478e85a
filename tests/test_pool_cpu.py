"""The host memory pool (libzkp_pinned_pool.so) on the CPU: the buddy system's
rounding, splitting, merging, shrink / clear / preallocate and its refusals,
with file-backed slabs (the reference's use_mmap backing), which need no GPU.

Mirrors memory_pool/src/cpu_pool.rs:212-327 (random lengths written and read
back after a shuffle, then freed; preallocate; a zero-length request is
refused) and adds the buddy invariants: blocks never overlap, a block's
address is aligned to its size inside its slab, everything merges back into
whole slabs once freed, and shrink then returns them.
"""
import random
import threading

import pytest

from zkpoly_amd import hip
from zkpoly_amd.pool import MMAP, CpuMemoryPool


@pytest.fixture
def pool_factory(tmp_path):
    made = []

    def make(max_log, base):
        p = CpuMemoryPool(max_log, base, backing=MMAP, tmp_dir=str(tmp_path))
        made.append(p)
        return p

    yield make
    for p in made:
        p.destroy()


def test_random_lengths_written_and_read_back(pool_factory):
    max_log, base = 10, 4
    pool = pool_factory(max_log, base)
    rng = random.Random(7)
    for _ in range(10):
        blocks = []
        for _ in range(100):
            n = rng.randint(1, (base << max_log) // 4)
            ptr = pool.allocate(n * 4)
            v = pool.view(ptr, n * 4, dtype="uint32")
            v[:] = range(n)
            blocks.append((ptr, n))
        rng.shuffle(blocks)
        for ptr, n in blocks:
            v = pool.view(ptr, n * 4, dtype="uint32")
            assert list(v[:: max(1, n // 7)]) == list(range(0, n, max(1, n // 7)))
        for ptr, _ in blocks:
            pool.free(ptr)
    # everything merged back into whole slabs
    assert all(pool.free_blocks(k) == 0 for k in range(max_log))
    assert pool.free_blocks(max_log) * (base << max_log) == pool.used_bytes()
    pool.shrink()
    assert pool.used_bytes() == 0
    pool.clear()


def test_rounding_split_and_merge(pool_factory):
    pool = pool_factory(4, 64)  # slabs of 1 KiB
    a = pool.allocate(1)
    assert pool.block_size(a) == 64
    assert pool.used_bytes() == 1024
    # the slab was split down to class 0: one free block left in every class below 4
    assert [pool.free_blocks(k) for k in range(5)] == [1, 1, 1, 1, 0]
    b = pool.allocate(64)
    assert abs(b - a) == 64  # the buddy
    c = pool.allocate(65)  # rounds up to 128
    assert pool.block_size(c) == 128 and c % 128 == a % 128
    pool.free(a)
    assert pool.free_blocks(0) == 1  # its buddy b is still live: no merge
    pool.free(b)
    pool.free(c)
    assert [pool.free_blocks(k) for k in range(5)] == [0, 0, 0, 0, 1]


def test_blocks_never_overlap_and_are_aligned(pool_factory):
    max_log, base = 8, 32
    pool = pool_factory(max_log, base)
    rng = random.Random(11)
    live = {}
    for step in range(3000):
        if live and (rng.random() < 0.45 or len(live) > 60):
            ptr = rng.choice(list(live))
            pool.free(ptr)
            del live[ptr]
        else:
            k = rng.randint(0, max_log)
            ptr = pool.allocate_log(k)
            live[ptr] = base << k
        spans = sorted(live.items())
        for (p0, s0), (p1, _) in zip(spans, spans[1:]):
            assert p0 + s0 <= p1, step
    slab = base << max_log
    for ptr, size in live.items():
        # slabs are page-aligned mappings and a block's offset inside its slab
        # is a multiple of its size
        assert ptr % min(size, 4096) == 0
    for ptr in list(live):
        pool.free(ptr)
    assert pool.free_blocks(max_log) * slab == pool.used_bytes()


def test_refusals(pool_factory):
    pool = pool_factory(5, 4)
    with pytest.raises(hip.HipError):
        pool.allocate(0)  # the reference panics on a zero-length request
    with pytest.raises(hip.HipError):
        pool.allocate((4 << 5) + 1)  # above the largest class
    with pytest.raises(hip.HipError):
        pool.allocate_log(6)
    p = pool.allocate(8)
    pool.free(p)
    with pytest.raises(hip.HipError):
        pool.free(p)  # double free
    with pytest.raises(hip.HipError):
        pool.free(p + 4)  # not a block start


def test_preallocate_shrink_clear(pool_factory):
    pool = pool_factory(10, 4)
    pool.preallocate(5)
    assert pool.used_bytes() == 5 * 4096
    ptr = pool.allocate(40)
    v = pool.view(ptr, 4, dtype="uint32")
    v[0] = 123
    assert v[0] == 123
    pool.shrink()  # four slabs are wholly free, the fifth holds ptr
    assert pool.used_bytes() == 4096
    pool.free(ptr)
    pool.shrink()
    assert pool.used_bytes() == 0
    pool.allocate(4)
    pool.clear()
    assert pool.used_bytes() == 0


def test_concurrent_threads(pool_factory):
    pool = pool_factory(12, 64)
    errors = []
    live = []
    lock = threading.Lock()

    def worker(seed):
        rng = random.Random(seed)
        mine = []
        try:
            for _ in range(400):
                if mine and rng.random() < 0.5:
                    pool.free(mine.pop(rng.randrange(len(mine))))
                else:
                    mine.append(pool.allocate_log(rng.randint(0, 6)))
            with lock:
                live.extend((p, pool.block_size(p)) for p in mine)
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(s,)) for s in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors
    spans = sorted(live)
    for (p0, s0), (p1, _) in zip(spans, spans[1:]):
        assert p0 + s0 <= p1
    for p, _ in live:
        pool.free(p)
    assert pool.free_blocks(12) * (64 << 12) == pool.used_bytes()
