"""Host logic of the batched form: how the MSMs of a batch are spread over the
graph's parallel branches (zkpoly_amd.batch.msm_branches)."""
import random

from zkpoly_amd.batch import msm_branches


def check(lengths, branches):
    groups = msm_branches(lengths, branches)
    flat = sorted(j for g in groups for j in g)
    assert flat == list(range(len(lengths)))  # every MSM exactly once
    assert len(groups) <= branches and all(groups)  # no empty branch
    assert all(g == sorted(g) for g in groups)  # index order within a branch
    loads = [sum(lengths[j] for j in g) for g in groups]
    if len(groups) == branches:
        # greedy largest-first: no branch exceeds another by more than its largest MSM
        assert max(loads) - min(loads) <= max(lengths)
    return groups


def test_spread_of_uneven_sizes():
    rng = random.Random(5)
    for _ in range(200):
        n = rng.randrange(1, 20)
        lengths = [1 << rng.randrange(10, 21) for _ in range(n)]
        for b in (1, 2, 4, 7):
            check(lengths, b)


def test_fewer_msms_than_branches():
    assert check([1 << 12, 1 << 16], 4) == [[1], [0]]


def test_largest_alone():
    # one MSM as large as all the others together gets a branch of its own
    groups = check([1 << 20] + [1 << 16] * 16, 2)
    assert [0] in groups
