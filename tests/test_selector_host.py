"""Host logic of the selector trainer (no GPU): the schedule-aware uses restatement against a
direct walk of every row, and the row grouping."""
import numpy as np
import pytest

from simdgraphprocessing_amd import selector


def _random_oriented(n, seed, cap=40):
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        k = int(rng.integers(0, min(i, cap) + 1))
        rows.append(np.sort(rng.choice(i, k, replace=False)) if k else np.zeros(0, np.int64))
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(r) for r in rows])
    return rows, off, np.concatenate(rows).astype(np.uint32)


@pytest.mark.parametrize("split,bsh", [(50, 10), (100, 20), (10, 0), (299, 5), (400, 1)])
def test_tuned_uses_matches_direct_walk(split, bsh):
    rows, off, nbr = _random_oriented(300, split + bsh)
    card = np.diff(off)
    want = np.zeros(len(rows), np.int64)
    for r in rows:
        hi = [j for j in r if j >= split]
        if any(card[j] > bsh for j in hi):
            for j in hi:
                want[j] += card[j] > 0
    assert np.array_equal(selector.tuned_uses(off, nbr, split, bsh), want)


def test_groups_partition_the_candidates():
    rng = np.random.default_rng(3)
    card = rng.integers(0, 600, 5000)
    uses = rng.integers(0, 2000, 5000)
    words = card * rng.integers(1, 100, 5000)
    groups = selector.groups_of(card, uses, words, 1)
    allrows = np.concatenate(groups)
    assert len(allrows) == len(np.unique(allrows))
    assert np.array_equal(np.sort(allrows), np.nonzero((card >= 2) & (uses >= 1))[0])


def test_significant_keeps_only_differences_above_the_noise():
    assert selector.significant([0.5, 0.52, 0.49, 0.51], 10.0) == pytest.approx(0.505)
    assert selector.significant([0.01, -0.02, 0.015, -0.01], 10.0) == 0.0  # within 0.1% of the pass
    assert selector.significant([0.3, -0.3, 0.31, -0.29], 1.0) == 0.0      # spread above the median
    assert selector.significant([-0.2, -0.21, -0.19, -0.2], 5.0) == pytest.approx(-0.2)
