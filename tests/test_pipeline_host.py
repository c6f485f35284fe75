"""Host-side logic of the batched sequence (no GPU): cache-sized keyframe groups."""
import pytest

pytest.importorskip("torch")
from probabilistic_mapping_amd.pipeline import GROUP_CACHE_BYTES, cache_groups  # noqa: E402


def _spans(groups):
    return [(r.start, r.stop) for r in groups]


def test_cache_groups_fp16_workload():
    """752x480x64 FP16 cost volumes (46.2 MB): five per group fit the cache budget."""
    assert _spans(cache_groups([752 * 480 * 64 * 2] * 20)) == [(0, 5), (5, 10), (10, 15),
                                                               (15, 20)]


def test_cache_groups_minimum_and_cover():
    big = GROUP_CACHE_BYTES   # one volume fills the budget: the minimum group size rules
    assert _spans(cache_groups([big] * 9)) == [(0, 4), (4, 8), (8, 9)]
    assert _spans(cache_groups([big] * 9, min_group=1)) == [(k, k + 1) for k in range(9)]
    sizes = [3 << 20, 200 << 20, 50 << 20, 10 << 20, 240 << 20, 1 << 20, 1 << 20]
    g = cache_groups(sizes, min_group=1)
    assert [i for r in g for i in r] == list(range(len(sizes)))           # a partition, in order
    for r in g:
        assert len(r) == 1 or sum(sizes[i] for i in r) <= GROUP_CACHE_BYTES
    assert cache_groups([]) == []
