"""Host logic of bench.py that needs no GPU: the split of the K timed steps
into runs (every run of odd length, so that with two steps per launch each
run ends with exactly one single step; exactly K steps in all)."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_timed_segments_are_odd_and_sum_to_k():
    b = load_bench()
    for k in range(0, 3001):
        segs = b.timed_segments(k)
        assert sum(segs) == k, k
        assert all(v > 0 and v % 2 == 1 for v in segs), (k, segs)
        assert len(segs) <= 5, (k, segs)
    assert b.timed_segments(2000) == [501, 499, 501, 499]
    assert b.timed_segments(20) == [9, 11]
    assert b.timed_segments(21) == [21]
    # near-equal runs: the K = 2000 default's median segment is a full run
    for k in (200, 2000, 2001, 20000):
        segs = b.timed_segments(k)
        assert max(segs) - min(segs) <= 4, (k, segs)
