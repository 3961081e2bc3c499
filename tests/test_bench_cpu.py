"""Host logic of bench.py that needs no GPU: the split of the K timed steps
into runs (every run of even length but the last when K is odd, so that with
two steps per launch the runs are whole two-step launches and at most one
single step is timed; exactly K steps in all)."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_timed_segments_are_even_and_sum_to_k():
    b = load_bench()
    for k in range(0, 3001):
        segs = b.timed_segments(k)
        assert sum(segs) == k, k
        assert all(v > 0 for v in segs), (k, segs)
        assert sum(v % 2 for v in segs) == k % 2, (k, segs)
        assert len(segs) <= 4, (k, segs)
    assert b.timed_segments(2000) == [500, 500, 500, 500]
    assert b.timed_segments(20) == [20]
    assert b.timed_segments(21) == [21]
    # near-equal runs: the K = 2000 default's median segment is a full run
    for k in (200, 2000, 2001, 20000):
        segs = b.timed_segments(k)
        assert max(segs) - min(segs) <= 4, (k, segs)
