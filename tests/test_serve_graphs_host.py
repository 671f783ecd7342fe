"""The serving loop's per-shape graph cache (serve.GraphCache), on the CPU.

The token mock of tests/test_serve_host.py stands in for the GPU, extended
with graph capture: calls issued between begin_capture and end_capture are
recorded instead of executed, and a launch re-executes them on the cache
backend's own memory.  A request served by a replay must end with the token
plain program-order evaluation gives, exactly like one served by the
batched steps, whatever mix of the two a run produces.
"""
import numpy as np
import pytest

from fideslib_amd import serve as S
from tests.test_serve_host import COEF, PRM, MockBackend, _MockServer, make_requests, sequential

OPS = ("hmul", "rot", "lin", "lincomb", "rescale", "drop")


class MockGraph:
    def __init__(self, calls):
        self.calls = calls
        self.closed = False

    def close(self):
        self.closed = True


class CaptureMock(MockBackend):
    """MockBackend with capture/replay and always-complete events."""

    def __init__(self, words):
        super().__init__(words)
        self.rec = None
        self.launches = 0
        self.scratch = []

    def __getattribute__(self, name):
        if name in OPS:
            rec = object.__getattribute__(self, "rec")
            if rec is not None:
                return lambda *args: rec.append((name, args))
        return object.__getattribute__(self, name)

    def begin_capture(self, scratch_bytes, task_bytes, scratch_ptr=None):
        assert self.rec is None
        self.scratch.append(scratch_ptr)
        self.rec = []

    def end_capture(self):
        g = MockGraph(self.rec)
        self.rec = None
        return g

    def launch(self, graph):
        assert not graph.closed
        self.launches += 1
        for name, args in graph.calls:
            getattr(MockBackend, name)(self, *args)

    def sync(self):
        pass

    def record(self):
        return object()

    def query(self, ev):
        return True

    def release(self, ev):
        pass


class TokenCache(S.GraphCache):
    def _read_output(self, be, a, out):
        return be._read(be.ptr(a.reg_ptr_words(out)), out.level)


def cache_for(programs, copies=1):
    words = S.GraphCache.words_needed(programs, copies)
    be = CaptureMock(words)
    c = TokenCache(be, 1, scratch_bytes=1 << 20, scratch_ptr=12345)
    for P in programs:
        c.add(P, copies)
    return c


def test_program_signature_identifies_shapes():
    a = S.chebyshev_program(PRM, 12, COEF, -8, 8)
    b = S.chebyshev_program(PRM, 12, COEF, -8, 8)
    c = S.chebyshev_program(PRM, 13, COEF, -8, 8)
    assert S.program_signature(a) == S.program_signature(b)
    assert S.program_signature(a) != S.program_signature(c)
    r1, r2 = S.rotsum_program(PRM, 20, 3), S.rotsum_program(PRM, 20, 4)
    assert S.program_signature(r1) != S.program_signature(r2)
    assert S.program_signature(r1) == S.program_signature(S.rotsum_program(PRM, 20, 3))


@pytest.mark.parametrize("realtime,seed", [(False, 1), (True, 2), (False, 3)])
def test_graph_and_batched_requests_equal_sequential(realtime, seed):
    rng = np.random.default_rng(seed)
    reqs = make_requests(rng, 40, t_scale=0.002 if realtime else 0.0)
    # capture the shapes of half of the distinct programs
    shapes = {}
    for r in reqs:
        shapes.setdefault(S.program_signature(r.program), r.program)
    cached = list(shapes.values())[::2]
    cache = cache_for(cached)
    be = MockBackend(400000)
    srv = _MockServer(be, 1)
    srv.run(reqs, realtime=realtime, graphs=cache)
    assert all(r.status == "done" for r in reqs)
    for r in reqs:
        assert r.result == sequential(r.program, [t for t, _ in r.inputs]), "request %d" % r.rid
    n_graph = sum(r.graph for r in reqs)
    assert n_graph == cache.stats["replays"] == cache.be.launches > 0
    assert n_graph < len(reqs)
    # every graph read the shared scratch it was given; nothing is left running
    assert set(cache.be.scratch) == {12345}
    assert not cache.inflight and not any(e.busy for l in cache.entries.values() for e in l)


def test_busy_entries_send_requests_to_the_batched_loop():
    # ten requests of one shape at once: with completion events that are never
    # ready, one entry (copies=1) serves the first, the rest join the steps
    P = S.chebyshev_program(PRM, 15, COEF, -8, 8)
    cache = cache_for([P])
    cache.be.query = lambda ev: False
    cache.be.wait = lambda ev: None
    reqs = [S.Request(i, S.chebyshev_program(PRM, 15, COEF, -8, 8), 0.0,
                      inputs=[(hash(("in", i)), 15)], keep_result=True) for i in range(10)]
    srv = _MockServer(MockBackend(200000), 1)
    srv.run(reqs, realtime=False, graphs=cache)
    assert [r.graph for r in reqs] == [True] + [False] * 9
    for r in reqs:
        assert r.result == sequential(r.program, [hash(("in", r.rid))])


def test_copies_serve_concurrent_requests():
    P = S.rotsum_program(PRM, 10, 5)
    cache = cache_for([P], copies=3)
    cache.be.query = lambda ev: False
    cache.be.wait = lambda ev: None
    reqs = [S.Request(i, S.rotsum_program(PRM, 10, 5), 0.0, inputs=[(hash(("x", i)), 10)], keep_result=True)
            for i in range(5)]
    srv = _MockServer(MockBackend(100000), 1)
    srv.run(reqs, realtime=False, graphs=cache)
    assert sum(r.graph for r in reqs) == 3
    for r in reqs:
        assert r.result == sequential(r.program, [hash(("x", r.rid))])


def test_words_needed_counts_each_shape_once():
    a = S.chebyshev_program(PRM, 12, COEF, -8, 8)
    b = S.chebyshev_program(PRM, 12, COEF, -8, 8)
    c = S.rotsum_program(PRM, 7, 2)
    one = S.GraphCache.words_needed([a])
    assert S.GraphCache.words_needed([a, b]) == one
    assert S.GraphCache.words_needed([a, b, c], copies=2) == 2 * (one + S.GraphCache.words_needed([c]))


def test_add_tops_up_to_copies():
    P = S.rotsum_program(PRM, 6, 2)
    cache = cache_for([P, S.rotsum_program(PRM, 6, 2)], copies=2)
    assert cache.stats["graphs"] == 2 and cache.stats["shapes"] == 1


def test_cached_shape_larger_than_the_arena_is_served_by_its_graph():
    # the loop's arena holds none of these programs; the cached shape still
    # runs (its graph has storage of its own), the uncached one is rejected
    P = S.chebyshev_program(PRM, 20, COEF, -8, 8)
    cache = cache_for([P])
    small = MockBackend(2 * 20)
    reqs = [S.Request(0, S.chebyshev_program(PRM, 20, COEF, -8, 8), 0.0, inputs=[(hash("a"), 20)], keep_result=True),
            S.Request(1, S.rotsum_program(PRM, 20, 3), 0.0, inputs=[(hash("b"), 20)], keep_result=True),
            S.Request(2, S.chebyshev_program(PRM, 20, COEF, -8, 8), 0.0, inputs=[(hash("c"), 20)], keep_result=True)]
    srv = _MockServer(small, 1)
    srv.run(reqs, realtime=False, graphs=cache)
    assert reqs[0].status == "done" and reqs[0].graph
    assert reqs[0].result == sequential(reqs[0].program, [hash("a")])
    assert reqs[1].status == "rejected"
    assert reqs[2].status == "done" and reqs[2].graph   # waited for the entry, then replayed
    assert reqs[2].result == sequential(reqs[2].program, [hash("c")])


def test_lanes_run_their_graphs_independently():
    # a long graph that never finishes on lane 1 does not hold up lane 0:
    # requests of lane 0's shape keep replaying, those of lane 1's shape
    # after the first join the batched steps
    long_p = S.chebyshev_program(PRM, 25, COEF, -8, 8)
    short_p = S.rotsum_program(PRM, 9, 2)
    cache = cache_for([short_p])
    be1 = CaptureMock(S.GraphCache.words_needed([long_p]))
    lane = cache.add_lane(be1, scratch_ptr=777)
    cache.add(long_p, lane=lane)
    be1.query = lambda ev: False
    be1.wait = lambda ev: None
    reqs = [S.Request(0, S.chebyshev_program(PRM, 25, COEF, -8, 8), 0.0, inputs=[(hash("L0"), 25)], keep_result=True)]
    reqs += [S.Request(1 + i, S.rotsum_program(PRM, 9, 2), 0.005 * (i + 1), inputs=[(hash(("s", i)), 9)],
                       keep_result=True) for i in range(4)]
    reqs.append(S.Request(9, S.chebyshev_program(PRM, 25, COEF, -8, 8), 0.001, inputs=[(hash("L1"), 25)],
                          keep_result=True))
    srv = _MockServer(MockBackend(300000), 1)
    srv.run(reqs, realtime=True, graphs=cache)
    assert reqs[0].graph and not reqs[-1].graph
    assert all(r.graph for r in reqs[1:5])
    assert cache.be.launches == 4 and be1.launches == 1 and be1.scratch == [777]
    for r in reqs:
        assert r.result == sequential(r.program, [t for t, _ in r.inputs]), r.rid
