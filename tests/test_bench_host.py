"""Host logic of bench.py without a GPU: the pipelined served steps answer every request of every
step exactly once, keep at most the given number of steps in flight, and submit steps in order."""
import importlib.util
import random
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def load_bench():
    spec = importlib.util.spec_from_file_location('bench_under_test', ROOT / 'bench.py')
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


class Result:
    def __init__(self, ticket, req):
        self.ticket, self.status, self.result = ticket, 0, req


class FakeServer:
    """Answers submitted requests in random order, a few per poll; records in-flight steps."""

    def __init__(self, seed):
        self.rng = random.Random(seed)
        self.pending, self.next, self.order, self.req_of = [], 0, [], {}

    def submit_batch(self, reqs):
        self.order.append(list(reqs))
        out = []
        for r in reqs:
            self.pending.append(self.next)
            self.req_of[self.next] = r
            out.append(self.next)
            self.next += 1
        return out

    def poll(self, max_results, min_results, timeout_ms):
        self.rng.shuffle(self.pending)
        k = min(len(self.pending), self.rng.randint(1, 7))
        got, self.pending = self.pending[:k], self.pending[k:]
        return [Result(t, self.req_of[t]) for t in got]


def test_served_steps_answers_every_request_once():
    bench = load_bench()
    for seed in range(5):
        rng = random.Random(seed)
        steps = [[(rng.randrange(12), rng.randrange(10 ** 6)) for _ in range(rng.randint(1, 40))] for _ in range(9)]
        for depth in (1, 2, 4, 20):
            srv = FakeServer(seed)
            got, dts, wall = bench.served_steps(srv, steps, depth)
            assert len(got) == len(steps) and len(dts) == len(steps) and wall >= 0
            for reqs, g in zip(steps, got):
                assert [r.result for r in g] == reqs     # each request's own answer, in request order
            assert srv.order == steps                    # submitted in step order, each once
            assert not srv.pending


def test_served_steps_keeps_at_most_depth_in_flight():
    bench = load_bench()
    steps = [[(0, i + 100 * s) for i in range(5)] for s in range(8)]

    class Watch(FakeServer):
        def __init__(self):
            super().__init__(3)
            self.max_open, self.answered = 0, set()

        def submit_batch(self, reqs):
            out = super().submit_batch(reqs)
            open_steps = sum(1 for st in self.order if not all(r in self.answered for r in st))
            self.max_open = max(self.max_open, open_steps)
            return out

        def poll(self, max_results, min_results, timeout_ms):
            res = super().poll(max_results, min_results, timeout_ms)
            self.answered.update(r.result for r in res)
            return res

    for depth in (1, 3):
        w = Watch()
        bench.served_steps(w, steps, depth)
        assert w.max_open <= depth
