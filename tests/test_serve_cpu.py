"""Host logic of the serving loop (lbamd.serve) against a stand-in for the
device batch: admission by footprint in arrival order with look-ahead,
continuous batching (admissions between intervals while others run),
retirement, latency bookkeeping."""
import numpy as np

import lbamd
from lbamd.serve import Request, Server, run_stream


class FakeBatch:
    """Pool of `total` bytes; a solve finishes after `work` intervals."""

    def __init__(self, total, work):
        self.total, self.used, self.work = total, 0, work
        self.res = {}
        self.done = []
        self.log = []
        self.interval = 0

    def status(self):
        return {"largest_free": self.total - self.used, "resident": len(self.res)}

    def _add(self, tag, nbytes):
        if nbytes > self.total - self.used:
            return False
        self.used += nbytes
        self.res[tag] = [self.work[tag], nbytes]
        self.log.append(("add", tag, self.interval))
        return True

    def add(self, geo, periodic, tau, force, tol, max_steps, tag):
        return self._add(tag, lbamd.batch_footprint(lbamd.D3Q19, lbamd.FP32S, geo, periodic)[0])

    def add_twophase(self, geo, periodic, rho0, wett, tau0, tau1, sigma, beta, force, tol, max_steps, tag):
        return self._add(tag, lbamd.batch_footprint(lbamd.D3Q19, lbamd.FP32S, geo, periodic, 2)[0])

    def advance(self, n):
        self.interval += n
        for tag in list(self.res):
            self.res[tag][0] -= n
            if self.res[tag][0] <= 0:
                self.used -= self.res[tag][1]
                del self.res[tag]
                self.done.append({"tag": tag, "k": float(tag), "k_phase": (0.0, 0.0), "steps": 100, "status": 0})
                self.log.append(("done", tag, self.interval))
        return 0

    def collect(self):
        out, self.done = self.done, []
        return out

    def close(self):
        pass


def geo(n, seed):
    g, _ = lbamd.sphere_pack((n, n, n), 2.5, 0.45, seed, periodic="x", walls=True)
    return g


def test_admission_order_lookahead_and_latency():
    gs = [geo(16, 1), geo(24, 2), geo(16, 3), geo(16, 4)]
    sizes = [lbamd.batch_footprint(lbamd.D3Q19, lbamd.FP32S, g, "x")[0] for g in gs]
    two = lbamd.batch_footprint(lbamd.D3Q19, lbamd.FP32S, gs[2], "x", 2)[0]
    assert two > sizes[2]
    # the large second request fits the pool alone, but not beside the first
    pool = sizes[1] + 1
    assert sizes[0] + two <= pool < sizes[0] + sizes[1]
    fb = FakeBatch(pool, {0: 3, 1: 2, 2: 1, 3: 1})
    srv = Server(pool, batch=fb, relperm_precision=lbamd.FP32S)
    reqs = [Request(0, gs[0]), Request(1, gs[1]),
            Request(2, gs[2], kind="relperm", rho0=(gs[2] > 0).astype(float)), Request(3, gs[3])]
    for r in reqs:
        srv.submit(r)
    results = srv.step(1)
    # 0 admitted; 1 does not fit and is passed over; 2 (two-phase) fits behind it
    assert [e[1] for e in fb.log if e[0] == "add"][:2] == [0, 2]
    while not srv.idle():
        results += srv.step(1)
    assert sorted(r["tag"] for r in results) == [0, 1, 2, 3]
    assert all(r["latency"] >= 0 for r in results)
    # nothing was ever over-committed, and everything was freed
    assert fb.used == 0
    kinds = {r["tag"]: r["kind"] for r in results}
    assert kinds[2] == "relperm" and kinds[0] == "perm"


def test_request_larger_than_pool_is_refused():
    g = geo(16, 5)
    size = lbamd.batch_footprint(lbamd.D3Q19, lbamd.FP32S, g, "x")[0]
    srv = Server(size - 1, batch=FakeBatch(size - 1, {0: 1}))
    try:
        srv.submit(Request(0, g))
        raise AssertionError("accepted a request larger than the pool")
    except ValueError:
        pass


def test_stream_replay_admits_while_others_run():
    gs = [geo(16, s) for s in range(6)]
    work = {i: 4 for i in range(6)}
    pool = 10 << 30
    fb = FakeBatch(pool, work)
    srv = Server(pool, batch=fb, relperm_precision=lbamd.FP32S)
    reqs = [Request(i, g, t_arrival=0.0 if i < 3 else 0.01) for i, g in enumerate(gs)]
    done, wall = run_stream(srv, reqs)
    assert sorted(d["tag"] for d in done) == list(range(6))
    adds = {e[1]: e[2] for e in fb.log if e[0] == "add"}
    fin = {e[1]: e[2] for e in fb.log if e[0] == "done"}
    # every solve stayed resident for exactly its work, wherever it joined
    assert np.all([fin[i] - adds[i] == 4 for i in range(6)])
