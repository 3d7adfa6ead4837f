"""Serving form: continuous batching of permeability requests on one GPU.

A Server owns one Batch pool (lb_batch).  Requests wait in arrival order;
between check intervals every waiting request whose footprint fits into the
pool's free memory is admitted (first fit over the queue, so one large
request does not hold back smaller ones behind it), all resident solves are
advanced together, and finished solves are returned with their latency
(completion time - arrival time).  State stays resident until convergence.
"""
import time
from collections import deque

from .ffi import D3Q19, FP32S, Batch, batch_footprint


class Request:
    """A permeability request (kind "perm": single-phase BGK, tau) or a
    relative-permeability request (kind "relperm": colour-gradient two-phase
    with the phase-0 density rho0 per cell, wettability, tau0, tau1, sigma,
    beta)."""
    __slots__ = ("tag", "geo", "periodic", "tau", "force", "tol", "max_steps", "t_arrival", "bytes", "kind",
                 "rho0", "wett", "tau0", "tau1", "sigma", "beta")

    def __init__(self, tag, geo, periodic="x", tau=1.0, force=(1e-6, 0.0, 0.0), tol=1e-6, max_steps=100000,
                 t_arrival=None, kind="perm", rho0=None, wett=1.0, tau0=1.0, tau1=1.0, sigma=0.01, beta=0.7):
        self.tag = tag
        self.geo = geo
        self.periodic = periodic
        self.tau = tau
        self.force = force
        self.tol = tol
        self.max_steps = max_steps
        self.t_arrival = t_arrival
        self.bytes = None
        self.kind = kind
        self.rho0 = rho0
        self.wett = wett
        self.tau0, self.tau1, self.sigma, self.beta = tau0, tau1, sigma, beta

    def add_to(self, batch):
        if self.kind == "relperm":
            return batch.add_twophase(self.geo, self.periodic, self.rho0, self.wett, self.tau0, self.tau1,
                                      self.sigma, self.beta, self.force, self.tol, self.max_steps, self.tag)
        return batch.add(self.geo, self.periodic, self.tau, self.force, self.tol, self.max_steps, self.tag)


class Server:
    """precision: storage of the permeability solves; relperm_precision: of the
    relative-permeability solves (FP32S by default, whatever `precision` is:
    the colour-gradient step computes in FP64 on FP32 storage and stays
    within 1e-4 of FP64 over 5e4 steps, while FP16S storage moves the
    interfaces and the phase permeabilities by 1e-3..1e-2)."""

    def __init__(self, pool_bytes, lattice=D3Q19, precision=FP32S, check_every=100, device=0, lookahead=64,
                 batch=None, relperm_precision=FP32S):
        # batch: an object with the Batch interface (add, add_twophase,
        # advance, collect, status, close); a device Batch by default
        if batch is None:
            batch = Batch(lattice, precision, pool_bytes, check_every, device)
            batch.set_precision("relperm", relperm_precision)
        self.batch = batch
        self.lattice, self.precision = lattice, precision
        prec = getattr(batch, "prec", None)  # per kind, when the batch keeps it
        self.relperm_precision = prec[1] if prec is not None else relperm_precision
        self.pool_bytes = pool_bytes
        self.lookahead = lookahead
        self.waiting = deque()
        self.resident = {}
        self.done = []
        self.clock = time.perf_counter

    def submit(self, req):
        if req.t_arrival is None:
            req.t_arrival = self.clock()
        if req.bytes is None:
            two = req.kind == "relperm"
            prec = self.relperm_precision if two else self.precision
            req.bytes = batch_footprint(self.lattice, prec, req.geo, req.periodic, 2 if two else 1)[0]
        if req.bytes > self.pool_bytes:
            raise ValueError("request %r needs %d bytes, more than the whole pool" % (req.tag, req.bytes))
        self.waiting.append(req)

    def admit(self):
        """Admit waiting requests that fit, in arrival order, looking past
        ones that do not fit yet (at most `lookahead` deep)."""
        free = self.batch.status()["largest_free"]
        kept = deque()
        n = 0
        while self.waiting:
            r = self.waiting.popleft()
            if n < self.lookahead and r.bytes <= free and r.add_to(self.batch):
                r.geo = r.rho0 = None  # the device holds the solve now
                self.resident[r.tag] = r
                free = self.batch.status()["largest_free"]
            else:
                kept.append(r)
            n += 1
        self.waiting = kept

    def step(self, intervals=1):
        """Advance all resident solves by up to `intervals` check intervals and
        collect the finished ones.  Waiting requests are admitted while the
        intervals run on the device (their host-side set-up and uploads
        overlap the steps) and join at the next call; with nothing resident
        they are admitted first."""
        if not self.resident:
            self.admit()
            if not self.resident:
                return []
        launch = getattr(self.batch, "launch", None)
        if launch is None:
            self.admit()
            self.batch.advance(intervals)
        else:
            launch(intervals)
            self.admit()
            self.batch.wait()
        now = self.clock()
        out = []
        for r in self.batch.collect():
            req = self.resident.pop(r["tag"])
            r["latency"] = now - req.t_arrival
            r["kind"] = req.kind
            out.append(r)
        self.done.extend(out)
        return out

    def idle(self):
        return not self.waiting and not self.resident

    def close(self):
        self.batch.close()


def run_stream(server, requests, poll_sleep=1e-3):
    """Replay requests (with t_arrival offsets in seconds from the start) in
    real time through the server; returns (results, wall seconds)."""
    pending = deque(sorted(requests, key=lambda r: r.t_arrival))
    t0 = server.clock()
    offsets = {}
    while pending or not server.idle():
        now = server.clock() - t0
        while pending and pending[0].t_arrival <= now:
            r = pending.popleft()
            offsets[r.tag] = r.t_arrival
            r.t_arrival = t0 + r.t_arrival
            server.submit(r)
        if server.idle():
            if pending:
                time.sleep(min(poll_sleep, max(0.0, pending[0].t_arrival - (server.clock() - t0))))
            continue
        server.step(1)
    return server.done, server.clock() - t0
