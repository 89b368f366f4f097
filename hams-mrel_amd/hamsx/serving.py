"""Continuous-batching frequency-solve server over the non-blocking C-ABI
(hamsx_submit_frequency / hamsx_query / hamsx_collect).

Requests -- a mesh, one frequency, a depth and headings -- arrive over time.
The server owns `slots` contexts, one HIP stream each, so admitted requests
run concurrently on the device; a latency-bound LU of one overlaps the
compute-bound assembly of another.  A request is admitted when a slot is free
and its device footprint fits the HBM budget next to the requests already
running; otherwise it waits in arrival order.  A slot whose previous request
used the same mesh keeps that mesh, and its pair sort, resident, and replays
the solve's recorded HIP graph for every later frequency at the same depth
(HAMSX_OPT_GRAPHS).  The loop sleeps on the library's completion queue
(hamsx_queue_wait): it wakes when a solve finishes or the next request is
due, never to poll.  Per-size LU options come from the autotuner's table
(hamsx/tuning.json, written by hamsx/autotune.py) when one is present.

This has no counterpart in the reference, whose driver solves frequencies
one after another (src/HAMS_Prog.f90:613-733).
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

from . import autotune
from .mesh import Combined
from .solver import CompletionQueue, FrequencySolver


@dataclass
class Request:
    rid: int
    mesh: Combined
    w1: float                      # rad/s
    depth: float                   # < 0: infinite
    headings_rad: np.ndarray
    arrival: float = 0.0           # seconds after the server's start
    mesh_key: object = None        # requests with equal keys share a resident mesh


@dataclass
class Result:
    rid: int
    amas: np.ndarray
    bdmp: np.ndarray
    exfc: np.ndarray
    info: int
    arrival: float
    admitted: float
    done: float
    slot: int

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def footprint_bytes(n: int, nrhs: int, precision: str) -> int:
    """Device bytes a context needs for an N-panel solve: A and S
    (complex128), the right-hand sides, the finite-depth pair list and mode
    tables, and the complex64 copy in mixed precision."""
    b = 32 * n * n + 4 * n * n + 64 * n * max(nrhs, 1) + 3200 * n
    if precision == "mixed":
        b += 8 * n * n + 48 * n * max(nrhs, 1)
    return b


@dataclass
class _Slot:
    solver: FrequencySolver | None = None
    key: object = None
    req: Request | None = None
    admitted: float = 0.0
    bytes: int = 0                 # high-water footprint of this context


class Server:
    def __init__(self, slots: int = 4, hbm_budget: float = 200e9, precision: str = "mixed", device: int = 0,
                 tuning: dict | None = None):
        self.precision = precision
        self.device = device
        self.budget = hbm_budget
        self.slots = [_Slot() for _ in range(slots)]
        self.tuning = autotune.load_table() if tuning is None else tuning
        self.queue = CompletionQueue()

    def close(self):
        for s in self.slots:
            if s.solver is not None:
                s.solver.close()
                s.solver = None
        if self.queue is not None:
            self.queue.close()
            self.queue = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _in_use(self) -> int:
        return sum(s.bytes for s in self.slots if s.solver is not None)

    def _admit(self, s: _Slot, r: Request, now: float):
        need = footprint_bytes(r.mesh.nelem, 6 * r.mesh.nbody + len(r.headings_rad), self.precision)
        key = r.mesh_key if r.mesh_key is not None else id(r.mesh)
        if s.solver is None:
            s.solver = FrequencySolver(r.mesh, device=self.device, precision=self.precision)
            s.solver.set_queue(self.queue, self.slots.index(s))
        elif s.key != key:
            s.solver.set_mesh(r.mesh)
        if s.key != key:
            opts = autotune.options_for(self.tuning, r.mesh.nelem, r.depth, self.precision)
            if opts:
                s.solver.set_options(**opts)
        s.key = key
        s.bytes = max(s.bytes, need)
        s.solver.set_waves(r.headings_rad)
        s.solver.submit(r.w1, r.depth)
        s.req = r
        s.admitted = now

    def _fits(self, r: Request, s: _Slot) -> bool:
        need = footprint_bytes(r.mesh.nelem, 6 * r.mesh.nbody + len(r.headings_rad), self.precision)
        extra = max(0, need - s.bytes) if s.solver is not None else need
        return self._in_use() + extra <= self.budget

    def run(self, requests, clock=time.perf_counter) -> list[Result]:
        """Serve `requests` (arrival times relative to the start); returns the
        results in completion order."""
        pending = collections.deque(sorted(requests, key=lambda q: q.arrival))
        results = []
        t0 = clock()
        while pending or any(s.req is not None for s in self.slots):
            now = clock() - t0
            # admit arrived requests, in arrival order, while slots and memory allow
            blocked = False
            while pending and pending[0].arrival <= now:
                free = [s for s in self.slots if s.req is None]
                if not free:
                    blocked = True
                    break
                r = pending[0]
                # prefer a free slot that already holds this mesh
                key = r.mesh_key if r.mesh_key is not None else id(r.mesh)
                free.sort(key=lambda s: (s.key != key, s.solver is None))
                slot = next((s for s in free if self._fits(r, s)), None)
                if slot is None:
                    if all(s.req is None for s in self.slots):
                        raise ValueError(f"request {r.rid} does not fit the HBM budget even alone")
                    blocked = True
                    break
                pending.popleft()
                self._admit(slot, r, clock() - t0)
            busy = any(s.req is not None for s in self.slots)
            if not busy:
                if pending:  # nothing on the device: sleep until the next arrival
                    time.sleep(max(0.0, pending[0].arrival - (clock() - t0)))
                continue
            # sleep on the completion queue until a solve finishes, or until the
            # next request is due when it could be admitted on arrival
            timeout = None
            if pending and not blocked:
                timeout = max(0.0, pending[0].arrival - (clock() - t0))
            tag = self.queue.wait(timeout)
            if tag is None:
                continue
            s = self.slots[tag]
            am, bd, ex, info = s.solver.collect()
            r = s.req
            results.append(Result(r.rid, am, bd, ex, info, r.arrival, s.admitted, clock() - t0, tag))
            s.req = None
        return results


def poisson_arrivals(n: int, rate_hz: float, rng: np.random.Generator) -> np.ndarray:
    """Arrival times of a Poisson stream of n requests at rate_hz."""
    return np.cumsum(rng.exponential(1.0 / rate_hz, n))
