"""Continuous-batching frequency-solve server over the non-blocking C-ABI
(hamsx_submit_frequency / hamsx_query / hamsx_collect).

Requests -- a mesh, one frequency, a depth and headings -- arrive over time.
The server owns `slots` contexts, one HIP stream each, so admitted requests
run concurrently on the device; a latency-bound LU of one overlaps the
compute-bound assembly of another.  A request is admitted when a slot is free
and its device footprint fits the HBM budget next to the requests already
running; otherwise it waits in arrival order.  A slot whose previous request
used the same mesh keeps that mesh, and its pair sort, resident.

This has no counterpart in the reference, whose driver solves frequencies
one after another (src/HAMS_Prog.f90:613-733).
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

from .mesh import Combined
from .solver import FrequencySolver


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
    def __init__(self, slots: int = 4, hbm_budget: float = 200e9, precision: str = "mixed", device: int = 0):
        self.precision = precision
        self.device = device
        self.budget = hbm_budget
        self.slots = [_Slot() for _ in range(slots)]

    def close(self):
        for s in self.slots:
            if s.solver is not None:
                s.solver.close()
                s.solver = None

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
        elif s.key != key:
            s.solver.set_mesh(r.mesh)
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

    def run(self, requests, clock=time.perf_counter, poll_s: float = 2e-4) -> list[Result]:
        """Serve `requests` (arrival times relative to the start); returns the
        results in completion order."""
        pending = collections.deque(sorted(requests, key=lambda q: q.arrival))
        results = []
        t0 = clock()
        while pending or any(s.req is not None for s in self.slots):
            now = clock() - t0
            progressed = False
            # admit arrived requests, in arrival order, while slots and memory allow
            while pending and pending[0].arrival <= now:
                free = [s for s in self.slots if s.req is None]
                if not free:
                    break
                r = pending[0]
                # prefer a free slot that already holds this mesh
                key = r.mesh_key if r.mesh_key is not None else id(r.mesh)
                free.sort(key=lambda s: (s.key != key, s.solver is None))
                slot = next((s for s in free if self._fits(r, s)), None)
                if slot is None:
                    break
                pending.popleft()
                self._admit(slot, r, clock() - t0)
                progressed = True
            # collect finished requests
            for i, s in enumerate(self.slots):
                if s.req is not None and s.solver.query():
                    am, bd, ex, info = s.solver.collect()
                    done = clock() - t0
                    r = s.req
                    results.append(Result(r.rid, am, bd, ex, info, r.arrival, s.admitted, done, i))
                    s.req = None
                    progressed = True
            if not progressed:
                wait = poll_s
                if pending and all(s.req is None for s in self.slots):
                    wait = max(0.0, min(pending[0].arrival - (clock() - t0), 0.05))
                time.sleep(wait)
        return results


def poisson_arrivals(n: int, rate_hz: float, rng: np.random.Generator) -> np.ndarray:
    """Arrival times of a Poisson stream of n requests at rate_hz."""
    return np.cumsum(rng.exponential(1.0 / rate_hz, n))
