"""Continuous-batching solve server over the batched engine.

Requests (grid, obstacle, flow, step count) arrive over time.  The loop
admits queued requests whenever the device memory budget and the slot limit
allow, folds them into the running batch (one launch per kernel for all
resident cases; each case keeps its own CFL dt), retires a case as soon as it
has taken its step count, reads its state back and frees its memory for the
next admission.  State never leaves HBM between steps.  A case's arithmetic
does not depend on its neighbours in the batch, so a result is bitwise the
same alone and in the mix.

The host never waits for the device between chunks: the next chunk of steps
is queued before the cases finished by the previous one are read back, the
engine runs admissions and read-backs of finished cases on a side stream (a
read-back waits for its case's last step only), and at most two chunks are
queued at a time, so admissions join within about one chunk.
"""
from __future__ import annotations

import collections
import dataclasses
import sys
import time

import numpy as np

from . import params as P
from .batch import BatchEngine
from .lib import FvmError


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the run
    params: P.FvmParams
    steps: int
    admitted: float = -1.0
    completed: float = -1.0
    result: np.ndarray | None = None
    rejected: str = ""       # why a request that can never be admitted was refused

    @property
    def cells(self) -> int:
        return self.params.nx * self.params.ny


class SolveServer:
    def __init__(self, device: int = 0, mem_budget_bytes: int = 200 << 30, max_active: int = 1024,
                 chunk_steps: int = 25, keep_results: bool = False):
        self.engine = BatchEngine(device, mem_budget_bytes=mem_budget_bytes, max_cases=max_active)
        self.budget = mem_budget_bytes
        self.max_active = max_active
        self.chunk = chunk_steps
        self.keep_results = keep_results

    def close(self):
        self.engine.close()

    def run(self, requests: list[Request], time_limit: float | None = None, progress_every: float | None = None) -> dict:
        """Serve `requests` (arrival times relative to the call) until all are
        done or `time_limit` seconds pass.  Returns latency and throughput stats.
        progress_every: seconds between status lines on stderr (None: quiet)."""
        future = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue: collections.deque[Request] = collections.deque()
        active: dict[int, Request] = {}
        done: list[Request] = []
        rejected: list[Request] = []
        in_use = 0
        sizes: dict[int, int] = {}
        retire: list[int] = []   # finished by the chunks queued so far, not yet read back
        t0 = time.perf_counter()
        busy = 0.0
        next_report = progress_every
        while True:
            now = time.perf_counter() - t0
            if time_limit is not None and now > time_limit:
                break
            if next_report is not None and now >= next_report:
                next_report += progress_every
                print(f"[serve] t={now:.0f}s arrived={len(requests) - len(future)} completed={len(done)} "
                      f"active={len(active)} queued={len(queue)}", file=sys.stderr, flush=True)
            while future and future[0].arrival <= now:
                queue.append(future.popleft())
            # admission: FIFO against the memory budget and the slot limit; a
            # request that would not fit even alone is refused, not kept waiting
            while queue and len(active) < self.max_active:
                r = queue[0]
                need = BatchEngine.case_bytes(r.params) + r.cells * 2  # + ghost table allowance
                if need > self.budget:
                    r.rejected = f"needs {need} bytes, more than the {self.budget}-byte budget"
                    rejected.append(queue.popleft())
                    continue
                if in_use + need > self.budget:
                    break
                try:
                    cid = self.engine.add(r.params, steps_total=r.steps)
                except FvmError as e:
                    if e.code == -5 and active:  # budget exhausted (fragmentation): retry after a retirement
                        break
                    if e.code == -5:            # an empty engine cannot hold it either
                        r.rejected = str(e)
                        rejected.append(queue.popleft())
                        continue
                    raise
                queue.popleft()
                r.admitted = time.perf_counter() - t0
                active[cid] = r
                sizes[cid] = need
                in_use += need
            if not active:
                if not future and not queue:
                    break
                if future:
                    wait = future[0].arrival - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                continue
            running = [cid for cid in active if cid not in retire]
            if running:
                # queue the next chunk first (returns once at most one earlier
                # chunk is still queued): the device stays busy while the
                # host reads back and admits
                tb = time.perf_counter()
                self.engine.advance(self.chunk)
                busy += time.perf_counter() - tb
            # cases finished by the chunks queued before this one
            for cid in retire:
                r = active.pop(cid)
                state = self.engine.state(cid)
                if self.keep_results:
                    r.result = state
                r.completed = time.perf_counter() - t0
                self.engine.remove(cid)
                in_use -= sizes.pop(cid)
                done.append(r)
            retire = [cid for cid in active if self.engine.status(cid, with_time=False)[2]]
        self.engine.synchronize()
        wall = time.perf_counter() - t0
        lat = np.array([r.completed - r.arrival for r in done]) if done else np.zeros(1)
        work = sum(r.cells * r.steps for r in done)
        # steps taken by the cases still resident when the run stopped
        partial = sum(r.cells * self.engine.status(cid, with_time=False)[0] for cid, r in active.items())
        return {
            "completed": len(done),
            "submitted": len(requests),
            "rejected": len(rejected),
            "left_in_queue": len(queue) + len(future),
            "left_active": len(active),
            "wall_s": wall,
            "host_wait_in_advance_s": busy,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()),
            "cell_updates_per_s": work / wall if wall > 0 else 0.0,
            "executed_cell_updates_per_s": (work + partial) / wall if wall > 0 else 0.0,
            "done": done,
        }


def poisson_stream(rate_per_s: float, duration_s: float, seed: int = 5, request_fn=P.served_request,
                   step_scale: float = 1.0) -> list[Request]:
    """Seeded Poisson arrivals at `rate_per_s` over `duration_s` seconds."""
    rng = np.random.default_rng(seed)
    t, out, rid = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate_per_s)
        if t > duration_s:
            break
        p, steps = request_fn(rng)
        out.append(Request(rid, t, p, max(1, int(round(steps * step_scale)))))
        rid += 1
    return out
