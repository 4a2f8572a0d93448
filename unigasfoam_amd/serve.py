"""Served form: a request loop with continuous batching over one Batch.

Requests (parametric cases) arrive over time.  Each is admitted into the
running batch as soon as its HBM footprint and a slot fit, even while other
cases are mid-run; all admitted cases advance together in the same kernels,
keep their particles and samplers resident, and leave the batch when their
sample has converged.  A case's result does not depend on what else ran
beside it (every random stream and per-cell sum is case-local), which
tests/test_gpu_batch.py checks against the plain solver.

Convergence: after `warmup` steps the samplers are cleared, then the case
samples either a fixed number of steps (enough for `samples_per_cell`
simulator samples per cell on average) or, with an `error_target`, until the
statistical error estimate of the sampled fields (uniGasVolFields.C:
1242-1252, after Hadjiconstantinou et al. 2003; `error_field` picks density,
velocity, temperature or pressure) is at most the target in every cell that
holds gas.  The error is checked every `check_every` sampled steps, from
the write-time fields of the case, and sampling stops at `max_sample_steps`
in any case.  The reported latency of a request is completion time minus
arrival time; with the simulated clock (realtime=False), where requests are
released early whenever the loop is idle, it counts from the release.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from .batch import Batch, footprint
from .core import Case
from .spec import CaseSpec


@dataclass
class Request:
    rid: int
    spec: CaseSpec
    fill: dict
    expected: int            # expected simulators
    warmup: int
    sample_steps: int        # sampled steps; with an error target, the most taken
    arrival: float = 0.0     # seconds after the start of the run
    error_target: float = 0.0    # > 0: sample until the error estimate is at most this
    error_field: str = "density_error"
    check_every: int = 10

    @property
    def total_steps(self):
        return self.warmup + self.sample_steps

    def converged(self, fields):
        """The error criterion on the write-time fields of the case."""
        err = np.asarray(fields[self.error_field])
        live = np.asarray(fields["rhoN"]) > 0.0
        return bool(live.any()) and float(err[live].max()) <= self.error_target


@dataclass
class Result:
    rid: int
    samples: np.ndarray
    counters: dict
    arrival: float
    admitted: float
    done: float
    wall_samples: np.ndarray = None

    @property
    def latency(self):
        return self.done - self.arrival


def make_request(rid, spec, fill, per_cell, warmup=50, samples_per_cell=2000, arrival=0.0, error_target=0.0,
                 error_field="density_error", check_every=10, max_sample_steps=None):
    """A request; with error_target > 0 it samples until that statistical
    error is reached (at most max_sample_steps, by default four times the
    fixed-count estimate)."""
    n = spec.mesh.n_cells * per_cell
    steps = int(math.ceil(samples_per_cell / per_cell))
    if error_target > 0.0:
        steps = int(max_sample_steps) if max_sample_steps else 4 * steps
    return Request(rid, spec, fill, n, warmup, steps, arrival, float(error_target), error_field, int(check_every))


@dataclass
class ServeStats:
    results: dict = field(default_factory=dict)
    wall: float = 0.0
    busy: float = 0.0
    particle_steps: int = 0
    rebuilds: int = 0
    max_live: int = 0

    def latencies(self):
        return np.array([r.latency for r in self.results.values()])

    def summary(self):
        lat = self.latencies()
        return {
            "cases": len(self.results),
            "particle_steps": self.particle_steps,
            "wall_s": self.wall,
            "busy_s": self.busy,
            "particle_steps_per_s": self.particle_steps / self.busy if self.busy > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
            "max_concurrent_cases": self.max_live,
        }


class Server:
    """Continuous batching: admit on arrival while the budget allows, step all
    live cases together in chunks, retire converged cases."""

    def __init__(self, precision="fp32", hbm_budget=0, max_cases=0, chunk=10, use_graphs=True,
                 realtime=True):
        self.batch = Batch(precision=precision, hbm_budget=hbm_budget, max_cases=max_cases, use_graphs=use_graphs)
        self.precision = precision
        self.chunk = chunk
        self.realtime = realtime

    def close(self):
        self.batch.close()

    def run(self, requests):
        st = ServeStats()
        pending = sorted(requests, key=lambda r: r.arrival)
        queue = []
        live = {}          # slot -> [request, steps done, admitted time]
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        released = {}      # rid -> time the request entered the queue (its latency starts there)
        while pending or queue or live:
            t = now()
            if not self.realtime:
                # simulated clock: arrivals are released as soon as the loop is
                # idle, and their latency counts from that release
                while pending and (not live or pending[0].arrival <= t):
                    released[pending[0].rid] = t
                    queue.append(pending.pop(0))
            while pending and pending[0].arrival <= t:
                released[pending[0].rid] = pending[0].arrival
                queue.append(pending.pop(0))
            while queue:
                r = queue[0]
                slot = self.batch.admit(r.spec, r.fill, r.expected)
                if slot is None:
                    break
                queue.pop(0)
                live[slot] = [r, 0, now()]
            if not live:
                if queue:
                    raise RuntimeError(f"request {queue[0].rid} does not fit an empty batch")
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - now(), 0.05)))
                continue
            st.max_live = max(st.max_live, len(live))
            # step to the next milestone of any live case (end of warmup,
            # an error check, or the end of sampling)
            k = self.chunk
            for r, done, _ in live.values():
                k = min(k, next_milestone(r, done) - done)
            tb = time.perf_counter()
            self.batch.step(k)
            for v in live.values():
                v[1] += k
            finished = []
            for slot, (r, done, t_adm) in live.items():
                if done == r.warmup and r.sample_steps > 0:
                    self.batch.case(slot).reset_samples()
                if done >= r.total_steps:
                    finished.append(slot)
                elif r.error_target > 0.0 and done > r.warmup and (done - r.warmup) % r.check_every == 0:
                    if r.converged(self.batch.case(slot).fields()):
                        finished.append(slot)
            if finished:
                self.batch.sync()
            for slot in finished:
                r, done, t_adm = live.pop(slot)
                c = self.batch.case(slot)
                cnt = c.counters()
                st.results[r.rid] = Result(r.rid, c.samples(), cnt, released[r.rid], t_adm, now(), c.wall_samples())
                st.particle_steps += cnt["particle_steps"]
                self.batch.retire(slot)
            st.busy += time.perf_counter() - tb
        self.batch.sync()
        st.wall = now()
        return st


def next_milestone(r: Request, done: int) -> int:
    """The next step count at which the loop must look at request r."""
    if done < r.warmup:
        return r.warmup
    if r.error_target > 0.0:
        return min(r.total_steps, done + r.check_every - (done - r.warmup) % r.check_every)
    return r.total_steps


def run_alone(requests, precision="fp32", use_graphs=True):
    """The same requests one at a time through the plain solver."""
    st = ServeStats()
    t0 = time.perf_counter()
    for r in requests:
        tb = time.perf_counter()
        c = Case(r.spec, precision=precision, use_graphs=use_graphs)
        c.fill(**r.fill)
        c.step(r.warmup)
        if r.sample_steps > 0:
            c.reset_samples()
        done = r.warmup
        while done < r.total_steps:
            nxt = next_milestone(r, done)
            c.step(nxt - done)
            done = nxt
            if r.error_target > 0.0 and done < r.total_steps and r.converged(c.fields()):
                break
        c.sync()
        cnt = c.counters()
        st.results[r.rid] = Result(r.rid, c.samples(), cnt, 0.0, tb - t0, time.perf_counter() - t0, c.wall_samples())
        st.particle_steps += cnt["particle_steps"]
        c.close()
        st.busy += time.perf_counter() - tb
    st.wall = time.perf_counter() - t0
    return st


def footprint_of(r: Request, precision="fp32"):
    return footprint(r.spec, r.expected, precision)
