"""Serving: a request loop that admits solves as they arrive, by device memory.

Requests are dendrite (Kobayashi-93, the batched-sweep parameter draws) and
reactive Cahn-Hilliard solves (Gamma log-uniform in [1e-4, 1e-2], cbar uniform
in [-0.2, 0.2]); grid sides are log-uniform, step counts log-uniform, arrivals
Poisson. The loop is continuous batching:

  * a request that has arrived is admitted (FIFO) as soon as its device
    footprint fits the memory budget; admission creates the instance and
    uploads its initial state on the library's set-up stream, overlapping the
    steps already running for the resident requests;
  * every tick advances all resident requests together (one merged-worklist
    launch per model per step, graph-replayed) by
    min(chunk, fewest steps any resident request still needs), so each
    request runs exactly its own step count;
  * finished requests are downloaded and destroyed, freeing their memory.

A request's result does not depend on the others: instances share launches but
no data, so its fields are bit-identical alone and in the mix.

The arrival rate can be steered during the run toward a target memory
demand (an inhomogeneous Poisson process whose rate is multiplied after every
tick by exp(gain (target - demand) dt), by at most e^0.25 either way per tick and within
[rate/8, 8 rate]). Demand counts the footprint of the queued requests as well as
the resident ones: steering on occupancy alone keeps lowering the rate while a
backlog drains and raises it only after the memory has emptied, so the rate
swings between its clamps and the device idles between swings.
"""
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import abi, batch, configs


@dataclass
class Request:
    rid: int
    kind: str            # "dendrite" or "ch"
    side: int
    steps: int
    arrival: float = 0.0  # seconds after the start of the run
    delta: float = 0.02
    aniso_j: int = 4
    latent_k: float = 1.6
    seeds: list = field(default_factory=list)
    react_gamma: float = 1e-3
    cbar: float = 0.0
    noise_seed: int = 1234

    def params(self, tiles, threads):
        tile = tiles[self.kind]
        if self.kind == "dendrite":
            return configs.dendrite_k93_params(n=self.side, tile=tile, delta=self.delta, aniso_j=self.aniso_j,
                                               latent_k=self.latent_k, threads=threads.get(self.kind, 0))
        return configs.ch_params(n=self.side, tile=tile, react_gamma=self.react_gamma, cbar=self.cbar,
                                 threads=threads.get(self.kind, 0))

    def fields(self):
        if self.kind == "dendrite":
            return list(configs.dendrite_k93_ic(self.side, self.side, seeds=self.seeds))
        return [configs.ch_ic(self.side, self.side, cbar=self.cbar, seed=self.noise_seed)]

    @property
    def cells(self):
        return self.side * self.side


def make_requests(n, master_seed=7, side_lo=256, side_hi=4096, steps_lo=1000, steps_hi=50000, multiple=1):
    """n requests (arrival times unset): 50% dendrite, 50% CH, sides and step
    counts log-uniform. Draws come from splitmix64 streams of master_seed."""
    out = []
    for i in range(n):
        u = configs.uniform01(master_seed * 7919 + i, 14)
        side = batch.log_uniform_side(u[1], side_lo, side_hi, multiple)
        steps = int(round(math.exp(math.log(steps_lo) + u[2] * (math.log(steps_hi) - math.log(steps_lo)))))
        if u[0] < 0.5:
            nseed = 1 + min(2, int(3 * u[7]))
            seeds = [(16 + int(u[8 + 2 * k] * (side - 32)), 16 + int(u[9 + 2 * k] * (side - 32)))
                     for k in range(nseed)]
            out.append(Request(i, "dendrite", side, steps, delta=0.01 + 0.04 * u[3], aniso_j=4 if u[4] < 0.5 else 6,
                               latent_k=1.2 + 0.8 * u[5], seeds=seeds))
        else:
            g = math.exp(math.log(1e-4) + u[3] * (math.log(1e-2) - math.log(1e-4)))
            out.append(Request(i, "ch", side, steps, react_gamma=g, cbar=-0.2 + 0.4 * u[4],
                               noise_seed=int(1 + u[6] * 2 ** 31)))
    return out


def poisson_times(n, rate, seed=11):
    """n arrival times of a Poisson process of the given rate (1/s)."""
    u = configs.uniform01(seed, n)
    return np.cumsum(-np.log1p(-u) / rate)


def spread_order(n):
    """0 .. n-1 in van der Corput (bit-reversed) order: every prefix is spread over
    the whole range, so a time-limited pass over requests in this order samples the
    whole stream rather than its first requests."""
    bits = max(1, (n - 1).bit_length())
    out = []
    for k in range(1 << bits):
        r = int(format(k, "0%db" % bits)[::-1], 2)
        if r < n:
            out.append(r)
    return out


def percentile(values, q):
    """Nearest-rank percentile (q in [0, 100]) of a non-empty sequence."""
    v = sorted(values)
    k = max(1, int(math.ceil(q / 100.0 * len(v))))
    return v[k - 1]


@dataclass
class Completion:
    rid: int
    kind: str
    arrival: float
    admitted: float
    done: float
    steps: int
    cells: int
    active_updates: int
    fields: list = None

    @property
    def latency(self):
        return self.done - self.arrival


class LibraryBackend:
    """The DBA library as the server's device: one Context."""

    def __init__(self, ctx):
        self.ctx = ctx

    def footprint(self, params):
        return self.ctx.footprint(params)

    def admit(self, params, fields):
        inst = self.ctx.create(params)
        for f, a in enumerate(fields):
            inst.upload(f, a)
        return inst

    def step(self, handles, n, chunk, mode=None):
        if mode is None:
            self.ctx.step(handles, n, graph_chunk=chunk)
            return
        self.ctx.set_single_launch(mode)
        try:
            self.ctx.step(handles, n, graph_chunk=chunk)
        finally:
            self.ctx.set_single_launch(1)

    def sync(self):
        self.ctx.sync()

    # GPU time of a tick: events on the library stream around the tick's launches
    # (read after the tick's sync; an idle GPU runs the first event at once, so
    # the span is the tick's busy time)
    def tick_begin(self):
        self.ctx.marker(4)

    def tick_end(self):
        self.ctx.marker(5)

    def tick_gpu_s(self):
        return self.ctx.elapsed_ms(4, 5) / 1e3

    def updates(self, h):
        return h.info().cell_updates

    def download(self, h, nfields):
        return [h.download(f) for f in range(nfields)]

    def release(self, h):
        h.destroy()


class Server:
    """Continuous-batching loop over a backend (LibraryBackend on the GPU)."""

    def __init__(self, backend, capacity_bytes, chunk=100, tiles=None, threads=None, keep_fields=False,
                 target_occupancy=None, gain=2.0, log=None, modes=None):
        self.b = backend
        # single-launch mode per request kind (dba_set_single_launch), or None:
        # all kinds in one call in the library's default mode
        self.modes = modes
        self.capacity = int(capacity_bytes)
        self.chunk = chunk
        self.tiles = tiles or {"dendrite": 16, "ch": 32}
        self.threads = threads or {}
        self.keep_fields = keep_fields
        self.target = target_occupancy
        self.gain = gain
        self.log = log
        self.used = 0
        self.max_used = 0
        self.busy_s = 0.0     # GPU time of the ticks (backends that time them)
        self.sync_s = 0.0     # host time spent waiting for ticks
        self.occupancy = []   # (tick wall time, used / capacity)
        self.ticks = 0
        self.rate_trace = []

    def run(self, requests, rate=None, clock=time.perf_counter, sleep=time.sleep, max_wall=None):
        """Serve `requests`. With rate=None they arrive at their .arrival times;
        otherwise arrivals are drawn on line from a Poisson process of that rate
        (steered toward target_occupancy if set), in request order. Returns
        the completions in completion order."""
        t0 = clock()
        todo = deque(requests)
        rate0 = rate
        if rate is not None:
            u_iter = iter(configs.uniform01(99, len(requests) + 1))
            next_t = -math.log1p(-next(u_iter)) / rate
        queue = deque()
        resident = {}   # rid -> [handle, steps left, request, admitted at, footprint]
        done = []
        fps = {}
        last_tick = 0.0

        def arrivals(now):
            nonlocal next_t
            while todo:
                if rate is None:
                    if todo[0].arrival > now:
                        break
                else:
                    if next_t > now:
                        break
                    todo[0].arrival = next_t
                    next_t += -math.log1p(-next(u_iter)) / rate
                r = todo.popleft()
                if r.rid not in fps:
                    fps[r.rid] = self.b.footprint(r.params(self.tiles, self.threads))
                queue.append(r)

        def admit():
            """FIFO admission while the head of the queue fits the budget."""
            new = {}
            while queue:
                r = queue[0]
                if r.rid not in fps:
                    fps[r.rid] = self.b.footprint(r.params(self.tiles, self.threads))
                fp = fps[r.rid]
                if fp > self.capacity:
                    raise ValueError(f"request {r.rid} needs {fp} B, more than the {self.capacity} B budget")
                if self.used + fp > self.capacity:
                    break
                queue.popleft()
                h = self.b.admit(r.params(self.tiles, self.threads), r.fields())
                self.used += fp
                self.max_used = max(self.max_used, self.used)
                new[r.rid] = [h, r.steps, r, clock() - t0, fp]
            return new

        last_log = 0.0
        while todo or queue or resident:
            now = clock() - t0
            if max_wall is not None and now > max_wall:
                break
            if self.log and now - last_log >= 10.0:
                last_log = now
                self.log(f"serve t={now:.0f}s done={len(done)} resident={len(resident)} queued={len(queue)} "
                         f"used={self.used / self.capacity:.2f}" + (f" rate={rate:.2f}/s" if rate else ""))
            arrivals(now)
            k = 0
            timed = hasattr(self.b, "tick_begin")
            if resident:
                # one tick for the resident set, asynchronous on the device
                k = min(self.chunk, min(v[1] for v in resident.values()))
                if timed:
                    self.b.tick_begin()
                if self.modes is None:
                    self.b.step([v[0] for v in resident.values()], k, self.chunk)
                else:
                    for kind in sorted({v[2].kind for v in resident.values()}):
                        hs = [v[0] for v in resident.values() if v[2].kind == kind]
                        self.b.step(hs, k, self.chunk, self.modes.get(kind, 1))
                if timed:
                    self.b.tick_end()
            # admissions overlap the tick: new instances are set up on another stream
            incoming = admit()
            if k:
                ts = clock()
                self.b.sync()
                self.sync_s += clock() - ts
                if timed:
                    self.busy_s += self.b.tick_gpu_s()
                self.ticks += 1
                now = clock() - t0
                self.occupancy.append((now, self.used / self.capacity))
                for rid in [rid for rid, v in resident.items() if v[1] == k]:
                    h, _, r, adm, fp = resident.pop(rid)
                    upd = self.b.updates(h)
                    flds = self.b.download(h, 2 if r.kind == "dendrite" else 1) if self.keep_fields else None
                    self.b.release(h)
                    self.used -= fp
                    done.append(Completion(rid, r.kind, r.arrival, adm, clock() - t0, r.steps, r.cells, upd, flds))
                for v in resident.values():
                    v[1] -= k
                if rate is not None and self.target is not None:
                    demand = (self.used + sum(fps[q.rid] for q in queue)) / self.capacity
                    f = math.exp(min(0.25, max(-0.25, self.gain * (self.target - demand) * (now - last_tick))))
                    rate = min(8 * rate0, max(rate0 / 8, rate * f))
                    self.rate_trace.append((now, rate))
                last_tick = now
            resident.update(incoming)
            if not resident and not queue and todo:
                wait = (todo[0].arrival if rate is None else next_t) - (clock() - t0)
                if wait > 0:
                    sleep(min(wait, 0.05))
        self.wall = clock() - t0
        return done


def report(done, wall, occupancy, busy_s=None):
    lat = [c.latency for c in done]
    upd = sum(c.active_updates for c in done)
    full = sum(c.cells * c.steps for c in done)
    occ = np.array([o for _, o in occupancy]) if occupancy else np.zeros(1)
    return {
        "completed": len(done),
        "p50_latency_s": percentile(lat, 50) if lat else None,
        "p99_latency_s": percentile(lat, 99) if lat else None,
        "mean_latency_s": float(np.mean(lat)) if lat else None,
        "active_updates": int(upd),
        "active_rate": upd / wall if wall > 0 else None,
        "full_rate": full / wall if wall > 0 else None,
        "occupancy_mean": float(occ.mean()),
        "occupancy_p10": float(np.percentile(occ, 10)),
        "occupancy_p90": float(np.percentile(occ, 90)),
        "wall_s": wall,
        "device_busy_fraction": (busy_s / wall) if (busy_s is not None and wall > 0) else None,
    }
