"""Batched parameter sweeps: many independent dendrite instances stepped together.

All instances of one model, tile and workgroup size form one group in the
library: each step is ONE stencil launch over the merged cross-instance
worklist (ascending tile order per instance, instances in call order) plus ONE
finalize launch, replayed from a captured hipGraph. The alternatives run one
instance at a time through the same library: `plain` launches one step and
waits for it per time step, the shape of the reference's
one-simulation-per-process loop (example_Dendritic_growth/kernel.cu:913-1011);
`graph` replays captured chunks of steps.

The sweep (64 instances by default):
  side      log-uniform in [256, 2048], rounded to the nearest cell; a side one
            more than a multiple of 8 becomes one cell longer (a partial last
            tile must be two cells or more for tiles of 8, 16 and 32)
  delta     uniform in [0.01, 0.05]
  j         4 or 6 with equal probability
  K         uniform in [1.2, 2.0]
  seeds     1-3 disks of radius 5 cells, centres uniform in [16, side - 16)
  the rest  the 512^2 Kobayashi dendrite: dx 0.03, dt 1e-4, tau 3e-4,
            eps_bar 0.01, theta0 0, alpha 0.9, gamma 10, T_eq 1, T = 0,
            zero-flux walls, Metric_p 0.01, Metric_T 1e-5.
Every draw comes from splitmix64 streams of one master seed, so the sweep is
the same on every machine.
"""
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import abi, configs

TILES = (8, 16, 32)
THREADS = (64, 128, 256)
SINGLE_LAUNCH_MODES = (1, 2)


@dataclass
class SweepItem:
    index: int
    side: int
    delta: float
    aniso_j: int
    latent_k: float
    seeds: list = field(default_factory=list)

    def params(self, tile=16, threads=0, mode=abi.MODE_DBA, history_cap=0):
        p = configs.dendrite_k93_params(n=self.side, tile=tile, mode=mode, delta=self.delta, aniso_j=self.aniso_j,
                                        latent_k=self.latent_k, threads=threads, history_cap=history_cap)
        return p

    def fields(self):
        return configs.dendrite_k93_ic(self.side, self.side, seeds=self.seeds)

    @property
    def cells(self):
        return self.side * self.side


def log_uniform_side(u, lo, hi, multiple=1):
    """exp(log lo + u (log hi - log lo)) rounded to the nearest multiple, kept in
    [lo, hi]; a side of 8 k + 1 cells becomes 8 k + 2 (the library refuses a
    partial last tile one cell wide, and 8 | 16 | 32)."""
    s = math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))
    s = int(round(s / multiple)) * multiple
    s = max(lo, min(hi, s))
    if s % 8 == 1:
        s += 1
    return s


def make_sweep(n=64, master_seed=2024, side_lo=256, side_hi=2048, multiple=1):
    items = []
    for i in range(n):
        # one stream per instance: side, delta, j, K, seed count, 3 x (cx, cy)
        u = configs.uniform01(master_seed * 1000003 + i, 11)
        side = log_uniform_side(u[0], side_lo, side_hi, multiple)
        nseed = 1 + min(2, int(3 * u[4]))
        seeds = [(16 + int(u[5 + 2 * k] * (side - 32)), 16 + int(u[6 + 2 * k] * (side - 32))) for k in range(nseed)]
        items.append(SweepItem(index=i, side=side, delta=0.01 + 0.04 * u[1], aniso_j=4 if u[2] < 0.5 else 6,
                               latent_k=1.2 + 0.8 * u[3], seeds=seeds))
    return items


@dataclass
class RunStats:
    wall_s: float
    gpu_ms: float
    steps: int
    active_updates: int
    full_updates: int

    @property
    def active_rate(self):
        return self.active_updates / self.wall_s

    @property
    def full_rate(self):
        return self.full_updates / self.wall_s


def _create(ctx, items, tile, threads, fields_cache=None):
    insts = []
    for it in items:
        inst = ctx.create(it.params(tile=tile, threads=threads))
        phi, T = fields_cache[it.index] if fields_cache is not None else it.fields()
        inst.upload(abi.F_PHI, phi)
        inst.upload(abi.F_T, T)
        insts.append(inst)
    return insts


def run_batched(ctx, items, steps, tile=16, threads=0, chunk=100, fields_cache=None, keep=False, single_launch=1):
    """All items in one group (merged worklist, or with single_launch=2 one
    launch per step over all their tiles); returns (stats, instances or None)."""
    insts = _create(ctx, items, tile, threads, fields_cache)
    ctx.sync()
    u0 = sum(i.info().cell_updates for i in insts)
    ctx.set_single_launch(single_launch)
    try:
        ctx.marker(0)
        t0 = time.perf_counter()
        ctx.step(insts, steps, graph_chunk=min(chunk, steps) if steps else 0)
        ctx.marker(1)
        ctx.sync()
    finally:
        ctx.set_single_launch(1)
    wall = time.perf_counter() - t0
    gpu_ms = ctx.elapsed_ms(0, 1)
    u1 = sum(i.info().cell_updates for i in insts)
    st = RunStats(wall, gpu_ms, steps, u1 - u0, steps * sum(it.cells for it in items))
    if keep:
        return st, insts
    for i in insts:
        i.destroy()
    return st, None


def step_plain(ctx, inst, steps):
    """The reference's loop shape: one step launched, then waited for, per time
    step (example_Dendritic_growth/kernel.cu:913-1011 synchronises after every
    kernel); no graph."""
    for _ in range(steps):
        inst.step(1)
        ctx.sync()


def run_sequential(ctx, items, steps, tile=16, threads=0, chunk=100, fields_cache=None, keep=False, plain=False):
    """One instance at a time (create, upload, step, next); stepping time only.
    plain=True steps like the reference's loop (step_plain) instead of replaying graphs."""
    wall = gpu = 0.0
    upd = 0
    kept = []
    for it in items:
        (inst,) = _create(ctx, [it], tile, threads, fields_cache)
        ctx.sync()
        u0 = inst.info().cell_updates
        ctx.marker(0)
        t0 = time.perf_counter()
        if plain:
            step_plain(ctx, inst, steps)
        else:
            inst.step(steps, graph_chunk=min(chunk, steps) if steps else 0)
        ctx.marker(1)
        ctx.sync()
        wall += time.perf_counter() - t0
        gpu += ctx.elapsed_ms(0, 1)
        upd += inst.info().cell_updates - u0
        if keep:
            kept.append(inst)
        else:
            inst.destroy()
    st = RunStats(wall, gpu, steps, upd, steps * sum(it.cells for it in items))
    return st, (kept if keep else None)


def autotune(ctx, items, trial_steps=200, warm_steps=20, tiles=TILES, threads=THREADS, chunk=100,
             fields_cache=None, log=None, modes=SINGLE_LAUNCH_MODES):
    """Time every (tile, threads, single-launch mode) candidate on the sweep's
    first steps (mode 1: merged worklist, a step kernel and the finalize per
    step; mode 2: one launch per step over every tile of the batch).

    The objective is wall time per step of the whole batch (the same physical
    simulation finishing sooner), not active cell-updates/s: a larger tile wakes
    more cells per interface and would score higher on the latter while doing
    more work. Threads never change results (bit-identical); the tile sets the
    activation granularity, so results at different tiles agree only to the
    DBA tolerance of the regular-mode solution.
    Threads and the mode never change results (bit-identical).
    Returns (best (tile, threads, mode), [(tile, threads, mode, ms_per_step, active_rate)]).
    """
    if fields_cache is None:
        fields_cache = {it.index: it.fields() for it in items}
    table = []
    try:
        for tile in tiles:
            if any(it.side % tile == 1 or it.side < tile for it in items):
                continue  # the library refuses a partial tile one cell wide
            for nt in threads:
                for mode in modes:
                    ctx.set_single_launch(mode)
                    insts = _create(ctx, items, tile, nt, fields_cache)
                    ctx.step(insts, warm_steps, graph_chunk=min(chunk, warm_steps))
                    ctx.sync()
                    u0 = sum(i.info().cell_updates for i in insts)
                    t0 = time.perf_counter()
                    ctx.step(insts, trial_steps, graph_chunk=min(chunk, trial_steps))
                    ctx.sync()
                    dt = time.perf_counter() - t0
                    u1 = sum(i.info().cell_updates for i in insts)
                    table.append((tile, nt, mode, dt * 1e3 / trial_steps, (u1 - u0) / dt))
                    if log:
                        log(f"autotune tile {tile} threads {nt} mode {mode}: {dt * 1e3 / trial_steps:.4f} ms/step")
                    for i in insts:
                        i.destroy()
    finally:
        ctx.set_single_launch(1)
    best = min(table, key=lambda r: r[3])
    return (best[0], best[1], best[2]), table


def footprint(ctx, items, tile=16):
    return sum(ctx.footprint(it.params(tile=tile)) for it in items)


def summary(items):
    sides = np.array([it.side for it in items])
    return {
        "instances": len(items),
        "cells": int((sides * sides).sum()),
        "side_min": int(sides.min()),
        "side_max": int(sides.max()),
        "j6": int(sum(1 for it in items if it.aniso_j == 6)),
        "seeds": int(sum(len(it.seeds) for it in items)),
    }
