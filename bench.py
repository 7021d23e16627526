#!/usr/bin/env python3
"""Benchmark of the MI355X DBA time step (one JSON line on stdout).

Workloads (--workload):
  ch        (default) reactive Cahn-Hilliard, 1024^2 periodic FP64, 32x32 tiles,
            plain DBA, 9-point Laplacian, M=1, kappa=1, dx=1, dt=0.01,
            Gamma=1e-3, cbar=0, c = U[-0.05, 0.05] (seed 1234).
  ch_cons   same, conserved DBA.
  dendrite  2D Kobayashi dendrite 512^2 FP64, 16x16 tiles, DBA.
  kh        compressible Euler Kelvin-Helmholtz 2048^2 FP64, gamma 1.4, CFL 0.4,
            reflective y walls, 16x16 tiles, density-gradient wake-up (eps 0.1).
A step is one DBA time step of the whole instance (stencil kernel + finalize
kernel), replayed from a captured hipGraph. Inputs are resident in HBM before
the timed region. value = active cell-updates/s over the timed steps (device
counter of cells in awake tiles).

roofline: the step kernel's algorithmic bytes per launch (B_alg x active cells
of that launch: CH fused pass 16 B = read c + write c'; dendrite 32 B = read and
write phi and T; KH 64 B = read and write the 4 conserved fields) over its average duration, timed with hipEvents on the
library stream in a second, eager pass of the same workload.
cpu_baseline: the oracle's regular-mode loop (serial and y-slab threads on the
host cores) on a bounded sample of the same grid.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "ch": dict(model="ch", n=1024, tile=32, mode="dba", b_alg=16, warm=500, steps=5000),
    "ch_cons": dict(model="ch", n=1024, tile=32, mode="conserved", b_alg=16, warm=500, steps=5000),
    "dendrite": dict(model="k93", n=512, tile=16, mode="dba", b_alg=32, warm=200, steps=4000),
    "kh": dict(model="kh", n=2048, tile=16, mode="dba", b_alg=64, warm=100, steps=1000),
}


def make(workload):
    from dba_amd import abi, configs
    w = WORKLOADS[workload]
    mode = {"regular": abi.MODE_REGULAR, "dba": abi.MODE_DBA, "conserved": abi.MODE_CONSERVED}[w["mode"]]
    if w["model"] == "ch":
        p = configs.ch_params(n=w["n"], tile=w["tile"], mode=mode)
        fields = [configs.ch_ic(w["n"], w["n"])]
    elif w["model"] == "kh":
        p = configs.kh_params(n=w["n"], tile=w["tile"], mode=mode)
        fields = configs.kh_ic(w["n"])
    else:
        p = configs.dendrite_k93_params(n=w["n"], tile=w["tile"], mode=mode)
        fields = list(configs.dendrite_k93_ic(w["n"], w["n"]))
    return p, fields


def cpu_baseline(workload, budget_s=12.0):
    """Oracle regular-mode loop on the host cores, bounded sample."""
    from dba_amd import abi
    from oracle.oracle import OracleSim
    p, fields = make(workload)
    p.mode = abi.MODE_REGULAR
    ncell = p.nx * p.ny
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(cores, 16))
    res = {}
    for label, nth in (("serial", 1), ("threads", cores)):
        o = OracleSim(p)
        for f, a in enumerate(fields):
            o.set_field(f, a)
        t = o.bench(2, nth)  # calibrate
        steps = max(2, int(budget_s / 2 / max(t / 2, 1e-6)))
        steps = min(steps, 2000)
        t = o.bench(steps, nth)
        res[label] = (ncell * steps / t, steps, nth)
        o.close()
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    v, steps, nth = res["threads"]
    return {
        "value": v, "unit": "cell-updates/s", "cores": nth, "kind": "port",
        "sample": f"{steps} regular-mode steps of the {p.nx}x{p.ny} grid, {nth} y-slab threads ({cpu_model})",
        "serial_value": res["serial"][0],
        "serial_sample": f"{res['serial'][1]} regular-mode steps, 1 thread",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="ch", choices=sorted(WORKLOADS))
    ap.add_argument("--chunk", type=int, default=100, help="steps per captured hipGraph")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=500)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported in this build "
              "(instances shard across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from dba_amd import Context
    w = WORKLOADS[args.workload]
    K = args.steps if args.steps is not None else w["steps"]
    W = args.warmup if args.warmup is not None else w["warm"]
    p, fields = make(args.workload)
    ncell = p.nx * p.ny

    with Context(0) as ctx:
        inst = ctx.create(p)
        for f, a in enumerate(fields):
            inst.upload(f, a)
        chunk = max(1, min(args.chunk, K))
        inst.step(W, graph_chunk=chunk)
        ctx.sync()
        u0 = inst.info().cell_updates
        ctx.marker(0)
        t0 = time.perf_counter()
        inst.step(K, graph_chunk=chunk)
        ctx.marker(1)
        ctx.sync()
        t1 = time.perf_counter()
        u1 = inst.info().cell_updates
        gpu_ms = ctx.elapsed_ms(0, 1)
        wall_s = t1 - t0
        active_updates = u1 - u0
        value = active_updates / wall_s
        active_frac = active_updates / (K * ncell)

        # kernel timing pass (eager, hipEvents around every step kernel)
        TS = args.timing_steps
        ctx.timing(True)
        u2 = inst.info().cell_updates
        inst.step(TS)
        ctx.sync()
        u3 = inst.info().cell_updates
        k_ms, k_n, f_ms, f_n = ctx.timing_query()
        ctx.timing(False)
        finite = inst.check_finite()
        inst.destroy()

    avg_k_s = k_ms / 1e3 / max(1, k_n)
    cells_per_launch = (u3 - u2) / max(1, k_n)
    bytes_per_launch = w["b_alg"] * cells_per_launch
    achieved = bytes_per_launch / avg_k_s / 1e9
    out = {
        "metric": "active cell-updates/s (+ % HBM peak of the step kernel)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall_s * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {p.nx}x{p.ny} FP64, {w['tile']}x{w['tile']} tiles, {w['mode']}",
            "grid": [p.nx, p.ny], "tile": w["tile"], "mode": w["mode"], "graph_chunk": chunk,
        },
        "full_grid_updates_per_s": K * ncell / wall_s,
        "active_fraction": active_frac,
        "gpu_ms_per_step": gpu_ms / K,
        "finite": bool(finite),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "step kernel (fused stencil + detector)",
            "b_alg_per_cell": w["b_alg"],
            "avg_kernel_us": avg_k_s * 1e6,
            "avg_finalize_us": f_ms * 1e3 / max(1, f_n),
            "launches_timed": k_n,
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
