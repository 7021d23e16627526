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
  batch     64 independent 2D Kobayashi dendrites, sides log-uniform in [256, 2048]
            (to the nearest cell), per-instance delta, j, K and 1-3 seeds
            (dba_amd.batch), 5000 steps, one merged-worklist group per step;
            tile and workgroup size from the autotuner. Reports the batched
            rate next to the plain one-instance-at-a-time loop (16x16 tiles).
  serve     Poisson stream of dendrite and reactive-CH solves (dba_amd.serve):
            sides log-uniform in [256, 4096], steps log-uniform in [1e3, 5e4],
            admitted by device footprint against a memory budget (--budget-gib),
            arrival rate steered toward 80% budget occupancy; reports p50/p99
            completion latency, the aggregate rate, and the plain sequential
            rate of a sample of the same requests run alone.
  modelb    the reference's ternary Model B, 256^3 periodic FP64, 4x4x4 blocks,
            DBA (every block stays awake); CPU baseline: the reference's own CPU
            program compiled from its sources.
  dendrite3d the reference's 3D Kobayashi-94 dendrite, 256^3 FP64, 4x4x4 blocks,
            DBA (about 2% of the field-set updates); CPU baseline as modelb.
  copycal   no DBA step: device-to-device copies of known size, for the traffic
            counters' calibration.
  hbmref    no DBA step: what a plain device copy (read + write) and sum (read)
            of 256 MB - 2 GB FP64 arrays move on this GPU, the yardstick for
            the step kernels' achieved bandwidth.
A step is one DBA time step of the whole instance (stencil kernel + finalize
kernel), replayed from a captured hipGraph. Inputs are resident in HBM before
the timed region. value = active cell-updates/s over the timed steps (device
counter of cells in awake tiles).

Before the timed region the graph it replays is captured, instantiated and
replayed once, whatever --warmup is, so the timed region only replays it. Then
a second instance of the same workload, from the same initial state, replays its
own chunk graph, untimed, for at least --settle seconds (0.2 by default): the
GPU's clocks take about that long to reach their steady state, and a few
warm-up steps of a few microseconds each do not get there (CH 1024^2 with
--steps 20 --warmup 5: 9.3 us per step without, 8.5 with). The measured
instance has then still taken exactly --warmup steps plus the one chunk, the
line reports the settling steps ("settle_steps"), and the timed region runs
exactly --steps steps of the measured instance.

The default workload's line also carries "companions": the same measurement of
dendrite, ch_cons, kh, modelb and dendrite3d at their own step counts, in the
same process after the default one (--no-companions skips them).

roofline: the step kernel's algorithmic bytes per launch (B_alg x active cells
of that launch: CH fused pass 16 B = read c + write c'; dendrite 16 B per (cell,
field) update; KH 64 B = read and write the 4 conserved fields) over its average
duration. One launch per step (single-launch steps): the GPU time per step of a
separate graph-replayed run of at least --timing-steps steps, between hipEvents
on the library stream. Two launches per step: hipEvents around every step
kernel in an eager pass.
cpu_baseline: the oracle's regular-mode loop on a bounded sample of the same
grid: serial, y-slab threads on the job's CPU share, and for CH the MPI slab
restatement (oracle/mpi_slab.c) on as many ranks; value is the faster all-core
leg, with the core counts and the CPU model reported beside it.
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
    "modelb": dict(model="modelb", n=256, tile=4, mode="dba", b_alg=48, warm=20, steps=200),
    "dendrite3d": dict(model="k94", n=256, tile=4, mode="dba", b_alg=32, warm=100, steps=2000),
    "batch": dict(model="k93", b_alg=32, warm=0, steps=5000),
    "serve": dict(model="mix", b_alg=None, warm=0, steps=0),
    "copycal": dict(model="ch", b_alg=None, warm=0, steps=0),
    "hbmref": dict(model=None, b_alg=None, warm=0, steps=0),
}


def make(workload, n=None, tile=None, threads=0):
    import numpy as np
    from dba_amd import abi, configs
    w = dict(WORKLOADS[workload])
    if n:
        w["n"] = n
    if tile:
        w["tile"] = tile
    mode = {"regular": abi.MODE_REGULAR, "dba": abi.MODE_DBA, "conserved": abi.MODE_CONSERVED}[w["mode"]]
    if w["model"] == "ch":
        p = configs.ch_params(n=w["n"], tile=w["tile"], mode=mode, threads=threads)
        fields = [configs.ch_ic(w["n"], w["n"])]
    elif w["model"] == "kh":
        p = configs.kh_params(n=w["n"], tile=w["tile"], mode=mode, threads=threads)
        fields = configs.kh_ic(w["n"])
    elif w["model"] == "modelb":
        n = w["n"]
        p = configs.modelb_params(n=n, mode=mode)
        fields = list(configs.modelb_ic(n, n, n)) + [np.zeros((n, n, n))]
    elif w["model"] == "k94":
        n = w["n"]
        p = configs.dendrite_k94_params(n=n, mode=mode)
        fields = list(configs.dendrite_k94_ic(n, n, n))
    else:
        p = configs.dendrite_k93_params(n=w["n"], tile=w["tile"], mode=mode, threads=threads)
        fields = list(configs.dendrite_k93_ic(w["n"], w["n"]))
    return p, fields


def graph_chunk_for(K, cap):
    """Steps per captured graph for a K-step timed region: the largest divisor of K
    not above cap, so the timed region is whole replays of one cached graph."""
    cap = max(1, min(cap, K)) if K > 0 else 1
    for c in range(cap, 0, -1):
        if K % c == 0:
            return c
    return 1


def reference_cpu_baseline(workload, budget_s=15.0):
    """The reference's own CPU_version step loop (compiled by oracle/Makefile.ref into
    oracle/_ref), single-threaded as the reference runs it, on the workload's grid;
    None when that build is absent."""
    import ctypes as C
    import numpy as np
    from dba_amd import configs
    ref = os.path.join(ROOT, "oracle", "_ref")
    w = WORKLOADS[workload]
    n = w["n"]
    D = C.POINTER(C.c_double)
    if w["model"] == "modelb":
        path = os.path.join(ref, "libref_modelb.so")
        if not os.path.exists(path):
            return None
        lib = C.CDLL(path)
        lib.ref_modelb_run.argtypes = [C.c_int, C.c_int, C.c_int, D, C.c_int, D, D, D, D]
        p = configs.modelb_params(n=n)
        prm = np.array([p.dx, p.dt, p.chi_xy, p.chi_xz, p.chi_yz, p.chi_xb, p.chi_yb, p.chi_zb, p.v_x, p.v_y, p.v_z,
                        p.r_x, p.r_y, p.r_z, p.n_exp, p.m_exp, p.epsx_sq, p.epsy_sq, p.epsz_sq, p.epsphi_sq,
                        p.mob_x, p.mob_y, p.mob_z, p.mob_phi, p.k_eq, p.z_crit, p.p_gel, p.k_0])
        x, y, z = (np.ascontiguousarray(a) for a in configs.modelb_ic(n, n, n))
        phi = np.zeros_like(x)
        ptr = lambda a: a.ctypes.data_as(D)

        def run(k):
            lib.ref_modelb_run(n, n, n, ptr(prm), k, ptr(x), ptr(y), ptr(z), ptr(phi))
        label = "example_Reactive_ModelB/CPU_version/base_CPU.cpp"
    elif w["model"] == "k94":
        path = os.path.join(ref, "libref_dendrite.so")
        if not os.path.exists(path):
            return None
        lib = C.CDLL(path)
        F = C.POINTER(C.c_float)
        lib.ref_dendrite_init.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, F, F]
        lib.ref_dendrite_run.argtypes = [C.c_int, C.c_int, C.c_int, F, C.c_int, F, F]
        q = configs.dendrite_k94_params(n=n)
        pf = np.zeros((n + 2) ** 3, dtype=np.float32)
        Tf = np.zeros_like(pf)
        lib.ref_dendrite_init(n, n, n, 0.25, pf.ctypes.data_as(F), Tf.ctypes.data_as(F))
        prm = np.array([q.dx, q.dt, q.delta, q.tau, q.eps_bar, q.alpha, q.gamma], dtype=np.float32)

        def run(k):
            lib.ref_dendrite_run(n, n, n, prm.ctypes.data_as(F), k, pf.ctypes.data_as(F), Tf.ctypes.data_as(F))
        label = "example_Dendritic_growth/CPU_version/Base.cpp (float32)"
    else:
        return None
    t0 = time.perf_counter()
    run(1)
    t1 = time.perf_counter() - t0
    steps = max(1, min(200, int(budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    run(steps)
    t = time.perf_counter() - t0
    return {"value": n ** 3 * steps / t, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
            "sample": f"{steps} regular-mode steps of the {n}^3 grid through the compiled reference CPU program "
                      f"({label}, g++ -O2), 1 thread ({cpu_model()})"}


def host_cores():
    """(cores used, facts): the CPU share this process may use. The affinity mask
    can list more CPUs than the job's share; OMP_NUM_THREADS, when set, states the
    share (the GPU box sets it), and the smaller of the two is used."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS", "")
    cores = min(aff, int(share)) if share.isdigit() and int(share) > 0 else aff
    facts = {"affinity_cpus": aff, "nproc": os.cpu_count(), "omp_num_threads": share or None,
             "cpu_model": cpu_model(), "cpu_quota": cgroup_cpu_quota()}
    return max(1, cores), facts


def cgroup_cpu_quota():
    """CPUs' worth of the cgroup v2 quota (cpu.max), or None without one."""
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        return None if q == "max" else int(q) / int(per)
    except (OSError, ValueError):
        return None


def mpirun_path():
    import shutil
    return shutil.which("mpirun") or ("/opt/conda/bin/mpirun" if os.path.exists("/opt/conda/bin/mpirun") else None)


def mpi_slab_baseline(p, field, ranks, steps):
    """The MPI slab restatement (oracle/mpi_slab.c: y-slabs, periodic MPI_Sendrecv
    ring of halo rows, the reference MPI_version's decomposition) on `ranks`
    processes; None when MPICH or the program is absent."""
    import subprocess
    import tempfile
    import numpy as np
    exe = os.path.join(ROOT, "oracle", "_build", "mpi_slab")
    mr = mpirun_path()
    if mr is None or not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "p.bin"), "wb") as f:
            f.write(bytes(p))
        np.ascontiguousarray(field, dtype=np.float64).tofile(os.path.join(d, "ic.bin"))
        r = subprocess.run([mr, "-np", str(ranks), exe, os.path.join(d, "p.bin"), os.path.join(d, "ic.bin"),
                            str(steps)], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(workload, budget_s=18.0, n=None):
    """Oracle regular-mode loop on the host cores, bounded sample: serial, y-slab
    threads and (CH) the MPI slab restatement on every core of the CPU share; the
    reference's own CPU program for the 3D reference workloads when it is built."""
    from dba_amd import abi
    from oracle.oracle import OracleSim
    if n is None:
        r = reference_cpu_baseline(workload)
        if r is not None:
            return r
    p, fields = make(workload, n=n)
    p.mode = abi.MODE_REGULAR
    ncell = p.nx * p.ny * p.nz
    cores, facts = host_cores()
    nth_all = cores if p.nz == 1 else 1  # the oracle steps the 3D models serially
    per = budget_s / 3
    res = {}
    for label, nth in (("serial", 1), ("threads", nth_all)):
        if label == "threads" and nth == 1:
            res[label] = res["serial"]
            continue
        o = OracleSim(p)
        for f, a in enumerate(fields):
            o.set_field(f, a)
        k0 = 1 if p.nz > 1 else 2
        t = o.bench(k0, nth)  # calibrate
        steps = max(1, min(2000, int(per / max(t / k0, 1e-6))))
        t = o.bench(steps, nth)
        res[label] = (ncell * steps / t, steps, nth)
        o.close()
    out = {
        "value": res["threads"][0], "unit": "cell-updates/s", "cores": res["threads"][2], "kind": "port",
        "sample": f"{res['threads'][1]} regular-mode steps of the {p.nx}x{p.ny}x{p.nz} grid, "
                  f"{res['threads'][2]} y-slab threads ({facts['cpu_model']})",
        "serial_value": res["serial"][0],
        "serial_sample": f"{res['serial'][1]} regular-mode steps, 1 thread",
        "threads_value": res["threads"][0],
        "host": facts,
    }
    if p.model == abi.MODEL_CH and p.mode == abi.MODE_REGULAR and nth_all > 1:
        steps = max(2, min(5000, int(per * res["threads"][0] / ncell)))
        m = mpi_slab_baseline(p, fields[0], nth_all, steps)
        if m is not None:
            out["mpi_value"] = m["cell_updates_per_s"]
            out["mpi_sample"] = (f"{m['steps']} regular-mode steps, {m['ranks']} MPI ranks in y-slabs "
                                 "(oracle/mpi_slab.c, MPICH)")
            if m["cell_updates_per_s"] > out["value"]:
                out["value"] = m["cell_updates_per_s"]
                out["cores"] = m["ranks"]
                out["sample"] = (f"{m['steps']} regular-mode steps of the {p.nx}x{p.ny} grid, {m['ranks']} MPI "
                                 f"ranks in y-slabs ({facts['cpu_model']})")
        else:
            out["mpi_value"] = None
            out["mpi_sample"] = "MPICH or oracle/_build/mpi_slab absent: y-slab threads only"
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def kernel_roofline(ctx, insts, b_alg, steps):
    """Eager pass with hipEvents around every step kernel: average duration and
    algorithmic GB/s (b_alg x active cells of the launch / duration)."""
    ctx.timing(True)
    u0 = sum(i.info().cell_updates for i in insts)
    ctx.step(insts, steps)
    ctx.sync()
    u1 = sum(i.info().cell_updates for i in insts)
    k_ms, k_n, f_ms, f_n = ctx.timing_query()
    ctx.timing(False)
    avg_k_s = k_ms / 1e3 / max(1, k_n)
    achieved = b_alg * (u1 - u0) / max(1, k_n) / avg_k_s / 1e9
    return {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": None,
        "kernel": "step kernel (fused stencil + detector)", "b_alg_per_cell": b_alg,
        "avg_kernel_us": avg_k_s * 1e6, "avg_finalize_us": f_ms * 1e3 / max(1, f_n), "launches_timed": k_n,
    }


def copy_calibration(n=4096, reps=8):
    """Device-to-device copies of known size (the D2D half of dba_upload, the
    runtime's blit kernel), for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    from dba_amd import Context, configs
    with Context(0) as ctx:
        inst = ctx.create(configs.ch_params(n=n, tile=32))
        c = configs.ch_ic(n, n)
        for _ in range(reps):
            inst.upload(0, c)
        inst.destroy()
    return {"workload": "copycal", "copies": reps, "bytes_per_copy": n * n * 8,
            "kernel": "__amd_rocclr_copyBuffer (D2D half of each upload)"}


def hbm_reference(sizes_mb=(256, 1024, 2048), reps=50):
    """What plain streaming moves on this GPU, as a yardstick for the step
    kernels' achieved bandwidth: PyTorch's device copy (each byte read once and
    written once, like a stencil step's 16 B per cell) and its sum (read only),
    of FP64 arrays of the given sizes, timed with hipEvents over `reps` launches
    after 5 untimed ones."""
    import torch
    dev = torch.device("cuda:0")
    out = {"workload": "hbmref", "unit": "TB/s", "copy": {}, "sum": {}}
    for mb in sizes_mb:
        n = mb * 1024 * 1024 // 8
        x = torch.rand(n, dtype=torch.float64, device=dev)
        y = torch.empty_like(x)
        for name, fn, nbytes in (("copy", lambda: y.copy_(x), 2 * n * 8), ("sum", lambda: x.sum(), n * 8)):
            for _ in range(5):
                fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            out[name][f"{mb}MB"] = nbytes / (e0.elapsed_time(e1) / reps / 1e3) / 1e12
        del x, y
    return out


def traffic_from_profile(workload):
    """HBM bytes per step-kernel launch measured by the rocprofv3 --pmc passes
    committed in profiles/ (FETCH_SIZE scaled by the same run's finite_kernel
    read of known size, plus WRITE_SIZE), or None when no such measurement is
    committed."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def bench_batch(args, K):
    """64-instance dendrite sweep: autotuned batched run vs the plain loop."""
    from dba_amd import Context, batch, configs
    from oracle.oracle import OracleSim
    items = batch.make_sweep(64)
    t0 = time.perf_counter()
    cache = {it.index: it.fields() for it in items}
    ic_s = time.perf_counter() - t0
    log = (lambda m: print(m, file=sys.stderr, flush=True))
    with Context(0) as ctx:
        (tile, nt, mode), table = batch.autotune(ctx, items, trial_steps=args.tune_steps, fields_cache=cache,
                                                 log=log)
        log(f"autotune: tile {tile}, threads {nt}, single-launch mode {mode}")
        st, insts = batch.run_batched(ctx, items, K, tile=tile, threads=nt, chunk=args.chunk, fields_cache=cache,
                                      keep=True, single_launch=mode)
        ctx.set_single_launch(mode)
        log(f"batched: {st.wall_s:.3f} s, {st.active_rate:.4g} active/s")
        roof = kernel_roofline(ctx, insts, 32, args.timing_steps)
        ctx.set_single_launch(1)
        finite = all(i.check_finite() for i in insts)
        for i in insts:
            i.destroy()
        st16, _ = batch.run_batched(ctx, items, K, tile=16, threads=0, chunk=args.chunk, fields_cache=cache)
        log(f"batched 16x16: {st16.wall_s:.3f} s")
        if args.no_seq:
            seq = plain = None
        else:
            seq, _ = batch.run_sequential(ctx, items, K, tile=16, threads=0, chunk=args.chunk, fields_cache=cache)
            log(f"sequential: {seq.wall_s:.3f} s, {seq.active_rate:.4g} active/s")
            plain, _ = batch.run_sequential(ctx, items, K, tile=16, threads=0, fields_cache=cache, plain=True)
            log(f"sequential plain: {plain.wall_s:.3f} s, {plain.active_rate:.4g} active/s")
    out = {
        "metric": "active cell-updates/s of the batched 64-instance dendrite sweep (+ vs plain sequential DBA)",
        "value": st.active_rate, "unit": "cell-updates/s", "n_gpus": 1, "steps": K, "warmup": 0,
        "ms_per_step": st.wall_s * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batch: 64 Kobayashi-93 dendrites, sides log-U[256,2048], 5000 steps",
                   "tile": tile, "threads": nt, "single_launch": mode, "graph_chunk": args.chunk,
                   **batch.summary(items)},
        "full_grid_updates_per_s": st.full_rate,
        "active_fraction": st.active_updates / st.full_updates,
        "batched_vs_sequential_plain": st.active_rate / plain.active_rate if plain else None,
        "batched_vs_sequential_graph": st.active_rate / seq.active_rate if seq else None,
        "sequential_plain": None if plain is None else {
            "value": plain.active_rate, "wall_s": plain.wall_s, "tile": 16, "threads": "default",
            "note": "one instance at a time, one step launched and waited for per time step "
                    "(the reference loop's shape), stepping time only"},
        "sequential_graph": None if seq is None else {
            "value": seq.active_rate, "wall_s": seq.wall_s, "tile": 16, "threads": "default",
            "note": "one instance at a time, graph-replayed chunks of steps, stepping time only"},
        "batched_16x16": {"value": st16.active_rate, "wall_s": st16.wall_s,
                          "vs_sequential_wall": seq.wall_s / st16.wall_s if seq else None},
        "autotune": [{"tile": a, "threads": b, "single_launch": m, "ms_per_step": c, "active_rate": d}
                     for a, b, m, c, d in table],
        "ic_host_s": ic_s, "finite": bool(finite), "roofline": roof,
    }
    if not args.no_cpu:
        # the oracle's regular-mode loop on one 512^2 dendrite of the sweep's parameters
        it = min(items, key=lambda x: abs(x.side - 512))
        p = it.params(tile=16)
        p.mode = 0
        o = OracleSim(p)
        phi, T = cache[it.index]
        o.set_field(0, phi)
        o.set_field(1, T)
        cores, facts = host_cores()
        tc = o.bench(2, cores)
        steps = max(2, min(2000, int(10.0 / max(tc / 2, 1e-6))))
        tc = o.bench(steps, cores)
        o.close()
        out["cpu_baseline"] = {"value": it.cells * steps / tc, "unit": "cell-updates/s", "cores": cores,
                               "kind": "port", "sample": f"{steps} regular-mode steps of instance {it.index} "
                               f"({it.side}^2), {cores} y-slab threads ({cpu_model()})", "host": facts}
    return out


def bench_serve(args):
    """Continuous-batching server on a Poisson stream; sequential sample alone."""
    from dba_amd import Context, batch, serve
    log = (lambda m: print(m, file=sys.stderr, flush=True))
    rs = serve.make_requests(args.requests)
    tiles = {"dendrite": 16, "ch": 32}
    budget = int(args.budget_gib * 2 ** 30)
    with Context(0) as ctx:
        free, total = ctx.device_memory()
        be = serve.LibraryBackend(ctx)
        # fully active CH batches take one launch per step; sparse dendrite batches the merged worklist
        modes = {"ch": 2, "dendrite": 1} if args.serve_modes == "split" else None
        srv = serve.Server(be, budget, chunk=args.chunk, tiles=tiles, target_occupancy=0.8, gain=2.0, log=log,
                           modes=modes)
        done = srv.run(rs, rate=args.rate)
        rep = serve.report(done, srv.wall, srv.occupancy, busy_s=srv.busy_s)
        log(f"served {rep['completed']} in {srv.wall:.1f} s, p50 {rep['p50_latency_s']:.3f} s, "
            f"p99 {rep['p99_latency_s']:.3f} s, {rep['active_rate']:.4g} active/s, "
            f"device busy {rep['device_busy_fraction']:.3f}")
        # sequential loops: requests of the whole stream alone, in van der Corput
        # order (any prefix spans the stream), until the time budget
        seq = {}
        order = serve.spread_order(len(rs))
        for mode in ("graph", "plain"):
            seq_t = seq_u = 0.0
            nseq = 0
            for r in (rs[i] for i in order):
                if seq_t > args.seq_budget_s:
                    break
                inst = ctx.create(r.params(tiles, {}))
                for f, a in enumerate(r.fields()):
                    inst.upload(f, a)
                ctx.sync()
                u0 = inst.info().cell_updates
                t0 = time.perf_counter()
                if mode == "graph":
                    inst.step(r.steps, graph_chunk=args.chunk)
                else:
                    batch.step_plain(ctx, inst, r.steps)
                ctx.sync()
                seq_t += time.perf_counter() - t0
                seq_u += inst.info().cell_updates - u0
                nseq += 1
                inst.destroy()
            seq[mode] = (seq_u / seq_t, nseq, seq_t)
            log(f"sequential {mode} sample: {nseq} requests, {seq[mode][0]:.4g} active/s")
    lat = sorted(c.latency for c in done)
    out = {
        "metric": "aggregate active cell-updates/s of the served dendrite + CH stream (+ p50/p99 latency)",
        "value": rep["active_rate"], "unit": "cell-updates/s", "n_gpus": 1, "steps": 0, "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"serve: {args.requests} requests, 50% Kobayashi-93 dendrite / 50% reactive CH, "
                               "sides log-U[256,4096], steps log-U[1e3,5e4]",
                   "budget_bytes": budget, "device_total_bytes": total, "graph_chunk": args.chunk,
                   "initial_rate_per_s": args.rate, "target_occupancy": 0.8,
                   "single_launch_modes": modes or "library default, one call"},
        **rep,
        "latency_s": {"p50": rep["p50_latency_s"], "p90": serve.percentile(lat, 90), "p99": rep["p99_latency_s"],
                      "max": lat[-1]},
        "served_vs_sequential_plain": rep["active_rate"] / seq["plain"][0],
        "served_vs_sequential_graph": rep["active_rate"] / seq["graph"][0],
        "sequential_plain": {"value": seq["plain"][0], "requests": seq["plain"][1], "wall_s": seq["plain"][2],
                             "note": "requests of the same stream run alone, in van der Corput order over the "
                                     "whole stream until the time budget, one step launched and waited for per "
                                     "time step (the reference loop's shape)"},
        "sequential_graph": {"value": seq["graph"][0], "requests": seq["graph"][1], "wall_s": seq["graph"][2],
                             "note": "the same sample alone, graph-replayed chunks of steps"},
        "host_wait_s": srv.sync_s,
        "ticks": srv.ticks, "max_used_bytes": srv.max_used,
        "final_rate_per_s": srv.rate_trace[-1][1] if srv.rate_trace else args.rate,
    }
    return out


# The other single-instance configurations, measured after the default one in the
# same process (their own step counts, no CPU baseline): the dendrite of the
# plain-DBA configuration (sleeping tiles), conserved CH, KH, and the reference's
# own 3D models.
COMPANIONS = ("dendrite", "ch_cons", "kh", "modelb", "dendrite3d")


def companion(o):
    """The fields of a companion measurement kept in the default line."""
    r = o["roofline"]
    return {"workload": o["config"]["workload"], "value": o["value"], "unit": o["unit"], "steps": o["steps"],
            "warmup": o["warmup"], "settle_steps": o.get("settle_steps"), "ms_per_step": o["ms_per_step"],
            "gpu_ms_per_step": o["gpu_ms_per_step"],
            "active_fraction": o["active_fraction"], "full_grid_updates_per_s": o["full_grid_updates_per_s"],
            "finite": o["finite"],
            "roofline": {k: r[k] for k in ("achieved", "frac", "traffic", "b_alg_per_cell", "avg_kernel_us", "kernel")}}


def measure(ctx, workload, K, W, chunk_cap, timing_steps, single_launch, n=None, tile=None, threads=0,
            settle=0.2):
    """One instance of `workload`: W warm-up steps, then K timed steps replayed
    from a captured graph; returns the JSON line without cpu_baseline."""
    w = WORKLOADS[workload]
    p, fields = make(workload, n=n, tile=tile, threads=threads)
    ncell = p.nx * p.ny * p.nz

    chunk = graph_chunk_for(K, chunk_cap)
    ctx.set_single_launch(single_launch)
    inst = ctx.create(p)
    for f, a in enumerate(fields):
        inst.upload(f, a)
    inst.step(W)
    # Capture, instantiate and replay once the graph the timed region replays,
    # whatever W is, so the timed region launches a cached, warm graph only.
    inst.step(chunk, graph_chunk=chunk)
    ctx.sync()
    # clock settling on a second instance of the same workload, from the same
    # initial state, so the measured instance keeps the state of its own warm-up
    settle_steps = 0
    warm = None
    if settle > 0:
        warm = ctx.create(p)
        for f, a in enumerate(fields):
            warm.upload(f, a)
        ts = time.perf_counter()
        while time.perf_counter() - ts < settle:
            warm.step(chunk, graph_chunk=chunk)
            ctx.sync()
            settle_steps += chunk
    info0 = inst.info()
    u0 = info0.cell_updates
    single = bool(info0.single_launch)
    ctx.marker(0)
    t0 = time.perf_counter()
    inst.step(K, graph_chunk=chunk)
    ctx.marker(1)
    ctx.sync()
    t1 = time.perf_counter()
    u1 = inst.info().cell_updates
    if warm is not None:
        warm.destroy()
    gpu_ms = ctx.elapsed_ms(0, 1)
    wall_s = t1 - t0
    active_updates = u1 - u0
    value = active_updates / wall_s
    active_frac = active_updates / (K * ncell)

    # kernel time of a long graph-replayed run (the host enqueues every replay
    # before the first ends, so no launch latency sits between the two events)
    reps = max(1, -(-timing_steps // chunk))
    u4 = inst.info().cell_updates
    ctx.marker(2)
    inst.step(reps * chunk, graph_chunk=chunk)
    ctx.marker(3)
    ctx.sync()
    u5 = inst.info().cell_updates
    replay_ms = ctx.elapsed_ms(2, 3)
    replay_steps = reps * chunk

    # eager pass with hipEvents around every step kernel
    TS = timing_steps
    ctx.timing(True)
    u2 = inst.info().cell_updates
    inst.step(TS)
    ctx.sync()
    u3 = inst.info().cell_updates
    k_ms, k_n, f_ms, f_n = ctx.timing_query()
    ctx.timing(False)
    finite = inst.check_finite()
    inst.destroy()

    eager_k_s = k_ms / 1e3 / max(1, k_n)
    if single:
        # one launch per step: the kernel's duration is the GPU time per step of
        # a long graph-replayed run (hipEvents at its two ends; it also holds one
        # state-restoring launch per graph chunk and the gaps between launches)
        avg_k_s = replay_ms / 1e3 / replay_steps
        cells_per_launch = (u5 - u4) / replay_steps
        kernel = "single-launch step kernel (awake sets + fused stencil + detector)"
    else:
        # two launches per step: hipEvents around every step kernel in an eager pass
        avg_k_s = eager_k_s
        cells_per_launch = (u3 - u2) / max(1, k_n)
        kernel = "step kernel (fused stencil + detector)"
    bytes_per_launch = w["b_alg"] * cells_per_launch
    achieved = bytes_per_launch / avg_k_s / 1e9
    out = {
        "metric": "active cell-updates/s (+ % HBM peak of the step kernel)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "settle_steps": settle_steps,
        "ms_per_step": wall_s * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{workload}: {p.nx}x{p.ny}" + (f"x{p.nz}" if p.nz > 1 else "") +
                        f" FP64, {p.tile_x}x{p.tile_y}" + (f"x{p.tile_z}" if p.nz > 1 else "") + f" tiles, {w['mode']}",
            "grid": [p.nx, p.ny, p.nz], "tile": p.tile_x, "threads": p.threads or "default", "mode": w["mode"],
            "graph_chunk": chunk,
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
            "traffic": None if (n or tile or threads) else traffic_from_profile(workload),
            "kernel": kernel,
            "b_alg_per_cell": w["b_alg"],
            "avg_kernel_us": avg_k_s * 1e6,
            "eager_kernel_us": eager_k_s * 1e6,
            "avg_finalize_us": f_ms * 1e3 / max(1, f_n),
            "launches_timed": K if single else k_n,
        },
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="ch", choices=sorted(WORKLOADS))
    ap.add_argument("--chunk", type=int, default=100, help="steps per captured hipGraph")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-companions", action="store_true",
                    help="default workload: skip the other single-instance configurations")
    ap.add_argument("--timing-steps", type=int, default=500)
    ap.add_argument("--settle", type=float, default=0.2,
                    help="untimed replays of the timed graph before the timed region, in seconds")
    ap.add_argument("--n", type=int, default=None, help="grid side (ch, ch_cons, dendrite, kh); default: the workload's")
    ap.add_argument("--tile", type=int, default=None, help="activation tile edge override")
    ap.add_argument("--threads", type=int, default=0, help="workgroup size override (64/128/256)")
    ap.add_argument("--tune-steps", type=int, default=200, help="batch: timed steps per autotuner candidate")
    ap.add_argument("--serve-modes", default="split", choices=("split", "one"),
                    help="serve: step CH (single-launch mode 2) and dendrites (mode 1) apart, or all in one call")
    ap.add_argument("--single-launch", type=int, default=1, choices=(0, 1),
                    help="one-instance workloads: 0 forces the worklist path (step kernel + finalize)")
    ap.add_argument("--requests", type=int, default=300, help="serve: requests in the stream")
    ap.add_argument("--rate", type=float, default=4.0, help="serve: initial arrival rate (1/s)")
    ap.add_argument("--budget-gib", type=float, default=4.0, help="serve: device memory budget for admission")
    ap.add_argument("--no-seq", action="store_true", help="batch: skip the one-instance-at-a-time loops")
    ap.add_argument("--seq-budget-s", type=float, default=30.0, help="serve: time for the sequential sample")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(one context per GPU; independent instances would shard across GPUs with no exchange)",
              file=sys.stderr)
        sys.exit(2)

    from dba_amd import Context
    w = WORKLOADS[args.workload]
    K = args.steps if args.steps is not None else w["steps"]
    W = args.warmup if args.warmup is not None else w["warm"]
    if args.workload == "batch":
        print(json.dumps(bench_batch(args, K)))
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(args)))
        return
    if args.workload == "hbmref":
        print(json.dumps(hbm_reference()))
        return
    if args.workload == "copycal":
        print(json.dumps(copy_calibration()))
        return
    with Context(0) as ctx:
        out = measure(ctx, args.workload, K, W, args.chunk, args.timing_steps, args.single_launch,
                      n=args.n, tile=args.tile, threads=args.threads, settle=args.settle)
        if args.workload == "ch" and not (args.n or args.tile or args.threads or args.no_companions):
            out["companions"] = [companion(measure(ctx, c, WORKLOADS[c]["steps"], WORKLOADS[c]["warm"], args.chunk,
                                                   args.timing_steps, args.single_launch, settle=args.settle))
                                 for c in COMPANIONS]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.workload, n=args.n)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
