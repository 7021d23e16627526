"""GPU: single-launch steps (one launch per step, awake sets derived in every
tile's workgroup, canonicalising kernels per call or graph chunk) against the
worklist path (step kernel + finalize) on the same instances, alone and in
batches: fields, awake sets,
step and cell-update counters, history, simulated time and the rebuilt worklist
are identical. The worklist path is pinned to the oracle by tests/test_gpu_ch.py
(which itself runs through the single-launch path for eligible instances)."""
import numpy as np
import pytest

from dba_amd import abi, configs

pytestmark = pytest.mark.gpu


def ch_case(n, tile, mode, stencil, bc):
    p = configs.ch_params(n=n, tile=tile, mode=mode, stencil=stencil, eps=2e-3, history_cap=400)
    if bc == "noflux":
        p.bc_x = p.bc_y = abi.BC_NOFLUX
    c = configs.ch_ic(n, n, cbar=0.0, amp=0.05)
    # a quiet quarter: its tiles sleep, and wake as the interface reaches them
    c[: n // 2, : n // 2] = -0.6
    return p, c


def run(ctx, p, c, single, plan):
    ctx.set_single_launch(single)
    inst = ctx.create(p)
    inst.upload(0, c)
    masks = []
    for steps, chunk in plan:
        inst.step(steps, graph_chunk=chunk)
        masks.append(inst.active(0)[1])
    info = inst.info()
    out = dict(c=inst.download(0), masks=masks, steps=info.steps_done, updates=info.cell_updates, time=info.sim_time,
               hist=inst.history(), work=inst.worklist(), shift=info.cons_offset, single=info.single_launch)
    inst.destroy()
    ctx.set_single_launch(True)
    return out


@pytest.mark.parametrize("tile,stencil,bc,mode", [
    (32, 9, "periodic", abi.MODE_DBA),
    (16, 5, "periodic", abi.MODE_DBA),
    (8, 9, "noflux", abi.MODE_DBA),
    (16, 9, "periodic", abi.MODE_REGULAR),
])
def test_single_launch_equals_worklist(ctx, tile, stencil, bc, mode):
    n = 16 * tile   # the quiet 8x8-tile quarter has tiles two away from the noise, which sleep
    p, c = ch_case(n, tile, mode, stencil, bc)
    plan = [(1, 0), (7, 0), (100, 25), (37, 10), (3, 0)]
    a = run(ctx, p, c, False, plan)
    b = run(ctx, p, c, True, plan)
    assert np.array_equal(a["c"], b["c"])
    for x, y in zip(a["masks"], b["masks"]):
        assert np.array_equal(x, y)
    assert a["steps"] == b["steps"] == sum(s for s, _ in plan)
    assert a["updates"] == b["updates"]
    assert a["time"] == b["time"]
    assert np.array_equal(a["hist"], b["hist"])
    assert np.array_equal(a["work"], b["work"])
    if mode == abi.MODE_DBA:
        # every tile ran the first step; some sleep from the second on
        assert a["masks"][0].sum() < a["masks"][0].size


def k93_run(ctx, p, phi, T, single, plan):
    ctx.set_single_launch(single)
    inst = ctx.create(p)
    inst.upload(abi.F_PHI, phi)
    inst.upload(abi.F_T, T)
    masks = []
    for steps, chunk in plan:
        inst.step(steps, graph_chunk=chunk)
        masks.append((inst.active(0)[1], inst.active(1)[1]))
    info = inst.info()
    out = dict(phi=inst.download(0), T=inst.download(1), masks=masks, steps=info.steps_done,
               updates=info.cell_updates, time=info.sim_time, hist=inst.history(), work=inst.worklist(),
               single=info.single_launch)
    inst.destroy()
    ctx.set_single_launch(True)
    return out


# 16 x 16 tiles at the default 512 threads take the speculative staging (both
# buffers' rings loaded before the neighbour table); 8 and 32 do not
@pytest.mark.parametrize("tile,mode,bc", [(16, abi.MODE_DBA, abi.BC_NOFLUX), (8, abi.MODE_DBA, abi.BC_NOFLUX),
                                          (32, abi.MODE_DBA, abi.BC_NOFLUX), (16, abi.MODE_REGULAR, abi.BC_NOFLUX),
                                          (16, abi.MODE_DBA, abi.BC_PERIODIC)])
def test_k93_single_launch_equals_worklist(ctx, tile, mode, bc):
    n = 256
    p = configs.dendrite_k93_params(n=n, tile=tile, mode=mode, history_cap=400, bc=bc)
    phi, T = configs.dendrite_k93_ic(n, n, seeds=[(90, 100), (170, 150)])
    plan = [(1, 0), (9, 0), (150, 50), (33, 10), (2, 0)]
    a = k93_run(ctx, p, phi, T, False, plan)
    b = k93_run(ctx, p, phi, T, True, plan)
    assert a["single"] == 0 and b["single"] == 1
    assert np.array_equal(a["phi"], b["phi"]) and np.array_equal(a["T"], b["T"])
    for (a0, a1), (b0, b1) in zip(a["masks"], b["masks"]):
        assert np.array_equal(a0, b0) and np.array_equal(a1, b1)
    assert a["steps"] == b["steps"] and a["updates"] == b["updates"] and a["time"] == b["time"]
    assert np.array_equal(a["hist"], b["hist"])
    assert np.array_equal(a["work"], b["work"])
    if mode == abi.MODE_DBA:
        p_awake, t_awake = a["masks"][-1]
        assert 0 < p_awake.sum() < p_awake.size and 0 < t_awake.sum() < t_awake.size
        assert not np.array_equal(p_awake, t_awake)   # the two sets differ


@pytest.mark.parametrize("n,tile,stencil", [(1024, 16, 9), (2048, 32, 5)])
def test_multi_segment_single_launch_equals_worklist(ctx, n, tile, stencil):
    """4096 tiles, four 1024-tile segments: single-launch steps through the tile
    map, one canonicalising workgroup per segment and the per-instance merge."""
    p, c = ch_case(n, tile, abi.MODE_DBA, stencil, "periodic")
    plan = [(1, 0), (6, 0), (40, 20), (3, 0)]
    a = run(ctx, p, c, False, plan)
    b = run(ctx, p, c, True, plan)
    assert np.array_equal(a["c"], b["c"])
    for x, y in zip(a["masks"], b["masks"]):
        assert np.array_equal(x, y)
    assert a["steps"] == b["steps"] and a["updates"] == b["updates"] and a["time"] == b["time"]
    assert np.array_equal(a["hist"], b["hist"])
    assert np.array_equal(a["work"], b["work"])
    assert a["masks"][-1].sum() < a["masks"][-1].size


def batch_cases():
    """Uneven K93 dendrites (one over 1024 tiles) and CH grids in DBA and regular mode."""
    out = []
    for n, seeds, mode in ((128, [(40, 60)], abi.MODE_DBA), (544, [(100, 300), (400, 120)], abi.MODE_DBA),
                           (96, [(48, 48)], abi.MODE_REGULAR)):
        p = configs.dendrite_k93_params(n=n, tile=16, mode=mode, history_cap=200)
        phi, T = configs.dendrite_k93_ic(n, n, seeds=seeds)
        out.append(("k93", p, [phi, T]))
    for n, mode in ((256, abi.MODE_DBA), (160, abi.MODE_REGULAR)):
        p, c = ch_case(n, 16, mode, 9, "periodic")
        p.history_cap = 200
        out.append(("ch", p, [c]))
    return out


def batch_run(ctx, cases, single, plan):
    ctx.set_single_launch(2 if single else 0)
    insts = []
    for _, p, fields in cases:
        inst = ctx.create(p)
        for f, a in enumerate(fields):
            inst.upload(f, a)
        insts.append(inst)
    groups = [[i for i, (m, _, _) in zip(insts, cases) if m == kind] for kind in ("k93", "ch")]
    groups = [g for g in groups if g]
    masks = []
    for steps, chunk in plan:
        for g in groups:
            ctx.step(g, steps, chunk)
        masks.append([[inst.active(s)[1] for s in range(2 if kind == "k93" else 1)]
                      for inst, (kind, _, _) in zip(insts, cases)])
    out = []
    for inst, (_, _, fields) in zip(insts, cases):
        info = inst.info()
        out.append(dict(f=[inst.download(f) for f in range(len(fields))], steps=info.steps_done,
                        updates=info.cell_updates, time=info.sim_time, hist=inst.history(),
                        work=inst.worklist(), single=info.single_launch))
        inst.destroy()
    ctx.set_single_launch(True)
    return out, masks


def test_batched_single_launch_equals_worklist(ctx):
    """Batches of uneven instances (mixed DBA / regular modes, one instance over
    1024 tiles) in single-launch steps through the tile map give the worklist
    path's fields, awake sets, counters, times, history and rebuilt worklists."""
    cases = batch_cases()
    plan = [(1, 0), (5, 0), (60, 20), (2, 0)]
    a, ma = batch_run(ctx, cases, False, plan)
    b, mb = batch_run(ctx, cases, True, plan)
    for x, y in zip(a, b):
        assert x["single"] == 0 and y["single"] == 1
        for u, v in zip(x["f"], y["f"]):
            assert np.array_equal(u, v)
        assert x["steps"] == y["steps"] and x["updates"] == y["updates"] and x["time"] == y["time"]
        assert np.array_equal(x["hist"], y["hist"])
        assert np.array_equal(x["work"], y["work"])
    for sa, sb in zip(ma, mb):
        for ia, ib in zip(sa, sb):
            for u, v in zip(ia, ib):
                assert np.array_equal(u, v)
    # the large dendrite has sleeping tiles at the end
    p_awake = ma[-1][1][0]
    assert 0 < p_awake.sum() < p_awake.size


def test_single_launch_modes(ctx):
    """Mode 1 leaves batches on the worklist path, mode 2 takes them too, other
    modes are rejected; the steps agree in every mode."""
    cases = batch_cases()[3:]
    plan = [(3, 0), (20, 10)]
    ref, mref = batch_run(ctx, cases, False, plan)
    ctx.set_single_launch(1)
    insts = []
    for _, p, fields in cases:
        inst = ctx.create(p)
        inst.upload(0, fields[0])
        insts.append(inst)
    for steps, chunk in plan:
        ctx.step(insts, steps, chunk)
    for inst, r in zip(insts, ref):
        assert np.array_equal(inst.download(0), r["f"][0])
        inst.destroy()
    for bad in (-1, 3):
        with pytest.raises(abi.DbaError):
            ctx.set_single_launch(bad)
    ctx.set_single_launch(True)


@pytest.mark.parametrize("n,tile", [(128, 16), (512, 16), (1024, 16)])
def test_conserved_single_launch_equals_worklist(ctx, n, tile):
    """Conserved DBA: one CH instance of <= 1024 tiles takes single-launch steps
    (each workgroup recomputes the previous step's mean shift in the finalize's
    order); 1024^2 at 16x16 (4096 tiles) stays on the worklist path. Either way
    fields, awake sets, counters, history, time and the cumulative shift are
    bit-identical to the worklist path."""
    p, c = ch_case(n, tile, abi.MODE_CONSERVED, 9, "periodic")
    plan = [(1, 0), (4, 0), (30, 10), (3, 0)]
    a = run(ctx, p, c, False, plan)
    b = run(ctx, p, c, True, plan)
    assert np.array_equal(a["c"], b["c"])
    for x, y in zip(a["masks"], b["masks"]):
        assert np.array_equal(x, y)
    assert a["steps"] == b["steps"] and a["updates"] == b["updates"] and a["time"] == b["time"]
    assert a["shift"] == b["shift"]
    assert np.array_equal(a["hist"], b["hist"])
    assert np.array_equal(a["work"], b["work"])
    assert b["single"] == (1 if (n // tile) ** 2 <= 1024 else 0)
    # the mean is conserved (to the rounding of the sums)
    assert abs(a["c"].sum() - c.sum()) < 1e-12 * c.size
