"""GPU: single-launch steps (one launch per step, awake sets derived in every
tile's workgroup, one canonicalising kernel per call or graph chunk) against the
worklist path (step kernel + finalize) on the same instance: fields, awake sets,
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
               hist=inst.history(), work=inst.worklist())
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


@pytest.mark.parametrize("tile,mode", [(16, abi.MODE_DBA), (8, abi.MODE_DBA), (32, abi.MODE_DBA),
                                       (16, abi.MODE_REGULAR)])
def test_k93_single_launch_equals_worklist(ctx, tile, mode):
    n = 256
    p = configs.dendrite_k93_params(n=n, tile=tile, mode=mode, history_cap=400)
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


def test_conserved_and_batched_use_the_worklist_path(ctx):
    """Conserved DBA and multi-instance groups are not single-launch eligible;
    they run (and match the oracle elsewhere) with the switch on."""
    n = 64
    p = configs.ch_params(n=n, tile=16, mode=abi.MODE_CONSERVED, eps=1e-3)
    c = configs.ch_ic(n, n)
    a = ctx.create(p)
    a.upload(0, c)
    a.step(20, graph_chunk=10)
    b = ctx.create(p)
    b.upload(0, c)
    ctx.set_single_launch(False)
    b.step(20, graph_chunk=10)
    ctx.set_single_launch(True)
    assert np.array_equal(a.download(0), b.download(0))
    a.destroy()
    b.destroy()
