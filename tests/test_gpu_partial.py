"""GPU: grids that are not a multiple of the activation tile (a partial last
tile column and row, as the reference's ceil(N / block) blocks with bounds
guards, example_Dendritic_growth/kernel.cu:102-104, 337-407), against the
oracle, whose tiles are ceil(N / tile) as well.

Bar: CH bit-exact (conserved DBA <= 1e-12), K93 <= 1e-10 absolute; the awake
set after the run and the per-step active-tile history identical to the
oracle's; the single-launch and worklist paths bit-identical."""
import numpy as np
import pytest

from dba_amd import abi, configs
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu


def run(ctx, p, fields, steps, chunk=0):
    inst = ctx.create(p)
    for f, a in enumerate(fields):
        inst.upload(f, a)
    inst.step(steps, graph_chunk=chunk)
    out = [inst.download(f) for f in range(len(fields))]
    return inst, out


def oracle(p, fields, steps):
    o = OracleSim(p)
    for f, a in enumerate(fields):
        o.set_field(f, a)
    o.step(steps)
    return o


def ch_ic_local(nx, ny, seed=5, frac=0.4):
    c = configs.ch_ic(nx, ny, seed=seed, cbar=0.05)
    c[:, int(nx * frac):] = 0.05
    return c


@pytest.mark.parametrize("mode", [abi.MODE_REGULAR, abi.MODE_DBA, abi.MODE_CONSERVED])
@pytest.mark.parametrize("nx,ny,tile,stencil", [(200, 136, 32, 9), (100, 75, 16, 5), (58, 90, 8, 9)])
def test_ch_partial_tiles(ctx, mode, nx, ny, tile, stencil):
    p = configs.ch_params(n=nx, ny=ny, tile=tile, mode=mode, stencil=stencil, eps=1e-3, react_gamma=0.01,
                          cbar=0.05, history_cap=64)
    assert nx % tile and ny % tile
    c0 = ch_ic_local(nx, ny)
    steps = 40
    inst, (got,) = run(ctx, p, [c0], steps)
    o = oracle(p, [c0], steps)
    ref = o.get_field(0)
    if mode == abi.MODE_CONSERVED:
        assert np.max(np.abs(got - ref)) <= 1e-12
        assert abs(got.sum() - c0.sum()) < 1e-10
    else:
        assert np.array_equal(got, ref), np.max(np.abs(got - ref))
    if mode != abi.MODE_REGULAR:
        cnt, mask = inst.active(0)
        assert np.array_equal(mask, o.awake(0))
        assert 0 < cnt < inst.info().ntiles
        assert np.array_equal(inst.history()[:, 0], o.history()[:, 0])
    ntx, nty = -(-nx // tile), -(-ny // tile)
    assert inst.info().ntiles == ntx * nty
    if mode == abi.MODE_REGULAR:   # cell-updates count the partial tiles' own cells
        assert inst.info().cell_updates == steps * nx * ny
    inst.destroy()


@pytest.mark.parametrize("nx,ny", [(130, 130), (150, 120)])
def test_k93_partial_tiles(ctx, nx, ny):
    p = configs.dendrite_k93_params(n=nx, ny=ny, tile=16, history_cap=200)
    phi, T = configs.dendrite_k93_ic(nx, ny, seeds=[(nx // 2, ny // 2), (nx - 6, ny - 8)])
    steps = 150
    inst, (gp, gT) = run(ctx, p, [phi, T], steps, chunk=50)
    o = oracle(p, [phi, T], steps)
    for g, r in ((gp, o.get_field(0)), (gT, o.get_field(1))):
        assert np.max(np.abs(g - r)) <= 1e-10
    for s in (0, 1):
        assert np.array_equal(inst.active(s)[1], o.awake(s))
    assert np.array_equal(inst.history(), o.history())
    inst.destroy()


@pytest.mark.parametrize("kind", ["ch", "k93"])
def test_partial_single_launch_equals_worklist(ctx, kind):
    if kind == "ch":
        p = configs.ch_params(n=200, ny=136, tile=32, eps=1e-3, react_gamma=0.01, cbar=0.05, history_cap=64)
        fields = [ch_ic_local(200, 136)]
    else:
        p = configs.dendrite_k93_params(n=150, ny=120, tile=16, history_cap=64)
        fields = list(configs.dendrite_k93_ic(150, 120, seeds=[(144, 112)]))
    res = []
    for mode in (0, 1):
        ctx.set_single_launch(mode)
        try:
            inst, out = run(ctx, p, fields, 60, chunk=20)
            info = inst.info()
            res.append((out, [inst.active(s)[1] for s in range(len(fields))], inst.history(),
                        info.cell_updates, info.field_updates, info.steps_done))
            inst.destroy()
        finally:
            ctx.set_single_launch(1)
    a, b = res
    for x, y in zip(a[0], b[0]):
        assert np.array_equal(x, y)
    for x, y in zip(a[1], b[1]):
        assert np.array_equal(x, y)
    assert np.array_equal(a[2], b[2]) and a[3:] == b[3:]
