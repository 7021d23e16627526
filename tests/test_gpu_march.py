"""GPU: the column-march single-launch CH step (ch_march_kernel, taken by one
large instance of full 32x32 tiles, periodic, not conserved: >= 4096 tiles)
against the worklist path (step kernel + finalize, mode 0) and the oracle.

Bar: bit-identical to the worklist path in fields, awake sets, history and
counters, over runs with sleeping and waking tiles and both stencils; bit-exact
to the oracle."""
import numpy as np
import pytest

from dba_amd import abi, configs
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu


def local_ic(nx, ny, seed=3):
    c = configs.ch_ic(nx, ny, seed=seed, cbar=0.05)
    y, x = np.mgrid[0:ny, 0:nx]
    c[((x // 32) % 9 < 5) | ((y // 32) % 10 < 5)] = 0.05   # bands of quiet tiles: the inner ones sleep
    return c


def run(ctx, p, c0, steps, mode, chunk):
    ctx.set_single_launch(mode)
    try:
        inst = ctx.create(p)
        inst.upload(0, c0)
        inst.step(steps, graph_chunk=chunk)
        info = inst.info()
        out = (inst.download(0), inst.active(0)[1], inst.history(), info.cell_updates, info.steps_done,
               info.single_launch)
        inst.destroy()
        return out
    finally:
        ctx.set_single_launch(1)


@pytest.mark.parametrize("mode,stencil,nx,ny", [(abi.MODE_DBA, 9, 4096, 2048), (abi.MODE_DBA, 5, 4096, 2048),
                                                (abi.MODE_REGULAR, 9, 4096, 2048), (abi.MODE_DBA, 9, 2048, 2048),
                                                (abi.MODE_DBA, 9, 4096, 4096), (abi.MODE_DBA, 5, 8192, 4096)])
def test_march_equals_worklist(ctx, mode, stencil, nx, ny):
    # strips of 8 tiles on 4096 x 2048 (128 x 64 tiles), 4 on 2048^2, 16 on 4096^2,
    # 32 on 8192 x 4096: the most tiles per strip that keep >= 1024 strips
    p = configs.ch_params(n=nx, ny=ny, tile=32, mode=mode, stencil=stencil, eps=2e-3, react_gamma=0.01,
                          cbar=0.05, history_cap=64)
    c0 = local_ic(nx, ny)
    a = run(ctx, p, c0, 40, 1, 20)
    b = run(ctx, p, c0, 40, 0, 20)
    assert a[5] == 1   # stepped with single-launch steps
    assert np.array_equal(a[0], b[0]), np.max(np.abs(a[0] - b[0]))
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]) and a[3:5] == b[3:5]
    if mode == abi.MODE_DBA:
        assert 0 < a[1].sum() < a[1].size   # some tiles sleep, some run
        assert a[2][:, 0].min() < (nx // 32) * (ny // 32)


def test_march_matches_oracle(ctx):
    nx, ny = 2048, 4096   # 64 x 128 tiles
    p = configs.ch_params(n=nx, ny=ny, tile=32, mode=abi.MODE_DBA, eps=2e-3, react_gamma=0.01, cbar=0.05)
    c0 = local_ic(nx, ny, seed=9)
    got = run(ctx, p, c0, 12, 1, 12)[0]
    o = OracleSim(p)
    o.set_field(0, c0)
    o.step(12)
    ref = o.get_field(0)
    o.close()
    assert np.array_equal(got, ref), np.max(np.abs(got - ref))
