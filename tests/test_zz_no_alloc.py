"""GPU: dba_droplets_noalloc (all buffers in the library's static device
scratch) against dba_droplets (per-call allocations, host-side ordering) on
the same instances: threshold, volume, count and the ordered sizes identical.
Two calls in a row, the second on a smaller grid, so the second call cannot
pass on what the first left in the shared scratch. A grid above the scratch's
2^22 cells is refused before any launch."""
import numpy as np
import pytest

from dba_amd import abi, configs

pytestmark = pytest.mark.gpu


def same(inst, field):
    a = inst.droplets(field)
    b = inst.droplets(field, noalloc=True)
    assert a[0] == b[0]            # threshold
    assert a[1] == b[1]            # volume
    assert list(a[2]) == list(b[2])  # sizes in the reference's order
    assert (np.isnan(a[3]) and np.isnan(b[3])) or a[3] == b[3]
    return b


def droplet_field(n, seed, frac):
    rng = np.random.default_rng(seed)
    a = np.where(rng.random((n, n)) < frac, 0.9, -0.9)
    a[rng.random((n, n)) < 0.03] = 0.0   # cells equal to the threshold
    return a


def test_two_calls_second_smaller(ctx):
    big = ctx.create(configs.ch_params(n=512, tile=32, mode=abi.MODE_REGULAR))
    big.upload(0, droplet_field(512, 1, 0.35))
    r1 = same(big, 0)
    assert len(r1[2]) > 1000
    small = ctx.create(configs.ch_params(n=96, tile=16, mode=abi.MODE_REGULAR))
    small.upload(0, droplet_field(96, 2, 0.2))
    r2 = same(small, 0)
    assert 0 < len(r2[2]) < len(r1[2])
    # a field without droplets after one with many: counters start from zero
    small.upload(0, np.full((96, 96), 0.2))
    r3 = same(small, 0)
    assert r3[1] == 0 and len(r3[2]) == 0
    big.destroy()
    small.destroy()


def test_cap_and_3d(ctx):
    n = 32
    inst = ctx.create(configs.modelb_params(n=n, mode=abi.MODE_REGULAR))
    rng = np.random.default_rng(11)
    z = np.where(rng.random((n, n, n)) < 0.3, 0.9, 0.1)
    inst.upload(2, z)
    full = same(inst, 2)
    part = inst.droplets(2, cap=7, noalloc=True)
    assert list(part[2]) == list(full[2][:7])
    inst.destroy()


def test_ch_evolved(ctx):
    n = 256
    p = configs.ch_params(n=n, tile=32, mode=abi.MODE_DBA, react_gamma=0.0, eps=1e-6)
    inst = ctx.create(p)
    inst.upload(0, configs.ch_ic(n, n))
    inst.step(2000, graph_chunk=100)
    same(inst, 0)
    inst.destroy()


def test_refuses_grid_above_scratch(ctx):
    n = 2080  # 4,326,400 cells > 2^22
    inst = ctx.create(configs.ch_params(n=n, tile=32, mode=abi.MODE_REGULAR))
    with pytest.raises(abi.DbaError, match="static scratch"):
        inst.droplets(0, noalloc=True)
    inst.destroy()
