"""GPU: dba_droplets (device coarsening analysis) against an independent host
labelling of the downloaded field (tests/droplets_ref.py: scipy labelling plus
periodic unions, droplets ordered by their first cell > avg as
example_Reactive_ModelB/coarsening_analysis.cpp scans).

Bar: threshold, volume, count and the ordered sizes exact (integer work on the
same FP64 values); radius by the reference's formula."""
import numpy as np
import pytest

from dba_amd import abi, configs
from droplets_ref import host_droplets, radius

pytestmark = pytest.mark.gpu


def check(inst, field, a):
    avg, vol, sizes, r = inst.droplets(field)
    havg, hvol, hsizes = host_droplets(a)
    assert avg == havg
    assert vol == hvol
    assert list(sizes) == hsizes
    if hsizes:
        assert r == radius(hvol, len(hsizes), a.ndim == 3 and a.shape[0] > 1)
    return sizes


def test_crafted_2d(ctx):
    n = 64
    inst = ctx.create(configs.ch_params(n=n, tile=16, mode=abi.MODE_REGULAR))
    a = np.full((n, n), -1.0)
    a[0:3, 10:14] = 1.0            # wraps across y = 0 with the block below
    a[n - 2:n, 11:13] = 1.0
    a[20, 0] = a[20, n - 1] = 1.0  # wraps across x
    a[30:33, 30] = 1.0
    a[31, 31:36] = 0.0             # == avg cells bridge two seeds into one droplet
    a[31, 36] = 1.0
    a[50, 50:53] = 0.0             # only == avg cells: not a droplet
    a[40, 40] = 1.0                # single cells
    a[40, 42] = 1.0
    a[5, 5] = 1.0
    inst.upload(0, a)
    sizes = check(inst, 0, a)
    assert len(sizes) == 6
    # no droplets: everything below the threshold seeded at 0.5
    inst.upload(0, np.full((n, n), 0.2))
    avg, vol, sizes, r = inst.droplets(0)
    assert avg == 0.35 and vol == 0 and len(sizes) == 0 and np.isnan(r)
    inst.destroy()


def test_random_3d(ctx):
    n = 32
    inst = ctx.create(configs.modelb_params(n=n, mode=abi.MODE_REGULAR))
    rng = np.random.default_rng(11)
    z = np.where(rng.random((n, n, n)) < 0.3, 0.9, 0.1)
    z[rng.random((n, n, n)) < 0.05] = 0.5   # cells equal to the threshold
    inst.upload(2, z)
    check(inst, 2, z)
    inst.destroy()


def test_ch_evolved(ctx):
    # 16 disks of c = +1 (radius 5, jittered 32-cell lattice) in c = -1, evolved
    # by the CH step: the droplets stay apart, their interfaces are diffuse
    n = 128
    rng = np.random.default_rng(4)
    yy, xx = np.mgrid[0:n, 0:n]
    c = -np.ones((n, n))
    for i in range(4):
        for j in range(4):
            cx, cy = 16 + 32 * i + rng.integers(-4, 5), 16 + 32 * j + rng.integers(-4, 5)
            c[(xx - cx) ** 2 + (yy - cy) ** 2 <= 25] = 1.0
    p = configs.ch_params(n=n, tile=16, mode=abi.MODE_DBA, react_gamma=0.0, eps=1e-6)
    inst = ctx.create(p)
    inst.upload(0, c)
    inst.step(500, graph_chunk=100)
    c = inst.download(0).reshape(n, n)
    sizes = check(inst, 0, c)
    assert len(sizes) == 16
    inst.destroy()


def test_modelb_evolved(ctx):
    n = 32
    p = configs.modelb_params(n=n, mode=abi.MODE_DBA)
    inst = ctx.create(p)
    fields = list(configs.modelb_ic(n, n, n)) + [np.zeros((n, n, n))]
    for f, a in enumerate(fields):
        inst.upload(f, a)
    inst.step(300, graph_chunk=100)
    z = inst.download(2).reshape(n, n, n)
    check(inst, 2, z)
    inst.destroy()


def test_arguments(ctx):
    inst = ctx.create(configs.ch_params(n=64, tile=16))
    with pytest.raises(abi.DbaError):
        inst.droplets(3)   # CH has one field
    inst.destroy()
