"""GPU: the fixed-order reductions, bit for bit (dba_field_sum).

dba_field_sum fixes its whole order: 256 workgroups x 256 threads; thread
(b, t) adds the cells b * 256 + t + k * 65536, k = 0, 1, ... in order; each
wave reduces its 64 partials by the shift tree v[i] += v[i + off],
off = 32, 16, 8, 4, 2, 1 (wave_sum_down: the last four levels are DPP row
shifts), thread 0 adds the 4 wave sums in wave order, and one thread adds the
256 workgroup sums in order. The host mirror below restates that order; on
values of mixed magnitude (+-1e16, +-1, +-2^-30) every other order tried here
gives another double, so a changed tree, a DPP shift in the wrong direction or
a reordered block sum fails the test. The conserved shift, the KH CFL speeds
and the step kernels' per-tile sums use the same tree (common.h block_sum)."""
import numpy as np
import pytest

from dba_amd import abi, configs

pytestmark = pytest.mark.gpu

BLOCKS, NT = 256, 256


def tree64(v, offsets=(32, 16, 8, 4, 2, 1)):
    v = list(v)
    for off in offsets:
        for i in range(off):
            v[i] = v[i] + v[i + off]
    return v[0]


def mirror_sum(a, offsets=(32, 16, 8, 4, 2, 1), wave_order=range(4), block_order=range(BLOCKS)):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    n = a.size
    stride = BLOCKS * NT
    parts = []
    for b in range(BLOCKS):
        th = []
        for t in range(NT):
            s = 0.0
            for i in range(b * NT + t, n, stride):
                s = s + float(a[i])
            th.append(s)
        ws = [tree64(th[64 * w:64 * w + 64], offsets) for w in range(4)]
        r = 0.0
        for w in wave_order:
            r = r + ws[w]
        parts.append(r)
    total = 0.0
    for b in block_order:
        total = total + parts[b]
    return total


def adversarial(n, seed):
    rng = np.random.default_rng(seed)
    mags = rng.choice([1e16, 1.0, 2.0 ** -30], size=n, p=[0.05, 0.6, 0.35])
    return mags * rng.choice([-1.0, 1.0], size=n) * (1.0 + rng.random(n))


@pytest.mark.parametrize("nx,ny", [(256, 256), (512, 384)])
def test_field_sum_tree_bit_exact(ctx, nx, ny):
    a = adversarial(nx * ny, seed=nx + ny).reshape(ny, nx)
    inst = ctx.create(configs.ch_params(n=nx, ny=ny, tile=32, mode=abi.MODE_REGULAR))
    inst.upload(0, a)
    got = inst.field_sum(0)
    inst.destroy()
    expect = mirror_sum(a)
    assert got == expect, (got, expect)
    # the data tells the orders apart: other trees and orders give other doubles
    others = {
        "shift up (lanes i - off)": mirror_sum(a, offsets=(1, 2, 4, 8, 16, 32)),
        "waves reversed": mirror_sum(a, wave_order=range(3, -1, -1)),
        "blocks reversed": mirror_sum(a, block_order=range(BLOCKS - 1, -1, -1)),
        "sequential": float(sum(float(v) for v in a.reshape(-1))),
    }
    assert all(v != expect for v in others.values()), others
