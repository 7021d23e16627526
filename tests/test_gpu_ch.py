"""GPU parity of the reactive Cahn-Hilliard DBA step (libdba_hip.so vs the oracle).

Bar: bit-exact fields for regular and DBA modes (both sides are compiled with
FP contraction off and evaluate the same operations in the same order);
|diff| <= 1e-12 for conserved DBA, whose mean shift is a sum in a different
order on each side. Active-tile sets are checked bit-exactly every step
against a host re-evaluation of the activation rule on the GPU's own fields.
"""
import numpy as np
import pytest

from dba_amd import abi, configs
from oracle.oracle import OracleSim

from gpu_util import dilate_faces, lap_periodic, tile_any

pytestmark = pytest.mark.gpu


def local_noise_ic(n, frac=0.25, seed=1234):
    c = configs.ch_ic(n, n, seed=seed)
    c[:, int(n * frac):] = 0.0
    return c


def run_pair(ctx, p, c0, steps):
    inst = ctx.create(p)
    inst.upload(0, c0)
    inst.step(steps)
    got = inst.download(0)
    o = OracleSim(p)
    o.set_field(0, c0)
    o.step(steps)
    return inst, o, got, o.get_field(0)


@pytest.mark.parametrize("mode", [abi.MODE_REGULAR, abi.MODE_DBA, abi.MODE_CONSERVED])
@pytest.mark.parametrize("stencil", [5, 9])
@pytest.mark.parametrize("tile,threads", [(8, 64), (16, 128), (32, 256), (16, 256), (32, 512), (32, 1024), (16, 0)])
def test_ch_matches_oracle(ctx, mode, stencil, tile, threads):
    n = max(128, 8 * tile)
    p = configs.ch_params(n=n, tile=tile, mode=mode, stencil=stencil, eps=1e-3, threads=threads,
                          react_gamma=0.01, cbar=0.05)
    c0 = local_noise_ic(n)
    inst, o, got, ref = run_pair(ctx, p, c0, 40)
    if mode == abi.MODE_CONSERVED:
        assert np.max(np.abs(got - ref)) <= 1e-12
        assert abs(got.sum() - c0.sum()) < 1e-10
    else:
        assert np.array_equal(got, ref), np.max(np.abs(got - ref))
    if mode != abi.MODE_REGULAR:
        cnt, mask = inst.active(0)
        assert np.array_equal(mask, o.awake(0))
        assert cnt < inst.info().ntiles  # some tiles sleep: DBA exercised
    inst.destroy()


@pytest.mark.parametrize("stencil", [5, 9])
def test_ch_active_sets_bit_exact(ctx, stencil):
    n, tile = 256, 16
    p = configs.ch_params(n=n, tile=tile, mode=abi.MODE_DBA, stencil=stencil, eps=1e-3, history_cap=64)
    inst = ctx.create(p)
    inst.upload(0, local_noise_ic(n, 0.3))
    ntx = n // tile
    for step in range(40):
        c = inst.download(0)
        _, a_now = inst.active(0)
        wl = inst.worklist()
        assert np.array_equal(wl, np.nonzero(a_now)[0])  # ascending, copy_if order
        inst.step(1)
        _, a_next = inst.active(0)
        Ncell = np.abs(lap_periodic(c, stencil, p.dx)) > p.thresh_a
        N = tile_any(Ncell, tile) & (a_now == 1)
        expect = dilate_faces(N, ntx, ntx, periodic=True)
        assert np.array_equal(a_next, expect), step
    h = inst.history()
    assert h.shape[0] == 40 and h[0, 0] == ntx * ntx
    inst.destroy()


def test_ch_graph_equals_eager(ctx):
    n = 128
    p = configs.ch_params(n=n, tile=16, mode=abi.MODE_CONSERVED, eps=1e-3)
    c0 = local_noise_ic(n)
    a = ctx.create(p)
    b = ctx.create(p)
    a.upload(0, c0)
    b.upload(0, c0)
    a.step(37)
    b.step(37, graph_chunk=10)
    assert np.array_equal(a.download(0), b.download(0))
    assert np.array_equal(a.active(0)[1], b.active(0)[1])
    a.destroy()
    b.destroy()


def test_ch_batched_equals_alone(ctx):
    sizes = [64, 128, 96, 160]
    insts, alone = [], []
    for k, n in enumerate(sizes):
        p = configs.ch_params(n=n, tile=16, mode=abi.MODE_DBA, eps=1e-3, react_gamma=1e-3 * (k + 1))
        c0 = local_noise_ic(n, 0.4, seed=100 + k)
        a = ctx.create(p)
        a.upload(0, c0)
        b = ctx.create(p)
        b.upload(0, c0)
        insts.append(a)
        alone.append(b)
    ctx.step(insts, 25)
    for b in alone:
        b.step(25)
    for a, b in zip(insts, alone):
        assert np.array_equal(a.download(0), b.download(0))
        assert np.array_equal(a.history(), b.history())
        a.destroy()
        b.destroy()


def test_ch_full_size_mean_law(ctx):
    """1024^2 configuration: the discrete mean law holds through the GPU step."""
    n = 1024
    p = configs.ch_params(n=n, tile=32, mode=abi.MODE_DBA, react_gamma=1e-3, cbar=0.0)
    c0 = configs.ch_ic(n, n, cbar=0.02)
    inst = ctx.create(p)
    inst.upload(0, c0)
    inst.step(200, graph_chunk=50)
    c = inst.download(0)
    m = p.cbar + (1 - p.react_gamma * p.dt) ** 200 * (c0.mean() - p.cbar)
    assert abs(c.mean() - m) < 1e-13
    assert np.isfinite(c).all()
    inst.destroy()
