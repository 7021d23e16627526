"""GPU parity of the reference's own 3D models (libdba_hip.so vs the oracle).

Bar: |diff| <= 1e-10 (dendrite: absolute on p and T; Model B: absolute on the
volume fractions x, y, z of order 0.2), since these steps call atan (dendrite),
log and exp (Model B), which differ by a few ulp between the device library
and glibc; identical awake sets; conserved totals to 1e-11.
"""
import numpy as np
import pytest

from dba_amd import abi, configs
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu

TOL = 1e-10


def pair(ctx, p, fields):
    inst = ctx.create(p)
    o = OracleSim(p)
    for f, a in enumerate(fields):
        inst.upload(f, a)
        o.set_field(f, a)
    return inst, o


@pytest.mark.parametrize("mode", [abi.MODE_REGULAR, abi.MODE_DBA, abi.MODE_CONSERVED])
@pytest.mark.parametrize("nz", [8, 1, 24])
def test_modelb_matches_oracle(ctx, mode, nz):
    # 48 wide: the uniform half keeps tile columns that neither the interface
    # nor its periodic image wakes, so DBA modes have sleeping tiles; nz = 24
    # has blocks whose box needs no wrap (the direct-offset loads), next to
    # sleeping neighbours whose current buffer differs
    n = 48
    X = list(configs.modelb_ic(n, n, nz)) + [np.zeros((nz, n, n))]
    if mode != abi.MODE_REGULAR:
        X[0][..., n // 2:] = 0.2
        X[1][..., n // 2:] = 0.2
        X[2][..., n // 2:] = 0.01
    p = configs.modelb_params(n=n, nz=nz, mode=mode, eps=1e-4, history_cap=64)
    inst, o = pair(ctx, p, X)
    for chunk in (1, 4, 15):
        inst.step(chunk)
        o.step(chunk)
        for f in range(3):
            d = np.max(np.abs(inst.download(f) - o.get_field(f)))
            assert d <= TOL, (f, d)
        if mode != abi.MODE_REGULAR:
            assert np.array_equal(inst.active(0)[1], o.awake(0))
    if mode != abi.MODE_REGULAR:
        assert inst.active(0)[0] < inst.info().ntiles
    if mode == abi.MODE_CONSERVED:
        for f in range(3):
            assert abs(inst.download(f).sum() - X[f].sum()) < 1e-11
    inst.destroy()


@pytest.mark.parametrize("mode", [abi.MODE_REGULAR, abi.MODE_DBA, abi.MODE_CONSERVED])
def test_modelb_band_sweep_matches_oracle(ctx, mode):
    """192 x 192 x 128 holds 73728 blocks, more than the 65536-workgroup grid:
    with every block awake the list is taken in the band-sweep order (each
    XCD's band of block columns z-layer by z-layer, some workgroups taking two
    blocks), the order used for a full single-instance list."""
    n, nz = 192, 128
    X = list(configs.modelb_ic(n, n, nz)) + [np.zeros((nz, n, n))]
    p = configs.modelb_params(n=n, nz=nz, mode=mode, eps=1e-4)
    inst, o = pair(ctx, p, X)
    for chunk in (1, 2):
        inst.step(chunk)
        o.step(chunk)
        for f in range(3):
            d = np.max(np.abs(inst.download(f) - o.get_field(f)))
            assert d <= TOL, (f, d)
        if mode != abi.MODE_REGULAR:
            assert np.array_equal(inst.active(0)[1], o.awake(0))
            assert inst.active(0)[0] == inst.info().ntiles   # the full list
    inst.destroy()


@pytest.mark.parametrize("mode", [abi.MODE_REGULAR, abi.MODE_DBA])
@pytest.mark.parametrize("nz", [32, 1])
def test_k94_matches_oracle(ctx, mode, nz):
    n = 32
    p0, T0 = configs.dendrite_k94_ic(n, n, nz)
    p = configs.dendrite_k94_params(n=n, nz=nz, mode=mode, history_cap=64)
    inst, o = pair(ctx, p, [p0, T0])
    for chunk in (1, 5, 30):
        inst.step(chunk)
        o.step(chunk)
        for f in (0, 1):
            d = np.max(np.abs(inst.download(f) - o.get_field(f)))
            assert d <= TOL, (f, d)
        if mode == abi.MODE_DBA:
            for s in (0, 1):
                assert np.array_equal(inst.active(s)[1], o.awake(s))
    if mode == abi.MODE_DBA:
        h = inst.history()
        assert np.array_equal(h, o.history())
        assert h[-1, 0] < inst.info().ntiles
    inst.destroy()


def test_k94_latent_heat_dropped_matches_oracle(ctx):
    """Metric_T = 0.01: p-awake / T-asleep blocks occur (about 3500 block-steps
    over 36 steps), where the reference drops the latent heat
    (example_Dendritic_growth/kernel.cu:411-574)."""
    n = 32
    p0, T0 = configs.dendrite_k94_ic(n, n, n)
    p = configs.dendrite_k94_params(n=n, mode=abi.MODE_DBA, thresh_b=0.01)
    inst, o = pair(ctx, p, [p0, T0])
    for chunk in (1, 5, 30):
        inst.step(chunk)
        o.step(chunk)
        for f in (0, 1):
            d = np.max(np.abs(inst.download(f) - o.get_field(f)))
            assert d <= TOL, (f, d)
        for s in (0, 1):
            assert np.array_equal(inst.active(s)[1], o.awake(s))
    inst.destroy()


def test_modelb_reference_size_runs(ctx):
    """64^3 blocks of the reference Input.txt layout, 50 steps through a graph."""
    n = 64
    X = list(configs.modelb_ic(n, n, n)) + [np.zeros((n, n, n))]
    p = configs.modelb_params(n=n, mode=abi.MODE_DBA)
    inst = ctx.create(p)
    for f, a in enumerate(X):
        inst.upload(f, a)
    inst.step(50, graph_chunk=25)
    assert inst.check_finite()
    b = 1.0 - inst.download(0) - inst.download(1) - inst.download(2)
    assert b.min() > 0.0
    inst.destroy()


@pytest.mark.parametrize("mode", [abi.MODE_REGULAR, abi.MODE_DBA])
def test_modelb_superblocks_match_block_kernels(ctx, mode, monkeypatch):
    """The superblock kernels (2x2x2 blocks per workgroup) against the block
    kernels on the same steps: DBA_MODELB_SUPERBLOCKS=1 takes every step on
    superblocks, including the sparse ones with sleeping blocks inside a
    superblock, =0 none; fields bit-identical, same awake sets and history,
    and the superblock run within the bar of the oracle."""
    n, nz = 48, 24
    X = list(configs.modelb_ic(n, n, nz)) + [np.zeros((nz, n, n))]
    if mode != abi.MODE_REGULAR:
        X[0][..., n // 2:] = 0.2
        X[1][..., n // 2:] = 0.2
        X[2][..., n // 2:] = 0.01
    p = configs.modelb_params(n=n, nz=nz, mode=mode, eps=1e-4, history_cap=64)
    runs = {}
    for env in ("1", "0"):
        monkeypatch.setenv("DBA_MODELB_SUPERBLOCKS", env)
        inst = ctx.create(p)
        for f, a in enumerate(X):
            inst.upload(f, a)
        snaps = []
        for chunk in (1, 4, 15):
            inst.step(chunk)
            snaps.append(([inst.download(f) for f in range(3)], inst.active(0)[1]))
        runs[env] = (snaps, inst.history())
        inst.destroy()
    for (fa, aa), (fb, ab) in zip(runs["1"][0], runs["0"][0]):
        for f in range(3):
            assert np.array_equal(fa[f], fb[f]), f
        assert np.array_equal(aa, ab)
    assert np.array_equal(runs["1"][1], runs["0"][1])
    o = OracleSim(p)
    for f, a in enumerate(X):
        o.set_field(f, a)
    o.step(20)
    for f in range(3):
        assert np.max(np.abs(runs["1"][0][-1][0][f] - o.get_field(f))) <= TOL
    if mode != abi.MODE_REGULAR:
        h = runs["1"][1]
        assert h[:, 0].min() < p.nx * p.ny * p.nz // 64   # sleeping blocks on some step
