"""GPU parity of the Kobayashi-93 dendrite DBA step (libdba_hip.so vs the oracle).

Bar: |diff| <= 1e-10 absolute on phi and T after the stated steps (the only
operation that differs between the sides is the device atan against glibc's,
a few ulp per call). Active sets are checked bit-exactly every step against
the reference rule re-evaluated on the GPU's own fields:
Np = any(Metric_p <= phi' <= 1 - Metric_p) over p-awake tiles,
NT = any(|T' - T| >= Metric_T) over T-awake tiles, then face dilation without
wrap (zero-flux boundaries).
"""
import numpy as np
import pytest

from dba_amd import abi, configs
from oracle.oracle import OracleSim

from gpu_util import dilate_faces, tile_any

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.mark.parametrize("mode", [abi.MODE_REGULAR, abi.MODE_DBA])
@pytest.mark.parametrize("tile,threads", [(16, 256), (8, 64), (32, 256), (16, 128), (32, 512), (16, 1024), (32, 0)])
def test_k93_matches_oracle(ctx, mode, tile, threads):
    n = 128
    p = configs.dendrite_k93_params(n=n, tile=tile, mode=mode, threads=threads)
    phi, T = configs.dendrite_k93_ic(n, n)
    inst = ctx.create(p)
    inst.upload(0, phi)
    inst.upload(1, T)
    inst.step(150)
    o = OracleSim(p)
    o.set_field(0, phi)
    o.set_field(1, T)
    o.step(150)
    for f in (0, 1):
        d = np.max(np.abs(inst.download(f) - o.get_field(f)))
        assert d <= TOL, (f, d)
    if mode == abi.MODE_DBA:
        for s in (0, 1):
            assert np.array_equal(inst.active(s)[1], o.awake(s))
    inst.destroy()


@pytest.mark.parametrize("single_launch", [1, 0])
@pytest.mark.parametrize("tile,threads", [(16, 0), (32, 256)])
def test_k93_latent_heat_dropped_matches_oracle(ctx, tile, threads, single_launch):
    """Metric_T = 0.03 puts the T set to sleep while the interface keeps the p
    set awake: about 1700 tile-steps are p-awake and T-asleep, where the
    reference drops the latent heat (example_Dendritic_growth/kernel.cu:411-574).
    An oracle that kept it ends 1.06 apart (tests/test_oracle_stale.py)."""
    n = 128
    p = configs.dendrite_k93_params(n=n, tile=tile, threads=threads)
    p.thresh_b = 0.03
    phi, T = configs.dendrite_k93_ic(n, n)
    ctx.set_single_launch(single_launch)
    try:
        inst = ctx.create(p)
        inst.upload(0, phi)
        inst.upload(1, T)
        o = OracleSim(p)
        o.set_field(0, phi)
        o.set_field(1, T)
        dropped = 0
        for chunk in (1, 9, 40, 100):
            ap, at = o.awake(0), o.awake(1)
            dropped += int((ap & (1 - at)).sum())
            inst.step(chunk)
            o.step(chunk)
            for f in (0, 1):
                d = np.max(np.abs(inst.download(f) - o.get_field(f)))
                assert d <= TOL, (f, d)
            for s in (0, 1):
                assert np.array_equal(inst.active(s)[1], o.awake(s))
        assert dropped > 0
        inst.destroy()
    finally:
        ctx.set_single_launch(1)


@pytest.mark.parametrize("j,theta0,bc", [(4, 0.0, abi.BC_NOFLUX), (6, 0.2, abi.BC_PERIODIC)])
def test_k93_active_sets_bit_exact(ctx, j, theta0, bc):
    n, tile = 128, 16
    p = configs.dendrite_k93_params(n=n, tile=tile, aniso_j=j, theta0=theta0, bc=bc, history_cap=100)
    phi, T = configs.dendrite_k93_ic(n, n, seeds=[(40, 50), (90, 80)])
    inst = ctx.create(p)
    inst.upload(0, phi)
    inst.upload(1, T)
    ntx = n // tile
    for step in range(60):
        ph0, T0 = inst.download(0), inst.download(1)
        _, ap = inst.active(0)
        _, at = inst.active(1)
        inst.step(1)
        ph1, T1 = inst.download(0), inst.download(1)
        # sleeping tiles keep their values
        cell_ap = np.kron(ap.reshape(ntx, ntx), np.ones((tile, tile), dtype=np.uint8)).astype(bool)
        cell_at = np.kron(at.reshape(ntx, ntx), np.ones((tile, tile), dtype=np.uint8)).astype(bool)
        assert np.array_equal(ph1[~cell_ap], ph0[~cell_ap])
        assert np.array_equal(T1[~cell_at], T0[~cell_at])
        Np = tile_any((ph1 >= p.thresh_a) & (ph1 <= 1.0 - p.thresh_a), tile) & (ap == 1)
        NT = tile_any(np.abs(T1 - T0) >= p.thresh_b, tile) & (at == 1)
        per = bc == abi.BC_PERIODIC
        assert np.array_equal(inst.active(0)[1], dilate_faces(Np, ntx, ntx, per)), step
        assert np.array_equal(inst.active(1)[1], dilate_faces(NT, ntx, ntx, per)), step
    inst.destroy()


def test_k93_config_size_runs(ctx):
    """512^2 configuration, 400 steps through a graph: matches the oracle."""
    n = 512
    p = configs.dendrite_k93_params(n=n, tile=16)
    phi, T = configs.dendrite_k93_ic(n, n)
    inst = ctx.create(p)
    inst.upload(0, phi)
    inst.upload(1, T)
    inst.step(400, graph_chunk=100)
    o = OracleSim(p)
    o.set_field(0, phi)
    o.set_field(1, T)
    o.step(400)
    for f in (0, 1):
        assert np.max(np.abs(inst.download(f) - o.get_field(f))) <= TOL
    assert inst.active(0)[0] < inst.info().ntiles // 4
    inst.destroy()
