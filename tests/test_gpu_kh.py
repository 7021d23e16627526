"""GPU parity of the compressible-Euler (KH) DBA step (libdba_hip.so vs the oracle).

Bar: bit-exact conserved fields for regular and DBA modes (the step uses only
+, -, *, / and sqrt, correctly rounded on both sides, in the same order);
|diff| <= 1e-12 relative for conserved DBA (mean shift summed in a different
order); identical awake sets every step; identical CFL time.
"""
import numpy as np
import pytest

from dba_amd import abi, configs
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu


def pair(ctx, p, U):
    inst = ctx.create(p)
    o = OracleSim(p)
    for f in range(4):
        inst.upload(f, U[f])
        o.set_field(f, U[f])
    return inst, o


def compare(inst, o, exact=True):
    for f in range(4):
        a, b = inst.download(f), o.get_field(f)
        if exact:
            assert np.array_equal(a, b), (f, np.max(np.abs(a - b)))
        else:
            assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b)), f


@pytest.mark.parametrize("mode", [abi.MODE_REGULAR, abi.MODE_DBA, abi.MODE_CONSERVED])
def test_reference_model_matches_oracle(ctx, mode):
    """gamma 5/3, periodic, fixed dt, 8x8 blocks, 8-gradient detector."""
    n = 128
    p = configs.kh_reference_params(n=n, tile=8, mode=mode, dx=1.0 / n, dt=2e-4, eps=8.0, history_cap=64)
    U = configs.kh_reference_ic(n, dx=1.0 / n)
    inst, o = pair(ctx, p, U)
    for chunk in (1, 1, 3, 10, 20):
        inst.step(chunk)
        o.step(chunk)
        compare(inst, o, exact=mode != abi.MODE_CONSERVED)
        if mode != abi.MODE_REGULAR:
            assert np.array_equal(inst.active(0)[1], o.awake(0))
    if mode != abi.MODE_REGULAR:
        h = inst.history()
        assert np.array_equal(h[:, 0], o.history()[:, 0])
        assert h[-1, 0] < inst.info().ntiles  # DBA exercised
    if mode == abi.MODE_CONSERVED:
        for f in range(4):
            assert abs(inst.download(f).sum() - U[f].sum()) <= 1e-12 * max(1e-30, np.abs(U[f]).sum())
    inst.destroy()


def test_non_power_of_two_spacing_matches_oracle(ctx):
    """dx = dy = 0.003: the step divides by the spacings and the cell area (the
    power-of-two grids of the other tests multiply by exact reciprocals)."""
    n = 128
    p = configs.kh_reference_params(n=n, tile=8, mode=abi.MODE_DBA, dx=0.003, dt=2e-4, eps=8.0, history_cap=64)
    U = configs.kh_reference_ic(n, dx=0.003)
    inst, o = pair(ctx, p, U)
    for chunk in (1, 3, 10, 20):
        inst.step(chunk)
        o.step(chunk)
        compare(inst, o, exact=True)
        assert np.array_equal(inst.active(0)[1], o.awake(0))
    inst.destroy()


@pytest.mark.parametrize("mode", [abi.MODE_REGULAR, abi.MODE_DBA])
@pytest.mark.parametrize("tile,threads", [(16, 256), (8, 64), (16, 128)])
def test_config_model_matches_oracle(ctx, mode, tile, threads):
    """gamma 1.4, reflective y walls, CFL 0.4, density-gradient detector."""
    n = 256
    p = configs.kh_params(n=n, tile=tile, mode=mode, eps=20.0, threads=threads, history_cap=64)
    U = configs.kh_ic(n)
    inst, o = pair(ctx, p, U)
    for chunk in (1, 2, 7, 30):
        inst.step(chunk)
        o.step(chunk)
        compare(inst, o)
        assert inst.info().sim_time == o.time()
        if mode == abi.MODE_DBA:
            assert np.array_equal(inst.active(0)[1], o.awake(0))
    if mode == abi.MODE_DBA:
        assert inst.active(0)[0] < inst.info().ntiles
    inst.destroy()


def test_graph_equals_eager(ctx):
    n = 128
    p = configs.kh_params(n=n, tile=16, mode=abi.MODE_DBA, eps=20.0)
    U = configs.kh_ic(n)
    a, _ = pair(ctx, p, U)
    b, _ = pair(ctx, p, U)
    a.step(45)
    b.step(45, graph_chunk=20)
    for f in range(4):
        assert np.array_equal(a.download(f), b.download(f))
    assert a.info().sim_time == b.info().sim_time
    a.destroy()
    b.destroy()


def test_full_size_conservation(ctx):
    """2048^2 configuration in regular mode, 100 CFL steps: mass, x-momentum and
    energy conserved (plain DBA is not conservative at awake/asleep block
    faces, as in the reference; conserved DBA restores the totals)."""
    n = 2048
    p = configs.kh_params(n=n, tile=16, mode=abi.MODE_REGULAR)
    U = configs.kh_ic(n)
    inst = ctx.create(p)
    for f in range(4):
        inst.upload(f, U[f])
    inst.step(100, graph_chunk=50)
    for f in (0, 1, 3):
        V = inst.download(f)
        assert abs(V.sum() - U[f].sum()) <= 1e-11 * np.abs(U[f]).sum(), f
    assert inst.check_finite()
    info = inst.info()
    assert 0 < info.sim_time < 100 * 0.4 / n
    inst.destroy()
