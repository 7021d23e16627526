"""GPU: diagnostics computed on the device (dba_field_sum, dba_growth_front,
dba_kh_vy_mode) against the same quantities computed on the host from the
downloaded fields.

Bar: growth front bit-exact (same expression, same operands); sums within
1e-12 of the field's absolute sum (different summation order); the vy-mode
amplitude within 1e-9 relative (device sin/cos/exp vs the host's libm)."""
import numpy as np
import pytest

from dba_amd import abi, configs

pytestmark = pytest.mark.gpu


def growth_front(phi, ix):
    """growth_front of example_Dendritic_growth/kernel.cu:760-774 (0 when not found)."""
    nz, ny, nx = phi.shape
    col = phi[nz // 2, :, ix]
    with np.errstate(divide="ignore", invalid="ignore"):   # a flat column gives -inf, as on the device
        for y in range(ny // 2, ny):
            if col[y] < 0.5:
                return (col[y - 1] - 0.5) / (col[y - 1] - col[y]) + y - 1 - ny // 2
    return 0.0


@pytest.mark.parametrize("model", ["k93", "k94"])
def test_growth_front_and_sums(ctx, model):
    if model == "k93":
        n, nz = 128, 1
        p = configs.dendrite_k93_params(n=n, tile=16)
        phi0, T0 = configs.dendrite_k93_ic(n, n)
        steps = 400
    else:
        n, nz = 32, 32
        p = configs.dendrite_k94_params(n=n, nz=nz)
        phi0, T0 = configs.dendrite_k94_ic(n, n, nz)
        steps = 60
    inst = ctx.create(p)
    inst.upload(0, phi0)
    inst.upload(1, T0)
    inst.step(steps, graph_chunk=50)
    phi = inst.download(0).reshape(nz, n, n)
    xs = list(range(n // 2 - 10, n // 2 + 11)) + [0, n - 1]
    got = inst.growth_front(xs)
    want = np.array([growth_front(phi, x) for x in xs])
    assert np.array_equal(got, want)
    assert got[10] > 0.0     # the centre column has a front
    for f in (0, 1):
        a = inst.download(f)
        assert abs(inst.field_sum(f) - a.sum()) <= 1e-12 * max(1.0, np.abs(a).sum())
    with pytest.raises(abi.DbaError):
        inst.growth_front([n])   # outside the grid
    with pytest.raises(abi.DbaError):
        inst.kh_vy_mode()        # not a KH instance
    inst.destroy()


def test_conserved_sum_is_invariant(ctx):
    n = 128
    p = configs.ch_params(n=n, tile=32, mode=abi.MODE_CONSERVED, eps=0.05)
    c0 = configs.ch_ic(n, n, cbar=0.1)
    inst = ctx.create(p)
    inst.upload(0, c0)
    s0 = inst.field_sum(0)
    inst.step(200, graph_chunk=100)
    s1 = inst.field_sum(0)
    assert abs(s1 - s0) <= 1e-10 * np.abs(c0).sum()
    assert abs(s1 - inst.download(0).sum()) <= 1e-12 * np.abs(c0).sum()
    inst.destroy()


def test_kh_vy_mode(ctx):
    n = 256
    p = configs.kh_params(n=n, tile=16, mode=abi.MODE_DBA, eps=20.0)
    U = configs.kh_ic(n)
    inst = ctx.create(p)
    for f in range(4):
        inst.upload(f, U[f])
    inst.step(50, graph_chunk=25)
    Uc = [inst.download(f) for f in range(4)]
    rho, vx, vy, pr = configs.kh_primitive(Uc, 1.4, 1.0 / n / n)
    xs = np.arange(n)[None, :] / n
    ys = np.arange(n)[:, None] / n * np.ones((1, n))
    w = np.exp(-4.0 * np.pi * np.abs(ys - np.where(ys < 0.5, 0.25, 0.75)))
    ss, sc, sd = (vy * np.sin(4 * np.pi * xs) * w).sum(), (vy * np.cos(4 * np.pi * xs) * w).sum(), w.sum()
    want = 2.0 * np.sqrt(ss * ss / (sd * sd) + sc * sc / (sd * sd))
    got = inst.kh_vy_mode()
    assert got == pytest.approx(want, rel=1e-9)
    assert abs(inst.field_sum(abi.F_MASS) - Uc[0].sum()) <= 1e-12 * Uc[0].sum()
    inst.destroy()
