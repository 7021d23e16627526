"""Pins of the oracle's compressible-Euler (KH) step.

- a uniform state is a fixed point bit for bit (all face fluxes equal);
- periodic boxes conserve mass, both momenta and energy to round-off (the
  flux differences telescope); with reflective y walls mass, x-momentum and
  energy stay conserved (the wall flux of those vanishes exactly) while
  y-momentum does not (wall pressure);
- a problem that varies only in x stays y-independent bit for bit, keeps
  vy = 0, and its transpose (x <-> y, Momx <-> Momy) gives the transposed
  result to round-off;
- a small-amplitude right-moving sound wave rho = 1 + e sin(kx),
  u = c e sin(kx), p = 1 + gamma e sin(kx) is carried at c = sqrt(gamma) with
  second-order convergence of the error (pins gamma, the Rusanov flux and the
  half-step predictor);
- the CFL step of a uniform state is cfl dx / (sqrt(gamma p/rho) + |v|);
- DBA with a detector that always fires equals regular mode bit for bit.
"""
import numpy as np
import pytest

from dba_amd import abi, configs
from oracle.oracle import OracleSim


def run(p, U, n):
    o = OracleSim(p)
    for f in range(4):
        o.set_field(f, U[f])
    o.step(n)
    return o, [o.get_field(f) for f in range(4)]


def params(n, walls=False, mode=abi.MODE_REGULAR, dt_mode=abi.DT_FIXED, dt=1e-3, gamma=1.4, tile=8):
    p = configs.kh_params(n=n, tile=tile, mode=mode, walls=walls, dt_mode=dt_mode, dt=dt, gas_gamma=gamma)
    return p


def test_uniform_fixed_point():
    n = 32
    vol = 1.0 / n / n
    U = configs.kh_conserved(np.full((n, n), 1.3), np.full((n, n), 0.4), np.full((n, n), -0.2),
                             np.full((n, n), 2.0), 1.4, vol)
    for walls in (False,):
        _, V = run(params(n, walls), U, 10)
        for f in range(4):
            assert np.array_equal(V[f], U[f])


@pytest.mark.parametrize("walls", [False, True])
def test_conservation(walls):
    n = 64
    U = configs.kh_ic(n)
    _, V = run(params(n, walls, dt=2e-4), U, 100)
    keep = [0, 1, 3] if walls else [0, 1, 2, 3]
    for f in keep:
        assert abs(V[f].sum() - U[f].sum()) <= 1e-13 * max(1.0, abs(U[f].sum())) * 10, f
    if walls:
        # a pressure ramp in y: the walls push, so y-momentum changes
        y = (np.arange(n) + 0.5) / n
        Y = np.tile(y[:, None], (1, n))
        Up = configs.kh_conserved(np.ones((n, n)), np.zeros((n, n)), np.zeros((n, n)), 2.0 + Y, 1.4, 1.0 / n / n)
        _, Vp = run(params(n, walls, dt=2e-4), Up, 20)
        assert abs(Vp[0].sum() - Up[0].sum()) <= 1e-14
        assert Vp[2].sum() < -1e-6 * (1.0 / n / n)  # net push towards lower pressure (down)


def test_x_only_problem_is_y_independent_and_transposes():
    n = 64
    x = (np.arange(n) + 0.5) / n
    rho = np.tile(1.0 + 0.5 * (np.abs(x - 0.5) < 0.2), (n, 1))
    pr = np.tile(1.0 + 1.0 * (np.abs(x - 0.5) < 0.2), (n, 1))
    vx = np.tile(0.1 * np.sin(2 * np.pi * x), (n, 1))
    zero = np.zeros((n, n))
    vol = 1.0 / n / n
    U = configs.kh_conserved(rho, vx, zero, pr, 1.4, vol)
    _, V = run(params(n), U, 60)
    for f in range(4):
        assert np.array_equal(V[f], np.tile(V[f][0], (n, 1)))
    assert np.all(V[2] == 0.0)
    UT = configs.kh_conserved(rho.T, zero, vx.T, pr.T, 1.4, vol)
    _, W = run(params(n), UT, 60)
    np.testing.assert_allclose(W[0], V[0].T, rtol=0, atol=1e-13 * vol * 100)
    np.testing.assert_allclose(W[2], V[1].T, rtol=0, atol=1e-13 * vol * 100)
    np.testing.assert_allclose(W[3], V[3].T, rtol=0, atol=1e-13 * vol * 100)


def acoustic_error(n, gamma=1.4, e=1e-6, t_end=0.25):
    k = 2 * np.pi
    c = np.sqrt(gamma)
    x = (np.arange(n) + 0.5) / n
    X = np.tile(x, (8, 1))
    def prim(t):
        s = np.sin(k * (X - c * t))
        return 1 + e * s, c * e * s, np.zeros_like(X), 1 + gamma * e * s
    dt = 0.2 / n / (c + c * e)
    steps = int(round(t_end / dt))
    dt = t_end / steps
    p = configs.kh_params(n=n, ny=8, tile=8, mode=abi.MODE_REGULAR, walls=False, dt_mode=abi.DT_FIXED, dt=dt,
                          gas_gamma=gamma)
    vol = (1.0 / n) * (1.0 / n)
    U = configs.kh_conserved(*prim(0.0), gamma, vol)
    _, V = run(p, U, steps)
    r, u, _, pr = configs.kh_primitive(V, gamma, vol)
    r_ex = prim(t_end)[0]
    return np.max(np.abs(r - r_ex)) / e


def test_acoustic_wave_second_order():
    e1, e2, e3 = acoustic_error(32), acoustic_error(64), acoustic_error(128)
    assert e3 < 0.01, (e1, e2, e3)
    assert e1 / e2 > 3.0 and e2 / e3 > 3.0, (e1, e2, e3)


def test_cfl_time_step():
    n = 32
    g, r0, u0, v0, p0 = 1.4, 1.3, 0.4, -0.3, 2.0
    U = configs.kh_conserved(np.full((n, n), r0), np.full((n, n), u0), np.full((n, n), v0), np.full((n, n), p0), g,
                             1.0 / n / n)
    p = params(n, walls=False, dt_mode=abi.DT_CFL)
    o, _ = run(p, U, 1)
    expect = p.cfl * ((1.0 / n) / (np.sqrt(g * p0 / r0) + np.sqrt(u0 * u0 + v0 * v0)))
    assert abs(o.time() - expect) <= 1e-15 * expect


@pytest.mark.parametrize("walls", [False, True])
def test_dba_always_awake_equals_regular(walls):
    n = 64
    U = configs.kh_ic(n)
    pr = params(n, walls, dt=2e-4)
    pd = params(n, walls, mode=abi.MODE_DBA, dt=2e-4)
    pd.thresh_a = -1.0
    _, A = run(pr, U, 30)
    _, B = run(pd, U, 30)
    for f in range(4):
        assert np.array_equal(A[f], B[f])


def test_dba_conserved_totals_and_sleeping():
    n = 128
    U = configs.kh_ic(n)
    p = params(n, walls=False, mode=abi.MODE_CONSERVED, dt=2e-4)
    p.thresh_a = 5.0
    o, V = run(p, U, 40)
    for f in range(4):
        assert abs(V[f].sum() - U[f].sum()) <= 1e-12 * max(1.0, abs(U[f].sum()))
    assert o.awake(0).sum() < o.ntiles


@pytest.mark.parametrize("walls,dt_mode", [(True, abi.DT_CFL), (False, abi.DT_FIXED)])
def test_threaded_bench_equals_serial(walls, dt_mode):
    """The CPU baseline's y-slab threads run the same operations as the serial step."""
    p = configs.kh_params(n=96, tile=16, mode=abi.MODE_REGULAR, walls=walls, dt_mode=dt_mode)
    U = configs.kh_ic(96)
    a, b = OracleSim(p), OracleSim(p)
    for f in range(4):
        a.set_field(f, U[f])
        b.set_field(f, U[f])
    a.bench(6, 1)
    b.bench(6, 5)
    for f in range(4):
        assert np.array_equal(a.get_field(f), b.get_field(f))
