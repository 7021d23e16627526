"""Pins of the oracle's reactive Cahn-Hilliard step by closed forms.

dc/dt = M lap(c^3 - c - kappa lap c) - Gamma (c - cbar) on a periodic grid.
- a single Fourier mode of small amplitude around cbar grows by exactly
  G(k) = 1 + dt [M lam(k) (3 cbar^2 - 1 - kappa lam(k)) - Gamma], where lam is
  the stencil's symbol (5-point: -(4/dx^2)(sin^2(kx dx/2) + sin^2(ky dx/2)),
  9-point isotropic: (4(2cos kx + 2cos ky) + 4 cos kx cos ky - 20)/(6 dx^2));
- the mean obeys mean' - cbar = (1 - Gamma dt)(mean - cbar) (the Laplacians sum
  to zero on a periodic grid);
- uniform c = cbar is a fixed point;
- conserved DBA keeps the total exactly (example_Reactive_ModelB/Conserved/kernel.cu:526-553);
- DBA with a detector that always fires equals regular mode bit for bit, and a
  detector that never fires freezes the field after the first step.
"""
import numpy as np
import pytest

from dba_amd import abi, configs
from oracle.oracle import OracleSim


def symbol(stencil, kx, ky, dx):
    if stencil == 5:
        return -(4.0 / dx**2) * (np.sin(kx * dx / 2) ** 2 + np.sin(ky * dx / 2) ** 2)
    return (4 * (2 * np.cos(kx * dx) + 2 * np.cos(ky * dx)) + 4 * np.cos(kx * dx) * np.cos(ky * dx) - 20) / (
        6 * dx**2)


@pytest.mark.parametrize("stencil", [5, 9])
@pytest.mark.parametrize("mx,my", [(1, 0), (3, 2), (5, 7)])
@pytest.mark.parametrize("cbar", [0.0, 0.15])
def test_single_mode_growth_factor(stencil, mx, my, cbar):
    n = 32
    p = configs.ch_params(n=n, tile=16, mode=abi.MODE_REGULAR, stencil=stencil, react_gamma=0.02, cbar=cbar)
    p.kappa = 0.7
    p.mobility = 1.3
    p.dx = 0.9
    p.dt = 0.003
    y, x = np.mgrid[0:n, 0:n]
    kx, ky = 2 * np.pi * mx / (n * p.dx), 2 * np.pi * my / (n * p.dx)
    mode = np.cos(kx * x * p.dx + ky * y * p.dx)
    eps = 1e-7
    o = OracleSim(p)
    o.set_field(0, cbar + eps * mode)
    o.step(1)
    c1 = o.get_field(0)
    lam = symbol(stencil, kx, ky, p.dx)
    G = 1 + p.dt * (p.mobility * lam * (3 * cbar**2 - 1 - p.kappa * lam) - p.react_gamma)
    # the cubic term contributes O(eps^2) relative (with cbar != 0, 3 cbar eps^2 mode^2)
    np.testing.assert_allclose(c1 - cbar, G * eps * mode, rtol=0, atol=eps * 1e-5)


def test_mean_law():
    n = 64
    p = configs.ch_params(n=n, tile=16, mode=abi.MODE_REGULAR, react_gamma=0.05, cbar=0.1)
    c = configs.ch_ic(n, n, cbar=0.3, amp=0.05)
    o = OracleSim(p)
    o.set_field(0, c)
    m = c.mean()
    for _ in range(40):
        o.step(1)
        m = p.cbar + (1 - p.react_gamma * p.dt) * (m - p.cbar)
        assert abs(o.get_field(0).mean() - m) < 1e-13


def test_uniform_fixed_point():
    p = configs.ch_params(n=32, tile=16, mode=abi.MODE_DBA, cbar=0.2)
    o = OracleSim(p)
    o.set_field(0, np.full((32, 32), 0.2))
    o.step(5)
    assert np.all(o.get_field(0) == 0.2)


def test_conserved_total_invariant():
    n = 128
    p = configs.ch_params(n=n, tile=16, mode=abi.MODE_CONSERVED, react_gamma=0.05, cbar=0.2, eps=1e-3)
    c = configs.ch_ic(n, n, amp=0.05)
    c[:, 16:] = 0.0  # noise in one tile column only, so far tiles sleep
    o = OracleSim(p)
    o.set_field(0, c)
    o.step(60)
    assert abs(o.get_field(0).sum() - c.sum()) < 1e-11
    # some tiles fell asleep, so the run really exercised DBA
    assert o.awake(0).sum() < o.ntiles


@pytest.mark.parametrize("stencil", [5, 9])
def test_dba_always_awake_equals_regular(stencil):
    n = 64
    c = configs.ch_ic(n, n)
    pr = configs.ch_params(n=n, tile=16, mode=abi.MODE_REGULAR, stencil=stencil)
    pd = configs.ch_params(n=n, tile=16, mode=abi.MODE_DBA, stencil=stencil, eps=-1.0)
    a, b = OracleSim(pr), OracleSim(pd)
    a.set_field(0, c)
    b.set_field(0, c)
    a.step(30)
    b.step(30)
    assert np.array_equal(a.get_field(0), b.get_field(0))


def test_dba_never_firing_freezes():
    n = 64
    c = configs.ch_ic(n, n)
    p = configs.ch_params(n=n, tile=16, mode=abi.MODE_DBA, eps=1e9, history_cap=10)
    o = OracleSim(p)
    o.set_field(0, c)
    o.step(1)
    c1 = o.get_field(0)
    o.step(5)
    assert np.array_equal(o.get_field(0), c1)
    h = o.history()
    assert h[0, 0] == o.ntiles and np.all(h[1:, 0] == 0)
