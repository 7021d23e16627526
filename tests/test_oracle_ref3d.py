"""DBA semantics of the oracle's reference 3D models (the regular-mode
arithmetic is pinned by tests/test_golden.py against the compiled reference).

- a detector that always fires gives regular mode bit for bit (for the
  dendrite, where the DBA T update adds the latent heat in a separate rounding
  step, to round-off);
- conserved Model B keeps the totals of x, y and z;
- uniform states are fixed points;
- far-field blocks of a dendrite seed fall asleep.
"""
import numpy as np
import pytest

from dba_amd import abi, configs
from oracle.oracle import OracleSim


def run(p, fields, n):
    o = OracleSim(p)
    for f, a in enumerate(fields):
        o.set_field(f, a)
    o.step(n)
    return o, [o.get_field(f) for f in range(len(fields))]


@pytest.mark.parametrize("nz", [1, 8])
def test_modelb_always_awake_equals_regular(nz):
    n = 16
    X = list(configs.modelb_ic(n, n, nz)) + [np.zeros((nz, n, n))]
    pr = configs.modelb_params(n=n, nz=nz, mode=abi.MODE_REGULAR)
    pd = configs.modelb_params(n=n, nz=nz, mode=abi.MODE_DBA, eps=-1.0)
    _, A = run(pr, [a.reshape(-1) for a in X], 15)
    _, B = run(pd, [a.reshape(-1) for a in X], 15)
    for f in range(3):
        assert np.array_equal(A[f], B[f])


def test_modelb_conserved_totals():
    n = 32
    X = list(configs.modelb_ic(n, n, 8)) + [np.zeros((8, n, n))]
    X[0][:, :, 8:] = 0.2  # quiet three quarters so blocks sleep
    X[1][:, :, 8:] = 0.2
    X[2][:, :, 8:] = 0.01
    p = configs.modelb_params(n=n, nz=8, mode=abi.MODE_CONSERVED, eps=1e-4)
    o, V = run(p, [a.reshape(-1) for a in X], 20)
    for f in range(3):
        assert abs(V[f].sum() - X[f].sum()) < 1e-11
    assert o.awake(0).sum() < o.ntiles


def test_modelb_uniform_fixed_point():
    n = 8
    X = [np.full((n, n, n), 0.2), np.full((n, n, n), 0.2), np.full((n, n, n), 0.01), np.zeros((n, n, n))]
    p = configs.modelb_params(n=n, mode=abi.MODE_REGULAR)
    o, V = run(p, [a.reshape(-1) for a in X], 1)
    # mu is uniform, so only the (uniform) reaction acts
    for f in range(3):
        assert np.all(V[f] == V[f].reshape(-1)[0])


def test_k94_dba_close_to_regular_and_sleeps():
    n = 24
    p0, T0 = configs.dendrite_k94_ic(n, n, n)
    pr = configs.dendrite_k94_params(n=n, mode=abi.MODE_REGULAR)
    pd = configs.dendrite_k94_params(n=n, mode=abi.MODE_DBA, history_cap=40)
    pd.thresh_a = -1.0
    pd.thresh_b = -1.0
    _, A = run(pr, [p0, T0], 30)
    _, B = run(pd, [p0, T0], 30)
    assert np.array_equal(A[0], B[0]) or np.max(np.abs(A[0] - B[0])) < 1e-14
    assert np.max(np.abs(A[1] - B[1])) < 1e-14
    pd2 = configs.dendrite_k94_params(n=n, mode=abi.MODE_DBA, history_cap=40)
    o, _ = run(pd2, [p0, T0], 30)
    h = o.history()
    assert h[0, 0] == o.ntiles and h[-1, 0] < o.ntiles
