"""The oracle's reference-model paths against golden vectors of the compiled
reference CPU_version programs (tests/golden/*.npz, written by
tests/golden/make_golden.py from oracle/_ref).

- KH (example_Kelvin_Helmholtz/CPU_version/KH_CPU.cpp): gamma 5/3, periodic,
  fixed dt, 40 steps on 32^2: bit-exact (same FP64 operations).
- Model B (example_Reactive_ModelB/CPU_version/base_CPU.cpp): ternary
  Flory-Huggins + reaction, 27-point Laplacian, 25 steps on 8^3: bit-exact.
- Dendrite (example_Dendritic_growth/CPU_version/Base.cpp): the reference CPU
  program is float32, and its ghost frame is not a boundary condition: each
  step copies the whole padded p_new/T_new (Base.cpp:297-298), whose frame
  Evolve_p_and_T never writes, so the frame becomes freshly allocated zeros.
  The FP64 periodic oracle of the GPU program's equations
  (example_Dendritic_growth/kernel.cu:181-293), started from the same float32
  state, agrees to float32 resolution over 12 steps on 32^3 in the central
  12^3 block, which the frame cannot influence in 12 explicit steps (each step
  reaches one cell further; the influence at distance 10 is < 1e-7).
When oracle/_ref is built (this container), the same comparisons also run
directly against the compiled reference on a second case.
"""
import ctypes as C
import os

import numpy as np
import pytest

from dba_amd import abi, configs
from oracle.oracle import OracleSim

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref")


def oracle_kh(n, dx, dt, steps, U):
    p = configs.kh_reference_params(n=n, tile=8, mode=abi.MODE_REGULAR, dx=dx, dt=dt)
    o = OracleSim(p)
    for f in range(4):
        o.set_field(f, U[f])
    o.step(steps)
    return np.stack([o.get_field(f) for f in range(4)])


def test_kh_golden():
    g = np.load(os.path.join(GOLD, "kh_reference.npz"))
    out = oracle_kh(int(g["n"]), float(g["dx"]), float(g["dt"]), int(g["steps"]), g["U_in"])
    assert np.array_equal(out, g["U_out"])


def modelb_params_from(prm, n):
    names = ["dx", "dt", "chi_xy", "chi_xz", "chi_yz", "chi_xb", "chi_yb", "chi_zb", "v_x", "v_y", "v_z", "r_x",
             "r_y", "r_z", "n_exp", "m_exp", "epsx_sq", "epsy_sq", "epsz_sq", "epsphi_sq", "mob_x", "mob_y",
             "mob_z", "mob_phi", "k_eq", "z_crit", "p_gel", "k_0"]
    p = configs.modelb_params(n=n, tile=4, mode=abi.MODE_REGULAR)
    for k, v in zip(names, prm):
        setattr(p, k, float(v))
    p.dy = p.dz = p.dx
    return p


def oracle_modelb(p, X, steps):
    o = OracleSim(p)
    for f in range(3):
        o.set_field(f, X[f])
    o.set_field(3, np.zeros_like(X[0]))
    o.step(steps)
    return np.stack([o.get_field(f) for f in range(3)])


def test_modelb_golden():
    g = np.load(os.path.join(GOLD, "modelb_reference.npz"))
    n = int(g["n"])
    out = oracle_modelb(modelb_params_from(g["prm"], n), g["X_in"], int(g["steps"]))
    assert np.array_equal(out, g["X_out"])


def oracle_dendrite_from_ref(n, prm, p_in, T_in, steps):
    p = configs.dendrite_k94_params(n=n, mode=abi.MODE_REGULAR)
    p.dx = p.dy = p.dz = float(prm[0])
    p.dt = float(prm[1])
    p.delta, p.tau, p.eps_bar, p.alpha, p.gamma = (float(v) for v in prm[2:7])
    o = OracleSim(p)
    o.set_field(0, p_in[1:-1, 1:-1, 1:-1].astype(np.float64))
    o.set_field(1, T_in[1:-1, 1:-1, 1:-1].astype(np.float64))
    o.step(steps)
    return o.get_field(0), o.get_field(1)


def test_dendrite_golden():
    g = np.load(os.path.join(GOLD, "dendrite_reference.npz"))
    n = int(g["n"])
    p, T = oracle_dendrite_from_ref(n, g["prm"], g["p_in"], g["T_in"], int(g["steps"]))
    pr, Tr = g["p_out"][1:-1, 1:-1, 1:-1], g["T_out"][1:-1, 1:-1, 1:-1]
    c = slice(10, n - 10)
    assert np.max(np.abs(p[c, c, c] - pr[c, c, c])) < 2e-6
    assert np.max(np.abs(T[c, c, c] - Tr[c, c, c])) < 2e-6
    assert np.max(np.abs(p[c, c, c] - g["p_in"][1:-1, 1:-1, 1:-1][c, c, c])) > 1e-3  # the seed evolved
    assert np.max(np.abs(T[c, c, c] - g["T_in"][1:-1, 1:-1, 1:-1][c, c, c])) > 1e-4


needs_ref = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "libref_kh.so")),
                               reason="oracle/_ref not built (reference sources not mounted)")


@needs_ref
def test_kh_against_compiled_reference_second_case():
    lib = C.CDLL(os.path.join(REF, "libref_kh.so"))
    D = C.POINTER(C.c_double)
    lib.ref_kh_run.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, D, D, D, D]
    n, dx, dt, steps = 48, 0.002, 2e-4, 30
    U = [np.ascontiguousarray(a) for a in configs.kh_reference_ic(n, dx=dx)]
    V = [a.copy() for a in U]
    lib.ref_kh_run(n, n, dx, dt, steps, *[a.ctypes.data_as(D) for a in V])
    assert np.array_equal(oracle_kh(n, dx, dt, steps, U), np.stack(V))


@needs_ref
def test_modelb_against_compiled_reference_2d_plane():
    lib = C.CDLL(os.path.join(REF, "libref_modelb.so"))
    D = C.POINTER(C.c_double)
    lib.ref_modelb_run.argtypes = [C.c_int, C.c_int, C.c_int, D, C.c_int, D, D, D, D]
    g = np.load(os.path.join(GOLD, "modelb_reference.npz"))
    prm = g["prm"].copy()
    n, nz, steps = 16, 1, 20
    X = configs.modelb_ic(n, n, nz, seed=99)
    x, y, z = (a.reshape(-1).copy() for a in X)
    phi = np.zeros(n * n)
    lib.ref_modelb_run(n, n, nz, prm.ctypes.data_as(D), steps, x.ctypes.data_as(D), y.ctypes.data_as(D),
                       z.ctypes.data_as(D), phi.ctypes.data_as(D))
    p = modelb_params_from(prm, n)
    p.nz = 1
    p.tile_z = 1
    out = oracle_modelb(p, np.stack(X), steps)
    assert np.array_equal(out.reshape(3, -1), np.stack([x, y, z]))
