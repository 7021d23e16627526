"""Writes the golden fixtures of the reference models (tests/golden/*.npz).

Runs the reference CPU_version step functions, compiled from the reference
sources into oracle/_ref by oracle/Makefile.ref (this container only), on
small seeded inputs, and stores inputs and outputs. The fixtures are data;
tests/test_golden.py checks the oracle against them everywhere.

    python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from dba_amd import configs  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
D = C.POINTER(C.c_double)
F = C.POINTER(C.c_float)


def ptr(a, t=D):
    return a.ctypes.data_as(t)


# Model B parameters of example_Reactive_ModelB/Input.txt (dt, dx as there)
MODELB = dict(dx=1.0, dt=0.0005, chi_xy=0.0, chi_xz=0.0, chi_yz=0.0, chi_xb=0.0, chi_yb=0.0, chi_zb=4.0,
              v_x=1, v_y=1, v_z=2, r_x=1, r_y=1, r_z=2, n_exp=1, m_exp=1,
              epsx_sq=4.0, epsy_sq=4.0, epsz_sq=4.0, epsphi_sq=0.01,
              mob_x=1.0, mob_y=1.0, mob_z=1.0, mob_phi=1.0, k_eq=100.0, z_crit=0.4, p_gel=1.0, k_0=1.0,
              x0=0.2, y0=0.2, z0=0.01)
# Dendrite parameters of example_Dendritic_growth/Input.txt
DENDRITE = dict(dx=0.02, dt=0.00005, delta=0.5, tau=0.0003, epsilon=0.01, alpha=0.9, gamma=40.0, Delta=0.25)


def kh_case():
    lib = C.CDLL(os.path.join(REF, "libref_kh.so"))
    lib.ref_kh_run.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, D, D, D, D]
    n, steps, dt = 32, 40, 2e-4
    dx = 1.0 / n
    U = [np.ascontiguousarray(a) for a in configs.kh_reference_ic(n, dx=dx)]
    V = [a.copy() for a in U]
    lib.ref_kh_run(n, n, dx, dt, steps, *[ptr(a) for a in V])
    np.savez_compressed(os.path.join(HERE, "kh_reference.npz"), n=n, dx=dx, dt=dt, steps=steps,
                        U_in=np.stack(U), U_out=np.stack(V))


def modelb_case():
    lib = C.CDLL(os.path.join(REF, "libref_modelb.so"))
    lib.ref_modelb_run.argtypes = [C.c_int, C.c_int, C.c_int, D, C.c_int, D, D, D, D]
    n, steps = 8, 25
    m = MODELB
    prm = np.array([m["dx"], m["dt"], m["chi_xy"], m["chi_xz"], m["chi_yz"], m["chi_xb"], m["chi_yb"], m["chi_zb"],
                    m["v_x"], m["v_y"], m["v_z"], m["r_x"], m["r_y"], m["r_z"], m["n_exp"], m["m_exp"],
                    m["epsx_sq"], m["epsy_sq"], m["epsz_sq"], m["epsphi_sq"],
                    m["mob_x"], m["mob_y"], m["mob_z"], m["mob_phi"], m["k_eq"], m["z_crit"], m["p_gel"], m["k_0"]],
                   dtype=np.float64)
    X, Y, Z = configs.modelb_ic(n, n, n, m["x0"], m["y0"], m["z0"])
    phi = np.zeros(n ** 3)
    x, y, z = X.reshape(-1).copy(), Y.reshape(-1).copy(), Z.reshape(-1).copy()
    lib.ref_modelb_run(n, n, n, ptr(prm), steps, ptr(x), ptr(y), ptr(z), ptr(phi))
    np.savez_compressed(os.path.join(HERE, "modelb_reference.npz"), n=n, steps=steps, prm=prm,
                        X_in=np.stack([X, Y, Z]), X_out=np.stack([x, y, z]).reshape(3, n, n, n))


def dendrite_case():
    lib = C.CDLL(os.path.join(REF, "libref_dendrite.so"))
    lib.ref_dendrite_init.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, F, F]
    lib.ref_dendrite_run.argtypes = [C.c_int, C.c_int, C.c_int, F, C.c_int, F, F]
    n, steps = 32, 12
    d = DENDRITE
    size = (n + 2) ** 3
    p = np.zeros(size, np.float32)
    T = np.zeros(size, np.float32)
    lib.ref_dendrite_init(n, n, n, d["Delta"], ptr(p, F), ptr(T, F))
    p0, T0 = p.copy(), T.copy()
    prm = np.array([d["dx"], d["dt"], d["delta"], d["tau"], d["epsilon"], d["alpha"], d["gamma"]], np.float32)
    lib.ref_dendrite_run(n, n, n, ptr(prm, F), steps, ptr(p, F), ptr(T, F))
    shp = (n + 2, n + 2, n + 2)
    np.savez_compressed(os.path.join(HERE, "dendrite_reference.npz"), n=n, steps=steps, prm=prm,
                        p_in=p0.reshape(shp), T_in=T0.reshape(shp), p_out=p.reshape(shp), T_out=T.reshape(shp))


if __name__ == "__main__":
    kh_case()
    modelb_case()
    dendrite_case()
    print("fixtures written to", HERE)
