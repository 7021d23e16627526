"""CPU: the FP64 logarithm of dba_amd/csrc/mathf64.h (used by the Model B
chemical potentials) against the exact logarithm.

The header is plain C++ outside HIP, so the same source is compiled here with
g++ (-ffp-contract=off, explicit fma like the device code) and evaluated on
random and edge-case inputs; the reference is numpy's extended-precision log
(64-bit mantissa), so the measured error is that of log_f64 itself.
Bar: below 1 ulp everywhere; IEEE special values exact."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

WRAP = r"""
#include "mathf64.h"
extern "C" void t_log(const double* x, double* y, long n) {
    for (long i = 0; i < n; i++) y[i] = dba::log_f64(x[i]);
}
"""


@pytest.fixture(scope="module")
def log_f64(tmp_path_factory):
    d = tmp_path_factory.mktemp("mathf64")
    src = d / "w.cpp"
    src.write_text(WRAP)
    so = d / "w.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-I",
                    os.path.join(ROOT, "dba_amd", "csrc"), "-o", str(so), str(src)], check=True)
    lib = ctypes.CDLL(str(so))
    P = ctypes.POINTER(ctypes.c_double)

    def f(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        lib.t_log(x.ctypes.data_as(P), y.ctypes.data_as(P), ctypes.c_long(x.size))
        return y
    return f


def ulp_error(x, y):
    exact = np.log(x.astype(np.longdouble))
    ulp = np.spacing(np.abs(exact.astype(np.float64))).astype(np.longdouble)
    return np.abs((y.astype(np.longdouble) - exact) / ulp).astype(np.float64)


@pytest.mark.skipif(np.finfo(np.longdouble).nmant < 60, reason="needs x87 extended precision for the reference")
def test_log_within_one_ulp(log_f64):
    rng = np.random.default_rng(7)
    x = np.concatenate([
        rng.uniform(0.0, 1.0, 400_000),                 # the concentrations' range
        np.exp(rng.uniform(-700.0, 700.0, 200_000)),    # every binade
        rng.uniform(0.70, 1.42, 200_000),               # around the reduction boundary
        1.0 + rng.uniform(-1e-8, 1e-8, 20_000),         # log near 0
        np.array([5e-324, 1e-310, 2.2250738585072014e-308, 0.5, 2.0, np.nextafter(1.0, 0.0),
                  np.nextafter(1.0, 2.0), float.fromhex("0x1.6a09e667f3bcdp-1"),
                  float.fromhex("0x1.6a09e667f3bccp-1"), 1.7976931348623157e308]),
    ])
    x = x[x > 0.0]
    err = ulp_error(x, log_f64(x))
    assert err.max() < 1.0, (err.max(), x[err.argmax()])
    assert log_f64(np.array([1.0]))[0] == 0.0


def test_log_special_values(log_f64):
    y = log_f64(np.array([0.0, -0.0, -1.0, -np.inf, np.inf, np.nan]))
    assert y[0] == -np.inf and y[1] == -np.inf
    assert np.isnan(y[2]) and np.isnan(y[3]) and np.isnan(y[5])
    assert y[4] == np.inf
