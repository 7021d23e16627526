"""Pins of the oracle's DBA data flow between awake and sleeping tiles.

The reference's DBA kernels (which cannot run here) read, from a sleeping
neighbour, whatever its global arrays last held:
  - the chemical potential stored when it last ran (example_Reactive_ModelB/kernel.cu:594-654);
  - its primitives, gradients, primes and face fluxes stored when it last ran
    (example_Kelvin_Helmholtz/KH_DBA.cu:747-923);
and they drop the latent heat of a p-awake / T-asleep block
(example_Dendritic_growth/kernel.cu:411-493, copy-back only where T is awake,
:496-574), and shift every cell, awake or not, by the mean change in conserved
mode (example_Reactive_ModelB/Conserved/kernel.cu:526-553,
example_Kelvin_Helmholtz/Conserved/KH_DBA.cu:1553-1575).

Each test prescribes the awake tiles of every step on a tiny grid (a tile that
ran one step and sleeps the next beside tiles that keep running), runs the oracle
and an independent statement of the reference's kernel sequence
(tests/dba_sequence.py, or a closed rule for the dendrites), and checks:
  1. the oracle matches it (tolerance 1e-12 of the field scale; bitwise where a
     value must not move at all);
  2. the schedule discriminates: the fresh-read answer differs by > 1e-6;
  3. the oracle built with the matching planted mistake (ORA_MUTANT in
     oracle/oracle.c) fails check 1.
"""
import os
import subprocess

import numpy as np
import pytest

from dba_amd import abi, configs
from oracle.oracle import OracleSim, load_path
from tests.dba_sequence import CHSeq, KHSeq, MBSeq, cell_mask, lap2d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_C = os.path.join(ROOT, "oracle", "oracle.c")
TOL = 1e-12


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    libs = {}
    for k in range(1, 6):
        out = str(d / f"liboracle_m{k}.so")
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", f"-DORA_MUTANT={k}", "-o", out,
                        ORACLE_C, "-lm", "-lpthread"], check=True)
        libs[k] = load_path(out)
    return libs


def schedule(nt, steps, seed, frac=0.5):
    """Step 1: every tile; then random sets that always hold a sleeping tile
    next to an awake one (tiles that ran one step and sleep the next)."""
    rng = np.random.default_rng(seed)
    out = [np.ones(nt, dtype=np.uint8)]
    while len(out) < steps:
        m = (rng.random(nt) < frac).astype(np.uint8)
        if 0 < m.sum() < nt:
            out.append(m)
    return out


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(1e-300, float(np.max(np.abs(b)))))


# ------------------------------------------------------------------ CH (stale mu)

def ch_case(stencil=9, mode=abi.MODE_DBA):
    p = configs.ch_params(n=32, tile=8, mode=mode, stencil=stencil, react_gamma=0.3, cbar=0.1)
    rng = np.random.default_rng(7)
    c0 = 0.1 + 0.3 * (2.0 * rng.random((32, 32)) - 1.0)
    return p, c0, schedule(16, 6, 11 + stencil)


def run_ch(lib, p, c0, sched, conserved):
    o = OracleSim(p, lib=lib)
    o.set_field(0, c0)
    seq = CHSeq(p, c0, conserved=conserved)
    fresh = CHSeq(p, c0, conserved=conserved, fresh=True)
    err = gap = shift_err = 0.0
    for m in sched:
        before = o.get_field(0).copy()
        o.set_awake(0, m)
        o.step(1)
        seq.step(m)
        fresh.step(m)
        got = o.get_field(0)
        err = max(err, rel_err(got, seq.c))
        gap = max(gap, rel_err(fresh.c, seq.c))
        if conserved:
            # every sleeping cell moves by exactly minus the mean change of the step
            asleep = ~cell_mask(m, got.shape, (p.tile_x, p.tile_y))
            if asleep.any():
                shift_err = max(shift_err, float(np.max(np.abs((got - before)[asleep] + seq.last_avg))))
    o.close()
    return err, gap, shift_err


@pytest.mark.parametrize("stencil", [5, 9])
def test_ch_stale_mu(stencil, mutants):
    p, c0, sched = ch_case(stencil)
    err, gap, _ = run_ch(None, p, c0, sched, False)
    assert err < TOL, err
    assert gap > 1e-6, gap
    err_m, _, _ = run_ch(mutants[1], p, c0, sched, False)
    assert err_m > 1e-6, f"fresh-mu oracle not caught ({err_m})"


def test_ch_conserved_shift_on_sleeping_cells(mutants):
    p, c0, sched = ch_case(9, abi.MODE_CONSERVED)
    err, gap, shift_err = run_ch(None, p, c0, sched, True)
    assert err < TOL, err
    assert gap > 1e-6, gap
    assert shift_err < 1e-15, shift_err
    for k in (1, 5):
        err_m, _, shift_m = run_ch(mutants[k], p, c0, sched, True)
        assert err_m > 1e-6 or shift_m > 1e-6, f"mutant {k} not caught ({err_m}, {shift_m})"


# --------------------------------------------------------- Model B (stale mu, 3D)

def mb_case(mode=abi.MODE_DBA):
    p = configs.modelb_params(n=8, mode=mode)
    x, y, z = configs.modelb_ic(8, 8, 8)
    phi = np.zeros_like(x)
    return p, (x, y, z, phi), schedule(8, 5, 23)


def run_mb(lib, p, ic, sched, conserved):
    o = OracleSim(p, lib=lib)
    for f, a in enumerate(ic):
        o.set_field(f, a)
    seq = MBSeq(p, *ic, conserved=conserved)
    fresh = MBSeq(p, *ic, conserved=conserved, fresh=True)
    err = gap = 0.0
    for m in sched:
        o.set_awake(0, m)
        o.step(1)
        seq.step(m)
        fresh.step(m)
        for k in range(3):
            err = max(err, rel_err(o.get_field(k), seq.X[k]))
            gap = max(gap, rel_err(fresh.X[k], seq.X[k]))
    o.close()
    return err, gap


@pytest.mark.parametrize("mode", [abi.MODE_DBA, abi.MODE_CONSERVED])
def test_modelb_stale_mu(mode, mutants):
    p, ic, sched = mb_case(mode)
    cons = mode == abi.MODE_CONSERVED
    err, gap = run_mb(None, p, ic, sched, cons)
    assert err < TOL, err
    assert gap > 1e-7, gap
    for k in ((1, 5) if cons else (1,)):
        err_m, _ = run_mb(mutants[k], p, ic, sched, cons)
        assert err_m > 1e-7, f"mutant {k} not caught ({err_m})"


# ------------------------------------------------ dendrites (dropped latent heat)

def neumann_lap5(f, dx):
    g = np.pad(f, 1, mode="symmetric")   # zero-flux mirror: f(-1) = f(0)
    return (g[1:-1, 2:] + g[1:-1, :-2] + g[2:, 1:-1] + g[:-2, 1:-1] - 4.0 * f) / (dx * dx)


def periodic_lap7(f, dx):
    s = sum(np.roll(f, k, axis=a) for a in range(3) for k in (1, -1))
    return (s - 6.0 * f) / (dx * dx)


def dendrite_schedule(nt, steps, seed):
    """(p set, T set) per step: every combination of p/T awake/asleep occurs."""
    rng = np.random.default_rng(seed)
    out = [(np.ones(nt, np.uint8), np.ones(nt, np.uint8))]
    while len(out) < steps:
        mp = (rng.random(nt) < 0.6).astype(np.uint8)
        mt = (rng.random(nt) < 0.5).astype(np.uint8)
        if (mp & (1 - mt)).any() and (mt & (1 - mp)).any() and (mp & mt).any():
            out.append((mp, mt))
    return out


def run_dendrite(lib, p, phi0, T0, sched, lap, latent):
    """Checks against the reference rule (example_Dendritic_growth/kernel.cu:411-574):
    p-asleep cells keep phi; T-asleep cells keep T bitwise (the latent heat of a
    p-awake tile is added to T_new, which is copied back only where T is awake);
    T-awake cells get T + dt lap T, plus latent (phi_new - phi) where p is awake."""
    o = OracleSim(p, lib=lib)
    o.set_field(0, phi0)
    o.set_field(1, T0)
    tile = (p.tile_x, p.tile_y, p.tile_z) if phi0.ndim == 3 else (p.tile_x, p.tile_y)
    frozen_T = frozen_phi = rule = 0.0
    moved = np.inf
    for mp, mt in sched:
        phi, T = o.get_field(0).copy(), o.get_field(1).copy()
        o.set_awake(0, mp)
        o.set_awake(1, mt)
        o.step(1)
        phi1, T1 = o.get_field(0), o.get_field(1)
        Ap, At = cell_mask(mp, phi.shape, tile), cell_mask(mt, phi.shape, tile)
        frozen_phi = max(frozen_phi, float(np.max(np.abs(phi1 - phi)[~Ap], initial=0.0)))
        frozen_T = max(frozen_T, float(np.max(np.abs(T1 - T)[~At], initial=0.0)))
        expect = T + p.dt * lap(T, p.dx) + np.where(Ap, latent * (phi1 - phi), 0.0)
        rule = max(rule, float(np.max(np.abs(T1 - expect)[At], initial=0.0)))
        if (Ap & ~At).any():
            # the dropped latent term is far above the tolerance
            moved = min(moved, float(np.max(np.abs(latent * (phi1 - phi))[Ap & ~At])))
    o.close()
    return frozen_phi, frozen_T, rule, moved


def k93_case():
    p = configs.dendrite_k93_params(n=32, tile=16)
    rng = np.random.default_rng(3)
    phi0 = 0.2 + 0.6 * rng.random((32, 32))
    T0 = 0.05 * (2.0 * rng.random((32, 32)) - 1.0)
    return p, phi0, T0, dendrite_schedule(4, 6, 5)


def k94_case():
    p = configs.dendrite_k94_params(n=8)
    rng = np.random.default_rng(4)
    phi0 = 0.2 + 0.6 * rng.random((8, 8, 8))
    T0 = -0.25 + 0.05 * rng.random((8, 8, 8))
    return p, phi0, T0, dendrite_schedule(8, 6, 9)


@pytest.mark.parametrize("model", ["k93", "k94"])
def test_dendrite_latent_heat_dropped(model, mutants):
    if model == "k93":
        p, phi0, T0, sched = k93_case()
        lap, latent = neumann_lap5, p.latent_k
    else:
        p, phi0, T0, sched = k94_case()
        lap, latent = periodic_lap7, 1.0
    frozen_phi, frozen_T, rule, moved = run_dendrite(None, p, phi0, T0, sched, lap, latent)
    assert frozen_phi == 0.0 and frozen_T == 0.0, (frozen_phi, frozen_T)
    assert rule < 1e-13, rule
    assert moved > 1e-6, moved
    _, frozen_T_m, _, _ = run_dendrite(mutants[2], p, phi0, T0, sched, lap, latent)
    assert frozen_T_m > 1e-6, f"latent-heat-kept oracle not caught ({frozen_T_m})"


# ------------------------------- KH (stale primitives, gradients, primes, fluxes)

def kh_case(mode=abi.MODE_DBA):
    n = 32
    p = configs.kh_reference_params(n=n, tile=8, mode=mode, dx=1.0 / n)
    rng = np.random.default_rng(17)
    rho = 1.0 + 0.5 * rng.random((n, n))
    vx = 0.3 * (2.0 * rng.random((n, n)) - 1.0)
    vy = 0.3 * (2.0 * rng.random((n, n)) - 1.0)
    pr = 1.0 + 0.5 * rng.random((n, n))
    cmax = float(np.max(np.sqrt(p.gas_gamma * pr / rho) + np.sqrt(vx * vx + vy * vy)))
    p.dt = 0.2 * p.dx / cmax
    U0 = configs.kh_conserved(rho, vx, vy, pr, p.gas_gamma, p.dx * p.dy)
    return p, U0, schedule(16, 6, 31)


def run_kh(lib, p, U0, sched, conserved):
    o = OracleSim(p, lib=lib)
    for f, a in enumerate(U0):
        o.set_field(f, a)
    seq = KHSeq(p, U0, conserved=conserved)
    fresh_p = KHSeq(p, U0, conserved=conserved, fresh_prims=True)
    fresh_f = KHSeq(p, U0, conserved=conserved, fresh_flux=True)
    err = gap_p = gap_f = 0.0
    for m in sched:
        o.set_awake(0, m)
        o.step(1)
        for s in (seq, fresh_p, fresh_f):
            s.step(m)
        for q in range(4):
            err = max(err, rel_err(o.get_field(q), seq.U[q]))
            gap_p = max(gap_p, rel_err(fresh_p.U[q], seq.U[q]))
            gap_f = max(gap_f, rel_err(fresh_f.U[q], seq.U[q]))
    o.close()
    return err, gap_p, gap_f


@pytest.mark.parametrize("mode", [abi.MODE_DBA, abi.MODE_CONSERVED])
def test_kh_stale_edge_state(mode, mutants):
    p, U0, sched = kh_case(mode)
    cons = mode == abi.MODE_CONSERVED
    err, gap_p, gap_f = run_kh(None, p, U0, sched, cons)
    assert err < TOL, err
    assert gap_p > 1e-6 and gap_f > 1e-6, (gap_p, gap_f)
    for k in ((3, 4, 5) if cons else (3, 4)):
        err_m, _, _ = run_kh(mutants[k], p, U0, sched, cons)
        assert err_m > 1e-6, f"mutant {k} not caught ({err_m})"


def test_sequences_reduce_to_regular_mode():
    """With every tile awake the kernel sequences are plain regular-mode steps, so
    the numpy statements can be checked against the oracle's regular mode, which
    is pinned bit-exactly to the compiled reference CPU programs (Model B, KH)
    and to closed forms (CH)."""
    p, c0, _ = ch_case(9, abi.MODE_REGULAR)
    o = OracleSim(p)
    o.set_field(0, c0)
    s = CHSeq(p, c0)
    for _ in range(3):
        o.step(1)
        s.step(np.ones(16, np.uint8))
    assert rel_err(o.get_field(0), s.c) < TOL
    p, U0, _ = kh_case(abi.MODE_REGULAR)
    o = OracleSim(p)
    for f, a in enumerate(U0):
        o.set_field(f, a)
    s = KHSeq(p, U0)
    for _ in range(3):
        o.step(1)
        s.step(np.ones(16, np.uint8))
    assert max(rel_err(o.get_field(q), s.U[q]) for q in range(4)) < TOL
    p, ic, _ = mb_case(abi.MODE_REGULAR)
    o = OracleSim(p)
    for f, a in enumerate(ic):
        o.set_field(f, a)
    s = MBSeq(p, *ic)
    for _ in range(3):
        o.step(1)
        s.step(np.ones(8, np.uint8))
    assert max(rel_err(o.get_field(k), s.X[k]) for k in range(3)) < TOL
    # the 9-point statement is the 27-point one on a single plane
    f = np.random.default_rng(0).random((6, 6))
    assert np.allclose(lap2d(f, 9, 1.0), __import__("tests.dba_sequence", fromlist=["lap27"]).lap27(
        np.stack([f, f, f]), 1.0)[1], rtol=0, atol=1e-13)


# ------------- the GPU parity cases would catch a kernel that reads fresh values

def _final(lib, p, fields, steps):
    o = OracleSim(p, lib=lib)
    for f, a in enumerate(fields):
        o.set_field(f, a)
    o.step(steps)
    out = [o.get_field(f).copy() for f in range(len(fields))]
    o.close()
    return out


def _gap(a, b):
    return max(float(np.max(np.abs(x - y))) for x, y in zip(a, b))


@pytest.mark.parametrize("case", ["ch", "ch_cons", "kh", "kh_cons", "k93", "k94", "modelb"])
def test_gpu_parity_cases_discriminate(case, mutants):
    """The workloads of tests/test_gpu_{ch,kh,k93,ref3d}.py (run here on the oracle
    only) end measurably apart under each planted mistake, beyond the tolerance
    those GPU tests hold the HIP path to (bit-exact for CH and KH, 1e-10 for the
    dendrite and Model B, 1e-12 for conserved modes). A HIP step that read fresh
    values, or kept the latent heat, or shifted only awake cells, would fail them."""
    if case in ("ch", "ch_cons"):
        n = 128
        mode = abi.MODE_DBA if case == "ch" else abi.MODE_CONSERVED
        p = configs.ch_params(n=n, tile=16, mode=mode, stencil=9, eps=1e-3, react_gamma=0.01, cbar=0.05)
        c = configs.ch_ic(n, n, seed=1234)
        c[:, n // 4:] = 0.0
        fields, steps, muts, bar = [c], 40, ((1, 5) if mode == abi.MODE_CONSERVED else (1,)), 1e-12
    elif case in ("kh", "kh_cons"):
        n = 128
        mode = abi.MODE_DBA if case == "kh" else abi.MODE_CONSERVED
        p = configs.kh_reference_params(n=n, tile=8, mode=mode, dx=1.0 / n, dt=2e-4, eps=8.0)
        fields, steps = configs.kh_reference_ic(n, dx=1.0 / n), 35
        muts, bar = ((3, 4, 5) if mode == abi.MODE_CONSERVED else (3, 4)), 1e-12 * 1e-4
    elif case == "k93":   # test_k93_latent_heat_dropped_matches_oracle
        p = configs.dendrite_k93_params(n=128, tile=16)
        p.thresh_b = 0.03
        fields, steps, muts, bar = list(configs.dendrite_k93_ic(128, 128)), 150, (2,), 1e-10
    elif case == "k94":   # test_k94_latent_heat_dropped_matches_oracle
        p = configs.dendrite_k94_params(n=32, mode=abi.MODE_DBA, thresh_b=0.01)
        fields, steps, muts, bar = list(configs.dendrite_k94_ic(32, 32, 32)), 36, (2,), 1e-10
    else:
        n, nz = 48, 8
        X = list(configs.modelb_ic(n, n, nz)) + [np.zeros((nz, n, n))]
        X[0][..., n // 2:] = 0.2
        X[1][..., n // 2:] = 0.2
        X[2][..., n // 2:] = 0.01
        p = configs.modelb_params(n=n, nz=nz, mode=abi.MODE_DBA, eps=1e-4)
        fields, steps, muts, bar = X, 20, (1,), 1e-10
    ref = _final(None, p, fields, steps)
    for k in muts:
        gap = _gap(_final(mutants[k], p, fields, steps), ref)
        assert gap > 100 * bar, f"{case}: mutant {k} ends only {gap} apart"
