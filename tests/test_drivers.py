"""Drop-in drivers (dba_amd/bin/{3D_dendritic_growth, CH, KH_DBA}).

CPU: Input.txt handling and the reference's exit codes.
GPU: each driver runs a small Input.txt in a scratch folder; the files it
writes (reference names and layouts) are parsed and checked against the oracle
run on the same parameters and initial state. Values are printed with the
stream's default 6 significant digits, so fields compare to rtol 1e-5.
"""
import os
import subprocess

import numpy as np
import pytest

from dba_amd import abi, configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "dba_amd", "bin")
PROGS = ("3D_dendritic_growth", "CH", "KH_DBA")


def prog(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built (run __graft_entry__.build())")
    return path


def run(name, folder, text, timeout=300):
    (folder / "Input.txt").write_text(text)
    r = subprocess.run([prog(name)], cwd=folder, capture_output=True, text=True, timeout=timeout)
    return r


# ----------------------------------------------------------------- Input.txt ---

KH_INPUT = """{n} {n} 1                    Nx Ny Nz

{tt} {tf} {dt}            t_total, t_freq, dt

{dx} {dx} {dx}             dx, dy, dz

{mode}                            Mode of computing

{eps} Metric_eps
{extra}"""

MB_INPUT = """{nx} {ny} {nz}                    Nx Ny Nz

{tt} {tf} {dt}           t_total, t_freq, dt

1 1 1                        dx, dy, dz

0 0 0 0 0 4                  chi_xy chi_xz chi_yz chi_xb chi_yb chi_zb

1 1 2                        vx vy vz

1 1 2                        rx ry rz

1 1                          n m

4 4 4 0.01                   epsilonx_sq epsilony_sq epsilonz_sq epsilonphi_sq

1 1 1 1                      Mobility_x Mobility_y Mobility_z Mobility_phi

0.2 0.2 0.01                 x0 y0 z0

100 0.4 1                    K z_crit p

1                            k_0

{eps} Metric_eps # comment after the value

{mode}                            Mode of computing
{extra}"""

DEND_INPUT = """{n} {n} {nz}                  Nx, Ny, Nz

{tt} {tf} {dt}          t_total, t_freq, dt

{dx} {dx} {dx}               dx, dy, dz

{delta} {tau} 0.01 0.9 {gamma}      delta, tau, epsilon, alpha, gamma

0.25                         Delta

{mode}                            Mode

0.01                         Metric_p

1e-5 Metric_T
{extra}"""


@pytest.mark.parametrize("name", PROGS)
def test_missing_input_exits_1(tmp_path, name):
    r = subprocess.run([prog(name)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "Can not openInput.txt" in r.stdout


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_device_exits_1_after_parsing(tmp_path):
    """Parsing succeeds (all groups and extras echoed), then the device error
    maps to exit code 1 like the reference's CUDACHECK."""
    r = run("CH", tmp_path, MB_INPUT.format(nx=16, ny=16, nz=16, tt=10, tf=5, dt=0.0005, eps=1e-6, mode=2,
                                            extra="model modelb\ntile 4\n"))
    assert r.returncode == 1
    assert "Nx = 16, Ny = 16, Nz = 16" in r.stdout
    assert "Metric_eps = 1e-06" in r.stdout
    assert "Mode = 2" in r.stdout
    assert "DBA error in dba_init" in r.stdout


def test_unknown_model_rejected(tmp_path):
    r = run("3D_dendritic_growth", tmp_path, DEND_INPUT.format(n=32, nz=32, tt=10, tf=5, dt=5e-5, dx=0.02,
                                                              delta=0.5, tau=3e-4, gamma=40., mode=1,
                                                              extra="model k95\n"))
    assert r.returncode == 1
    assert "unknown model k95" in r.stdout


# ------------------------------------------------------------------ parsing ---

def read_vtk_scalars(path):
    """{name: array} from a RECTILINEAR_GRID or STRUCTURED_POINTS file, plus 'vec'."""
    out = {}
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines):
        ln = lines[i]
        if ln.startswith("SCALARS "):
            name = ln.split()[1]
            out[name] = np.array(lines[i + 2].split(), dtype=np.float64)
            i += 3
        elif ln.startswith("VECTORS "):
            rows = [r.split() for r in lines[i + 1:] if r.strip()]
            out["vec"] = np.array(rows, dtype=np.float64)
            break
        else:
            i += 1
    return out


def read_table(path, skip_header=False):
    with open(path) as f:
        rows = f.read().strip().split("\n")
    if skip_header:
        rows = rows[1:]
    return np.array([r.split() for r in rows], dtype=np.float64)


def close(a, b, rtol=1e-5, atol=1e-12):
    a, b = np.asarray(a, dtype=np.float64).ravel(), np.asarray(b, dtype=np.float64).ravel()
    assert a.shape == b.shape
    assert np.allclose(a, b, rtol=rtol, atol=atol), np.max(np.abs(a - b))


def expand(mask, tiles, shape):
    """Per-point flags from per-tile flags (x fastest), as Output_wakeup writes them."""
    nz, ny, nx = shape
    tx, ty, tz = tiles
    m = np.asarray(mask).reshape(nz // tz, ny // ty, nx // tx)
    return np.repeat(np.repeat(np.repeat(m, tz, 0), ty, 1), tx, 2).ravel()


# ------------------------------------------------------------------ GPU runs ---

@pytest.mark.gpu
@pytest.mark.parametrize("mode", [abi.MODE_DBA, abi.MODE_CONSERVED])
def test_kh_driver_reference_model(tmp_path, mode):
    from oracle.oracle import OracleSim
    n, dt, tt, tf, eps = 128, 2e-4, 60, 20, 8.0
    dx = 1.0 / n
    r = run("KH_DBA", tmp_path, KH_INPUT.format(n=n, tt=tt, tf=tf, dt=dt, dx=repr(dx), mode=mode, eps=eps,
                                                 extra="dump 1\n"))
    assert r.returncode == 0, r.stdout + r.stderr
    p = configs.kh_reference_params(n=n, tile=8, mode=mode, dx=dx, dt=dt, eps=eps)
    U0 = configs.kh_reference_ic(n, dx=dx)
    o = OracleSim(p)
    for f in range(4):
        o.set_field(f, U0[f])
    # t = 0 output
    v0 = read_vtk_scalars(tmp_path / f"output_{mode}_0.vtk")
    rho, vx, vy, pr = configs.kh_primitive(U0, p.gas_gamma, dx * dx)
    close(v0["rho"], rho)
    close(v0["p"], pr)
    integ = []
    for t in range(tf, tt + 1, tf):
        o.step(tf)
        U = [o.get_field(f) for f in range(4)]
        for k, name in enumerate(("Mass", "Momx", "Momy", "Energy")):
            d = raw(tmp_path / f"{name}_{t}.bin")
            if mode == abi.MODE_CONSERVED:   # lazy vs eager mean shift: summation order
                assert np.max(np.abs(d - U[k].ravel())) <= 1e-12 * max(1.0, np.abs(U[k]).max())
            else:
                assert np.array_equal(d, U[k].ravel())
        rho, vx, vy, pr = configs.kh_primitive(U, p.gas_gamma, dx * dx)
        v = read_vtk_scalars(tmp_path / f"output_{mode}_{t}.vtk")
        close(v["rho"], rho)
        close(v["p"], pr)
        close(v["vec"][:, 0], vx, atol=1e-9)
        close(v["vec"][:, 1], vy, atol=1e-9)
        assert np.all(v["vec"][:, 2] == 0.0)
        act = read_vtk_scalars(tmp_path / f"Active_{t}.vtk")["wakeup"]
        assert np.array_equal(act, expand(o.awake(0), (8, 8, 1), (1, n, n)))
        # Calc_Vy_mode restated
        xs = np.arange(n)[None, :] / n
        ys = np.arange(n)[:, None] / n * np.ones((1, n))
        w = np.exp(-4.0 * np.pi * np.abs(ys - np.where(ys < 0.5, 0.25, 0.75)))
        ss, sc, sd = (vy * np.sin(4 * np.pi * xs) * w).sum(), (vy * np.cos(4 * np.pi * xs) * w).sum(), w.sum()
        integ.append((t, [u.sum() for u in U], 2.0 * np.sqrt(ss * ss / (sd * sd) + sc * sc / (sd * sd))))
    wp = read_table(tmp_path / "Wakeup_portion.out")
    assert list(wp[:, 0]) == [t for t, _, _ in integ]
    assert wp[-1, 1] == pytest.approx(o.awake(0).sum() / o.ntiles, abs=1e-6)
    assert np.all(wp[:, 2] > 0)
    vm = read_table(tmp_path / "VY_mode.out")
    close(vm[:, 1], [m for _, _, m in integ])
    if mode == abi.MODE_CONSERVED:
        it = read_table(tmp_path / "Integral.out")
        assert it.shape == (len(integ) + 1, 5)
        close(it[1:, 1], [s[0] for _, s, _ in integ])   # Mass
        close(it[1:, 2], [s[3] for _, s, _ in integ])   # Energy
    else:
        it = read_table(tmp_path / "Integral_mass.out")
        close(it[1:, 1], [s[0] for _, s, _ in integ])
    o.close()


@pytest.mark.gpu
def test_kh_driver_walls_cfl(tmp_path):
    from oracle.oracle import OracleSim
    n, tt, tf = 128, 40, 20
    extra = "gamma 1.4\ntile 16\nwalls 1\nmetric density\ncfl 0.4\nic sharp\n"
    r = run("KH_DBA", tmp_path, KH_INPUT.format(n=n, tt=tt, tf=tf, dt=2e-4, dx=repr(1.0 / n), mode=1, eps=20.0,
                                                extra=extra))
    assert r.returncode == 0, r.stdout + r.stderr
    p = configs.kh_params(n=n, tile=16, mode=abi.MODE_DBA, eps=20.0)
    U0 = configs.kh_ic(n)
    o = OracleSim(p)
    for f in range(4):
        o.set_field(f, U0[f])
    o.step(tt)
    rho, vx, vy, pr = configs.kh_primitive([o.get_field(f) for f in range(4)], 1.4, 1.0 / n / n)
    v = read_vtk_scalars(tmp_path / f"output_1_{tt}.vtk")
    close(v["rho"], rho)
    close(v["p"], pr)
    assert f"simulated time {o.time():.6g}" in r.stdout
    o.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [abi.MODE_REGULAR, abi.MODE_DBA, abi.MODE_CONSERVED])
def test_modelb_driver(tmp_path, mode):
    from oracle.oracle import OracleSim
    n, tt, tf, eps = 16, 40, 20, 1e-3
    r = run("CH", tmp_path, MB_INPUT.format(nx=n, ny=n, nz=n, tt=tt, tf=tf, dt=0.0005, eps=eps, mode=mode,
                                            extra="dump 1\ndroplets 1\n"))
    assert r.returncode == 0, r.stdout + r.stderr
    p = configs.modelb_params(n=n, mode=mode, eps=eps)
    X, Y, Z = configs.modelb_ic(n, n, n)
    o = OracleSim(p)
    for f, a in enumerate((X, Y, Z, np.zeros_like(X))):
        o.set_field(f, a)
    v0 = read_vtk_scalars(tmp_path / "output_0.vtk")
    close(v0["X"], X)
    close(v0["B"], 1.0 - X - Y - Z)
    sums = []
    for t in range(tf, tt + 1, tf):
        o.step(tf)
        x, y, z, phi = (o.get_field(f) for f in range(4))
        v = read_vtk_scalars(tmp_path / f"output_{t}.vtk")
        for k, a in (("X", x), ("Y", y), ("Z", z), ("B", 1.0 - x - y - z), ("PHI", phi)):
            close(v[k], a, atol=1e-9)
        for k, a in (("X", x), ("Y", y), ("Z", z), ("PHI", phi)):   # raw dump: the GPU parity bar
            assert np.max(np.abs(raw(tmp_path / f"{k}_{t}.bin") - a.ravel())) <= 1e-10
        if mode != abi.MODE_REGULAR:
            wk = read_vtk_scalars(tmp_path / f"Wakeup_{t}.vtk")["wakeup"]
            assert np.array_equal(wk, expand(o.awake(0), (4, 4, 4), (n, n, n)))
        sums.append([x.sum(), y.sum(), z.sum(), (1.0 - x - y - z).sum()])
    # coarsening_analysis skips 20 lines and reads Z
    with open(tmp_path / f"output_{tt}.vtk") as f:
        lines = f.read().split("\n")
    close(np.array(lines[20].split(), dtype=np.float64), z, atol=1e-9)
    it = read_table(tmp_path / "Integral.out")
    assert it.shape == (tt // tf + 1, 5)
    close(it[1:, 1:], np.array(sums), rtol=1e-5, atol=1e-6)
    if mode == abi.MODE_CONSERVED:
        assert not (tmp_path / "RTvalues.out").exists()
        wp = read_table(tmp_path / "Wakeup_portion.out")
    else:
        with open(tmp_path / "RTvalues.out") as f:
            assert f.readline().strip() == "Time Act_portion_p Steps_per_second"
        wp = read_table(tmp_path / "RTvalues.out", skip_header=True)
    assert list(wp[:, 0]) == list(range(tf, tt + 1, tf))
    expect = 1.0 if mode == abi.MODE_REGULAR else o.awake(0).sum() / o.ntiles
    assert wp[-1, 1] == pytest.approx(expect, abs=1e-6)
    check_droplets(tmp_path, "Z", [0] + list(range(tf, tt + 1, tf)), tf, (n, n, n))
    o.close()


def raw(path):
    return np.fromfile(path, dtype="<f8")


def check_droplets(folder, name, steps, tf, shape):
    """avg_radius.txt / size_dist.txt of the driver's device analysis against the
    host reference on the dumped field of each output step."""
    from droplets_ref import host_droplets, radius
    rad = [line.split(",") for line in (folder / "avg_radius.txt").read_text().splitlines()]
    dist = (folder / "size_dist.txt").read_text().splitlines()
    assert len(rad) == len(steps) == len(dist)
    for (k, r), sizes, t in zip(rad, dist, steps):
        _, vol, want = host_droplets(raw(folder / f"{name}_{t}.bin").reshape(shape))
        assert int(k) == t // tf
        assert [int(v) for v in sizes.split()] == want
        if want:
            assert float(r) == pytest.approx(radius(vol, len(want), shape[0] > 1), rel=1e-5)


@pytest.mark.gpu
def test_ch_driver(tmp_path):
    from oracle.oracle import OracleSim
    n, tt, tf, eps = 128, 60, 30, 0.05
    extra = "model ch\ntile 32\nmobility 1\nkappa 1\nreact_gamma 1e-3\ncbar 0\namp 0.05\nstencil 9\ndump 1\ndroplets 1\n"
    r = run("CH", tmp_path, MB_INPUT.format(nx=n, ny=n, nz=1, tt=tt, tf=tf, dt=0.01, eps=eps, mode=1, extra=extra))
    assert r.returncode == 0, r.stdout + r.stderr
    p = configs.ch_params(n=n, tile=32, mode=abi.MODE_DBA, eps=eps)
    o = OracleSim(p)
    o.set_field(0, configs.ch_ic(n, n))
    o.step(tt)
    v = read_vtk_scalars(tmp_path / f"output_{tt}.vtk")
    close(v["C"], o.get_field(0))
    assert np.array_equal(raw(tmp_path / f"C_{tt}.bin"), o.get_field(0).ravel())   # CH is bit-exact
    wk = read_vtk_scalars(tmp_path / f"Wakeup_{tt}.vtk")["wakeup"]
    assert np.array_equal(wk, expand(o.awake(0), (32, 32, 1), (1, n, n)))
    check_droplets(tmp_path, "C", [0, tf, tt], tf, (1, n, n))
    o.close()


def growth_front(phi, ix):
    """growth_front of example_Dendritic_growth/kernel.cu:760-774 (0 when not found)."""
    nz, ny, nx = phi.shape
    col = phi[nz // 2, :, ix]
    for y in range(ny // 2, ny):
        if col[y] < 0.5:
            return (col[y - 1] - 0.5) / (col[y - 1] - col[y]) + y - 1 - ny // 2
    return 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["k94", "k93"])
def test_dendrite_driver(tmp_path, model):
    from oracle.oracle import OracleSim
    if model == "k94":
        n, nz, tt, tf, dt, dx = 32, 32, 100, 50, 5e-5, 0.02
        text = DEND_INPUT.format(n=n, nz=nz, tt=tt, tf=tf, dt=dt, dx=dx, delta=0.5, tau=3e-4, gamma=40., mode=1,
                                 extra="vtk 1\ndump 1\n")
        p = configs.dendrite_k94_params(n=n, nz=nz)
        phi0, T0 = configs.dendrite_k94_ic(n, n, nz)
    else:
        n, nz, tt, tf, dt, dx = 128, 1, 200, 100, 1e-4, 0.03
        text = DEND_INPUT.format(n=n, nz=nz, tt=tt, tf=tf, dt=dt, dx=dx, delta=0.02, tau=3e-4, gamma=10., mode=1,
                                 extra="model k93\ntile 16\nK 1.6\nj 4\ntheta0 0\nT_eq 1\nbc noflux\nvtk 1\ndump 1\n")
        p = configs.dendrite_k93_params(n=n, tile=16)
        phi0, T0 = configs.dendrite_k93_ic(n, n)
        phi0, T0 = phi0[None], T0[None]
    r = run("3D_dendritic_growth", tmp_path, text)
    assert r.returncode == 0, r.stdout + r.stderr
    o = OracleSim(p)
    o.set_field(0, phi0)
    o.set_field(1, T0)
    close(read_vtk_scalars(tmp_path / "p_0.vtk")["p"], phi0)
    rows, tips, prev = [], [], 0.0
    for t in range(tf, tt + 1, tf):
        o.step(tf)
        phi = o.get_field(0).reshape(nz, n, n)
        T = o.get_field(1).reshape(nz, n, n)
        close(read_vtk_scalars(tmp_path / f"p_{t}.vtk")["p"], phi, atol=1e-9)
        assert np.max(np.abs(raw(tmp_path / f"p_{t}.bin") - phi.ravel())) <= 1e-10
        assert np.max(np.abs(raw(tmp_path / f"T_{t}.bin") - T.ravel())) <= 1e-10
        close(read_vtk_scalars(tmp_path / f"T_{t}.vtk")["T"], T, atol=1e-9)
        d = growth_front(phi, n // 2)
        rows.append([t, (d - prev) / tf, o.awake(0).sum() / o.ntiles, o.awake(1).sum() / o.ntiles])
        prev = d
        tips.append([growth_front(phi, ix) for ix in range(n // 2 - 10, n // 2 + 11)])
    with open(tmp_path / "RTValues.out") as f:
        assert f.readline().strip() == "Time Vel Act_portion_p Act_portion_T Steps_per_second"
    rt = read_table(tmp_path / "RTValues.out", skip_header=True)
    close(rt[:, :4], np.array(rows), atol=1e-6)
    close(read_table(tmp_path / "tip_profile.out"), np.array(tips), atol=1e-5)
    o.close()
