"""The reference's own analysis programs read the drop-in drivers' output files
unchanged (CPU; needs the reference sources, so skipped where they are absent).

oracle/Makefile.ref builds Compare_VTK and Extract_Tip
(example_Dendritic_growth/Compare_VTK.cpp, Extract_Tip.cpp) and
coarsening_analysis (example_Reactive_ModelB/coarsening_analysis.cpp) from the
reference sources into oracle/_ref. tests/vtk_harness.cpp writes files with the
drivers' own writers (dba_amd/host/host_common.h: write_structured, the 3D
dendrite's p_<t>.vtk; write_modelb_output, the Model B driver's
output_<t>.vtk) from oracle fields, and the tools' printed numbers are checked
against the same numbers computed here from the values as the files hold them
(6 significant digits).
"""
import os
import re
import resource
import subprocess

import numpy as np
import pytest

from dba_amd import abi, configs
from droplets_ref import host_droplets
from oracle.oracle import OracleSim

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
REF_OUT = os.path.join(ROOT, "oracle", "_ref")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")


@pytest.fixture(scope="module")
def tools(tmp_path_factory):
    subprocess.run(["make", "-s", "-f", "Makefile.ref", "tools"], cwd=os.path.join(ROOT, "oracle"), check=True,
                   capture_output=True)
    d = tmp_path_factory.mktemp("harness")
    exe = d / "vtk_harness"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), os.path.join(ROOT, "tests", "vtk_harness.cpp")],
                   check=True, capture_output=True)
    return {"harness": str(exe), **{t: os.path.join(REF_OUT, t)
                                    for t in ("Compare_VTK", "Extract_Tip", "coarsening_analysis")}}


def raw(path, a):
    np.ascontiguousarray(a, dtype=np.float64).tofile(path)
    return str(path)


def vtk_values(path, skip):
    """The values of a VTK file after `skip` header lines, as written (text)."""
    with open(path) as f:
        for _ in range(skip):
            f.readline()
        return np.array(f.readline().split(), dtype=np.float64)


def write_dendrite(tools, d, t, phi):
    nz, ny, nx = phi.shape
    subprocess.run([tools["harness"], "structured", "p", str(t), str(nx), str(ny), str(nz),
                    raw(d / f"phi_{t}.bin", phi)], cwd=d, check=True)
    return d / f"p_{t}.vtk"


@pytest.fixture(scope="module")
def dendrite_files(tools, tmp_path_factory):
    """p_0.vtk and p_200.vtk of the reference 3D dendrite (oracle, 24x24x84 periodic;
    Extract_Tip reads z in [Nz/2, Nz/2 + 40], so Nz >= 82)."""
    d = tmp_path_factory.mktemp("dendrite")
    nx, nz = 24, 84
    p = configs.dendrite_k94_params(n=nx, nz=nz, mode=abi.MODE_REGULAR)
    phi, T = configs.dendrite_k94_ic(nx, nx, nz)
    o = OracleSim(p)
    o.set_field(0, phi)
    o.set_field(1, T)
    f0 = write_dendrite(tools, d, 0, o.get_field(0))
    o.step(200)
    f1 = write_dendrite(tools, d, 200, o.get_field(0))
    o.close()
    return d, f0, f1


def test_compare_vtk_identical_and_perturbed(tools, dendrite_files):
    d, f0, f1 = dendrite_files
    r = subprocess.run([tools["Compare_VTK"], str(f1), str(f1)], cwd=d, capture_output=True, text=True, check=True)
    m = re.search(r"(\d+) elements have been compared, the total difference is: (\S+),", r.stdout)
    assert m, r.stdout
    assert int(m.group(1)) == 24 * 24 * 84
    assert float(m.group(2)) == 0.0
    # two different steps: the sum of |a - b| over the values the files hold
    r = subprocess.run([tools["Compare_VTK"], str(f0), str(f1)], cwd=d, capture_output=True, text=True, check=True)
    m = re.search(r"(\d+) elements have been compared, the total difference is: (\S+),", r.stdout)
    a, b = vtk_values(f0, 10), vtk_values(f1, 10)
    assert a.size == b.size == 24 * 24 * 84
    expect = float(np.sum(np.abs(a - b)))
    assert expect > 1.0
    assert float(m.group(2)) == pytest.approx(expect, rel=1e-5)   # printed with 6 digits
    # the difference-file form writes the per-cell differences in order
    r = subprocess.run([tools["Compare_VTK"], str(f0), str(f1), "diff.vtk"], cwd=d, capture_output=True, text=True,
                       check=True)
    diff = vtk_values(d / "diff.vtk", 10)
    np.testing.assert_allclose(diff, a - b, rtol=1e-5, atol=1e-12)


def test_extract_tip(tools, dendrite_files):
    d, f0, f1 = dendrite_files
    r = subprocess.run([tools["Extract_Tip"], str(f0), str(f1)], cwd=d, capture_output=True, text=True, check=True)
    tips = [int(v) for v in re.findall(r"Tip at x = (\d+)", r.stdout)]
    expect, prof = [], []
    for f in (f0, f1):
        phi = vtk_values(f, 10).reshape(84, 24, 24)
        loc = []
        for z in range(42, 42 + 41):
            xs = np.nonzero((phi[z] > 0.5).any(axis=0))[0]
            loc.append(int(xs.max()) if xs.size else 0)
        expect.append(max(loc))
        prof.append([max(loc) - v for v in loc])
    assert tips == expect
    assert expect[1] > expect[0]   # the tip advanced
    rows = [line.split() for line in open(d / "Tip_Profile.dat").read().splitlines()[1:]]
    assert [int(r[0]) for r in rows] == list(range(41))
    assert [[int(v) for v in r[1:]] for r in rows] == [list(c) for c in zip(*prof)]


def test_coarsening_analysis_reads_modelb_output(tools, tmp_path):
    """coarsening_analysis is hard-coded to 256^3 and output_<i*100000>.vtk,
    i = 0..100 (coarsening_analysis.cpp:12-15, 176-178). Only output_0.vtk is
    written; the later files are missing, which the tool reports and skips over,
    so only the first line of avg_radius.txt and size_dist.txt is checked."""
    n = 256
    rng = np.random.default_rng(7)
    z = np.full((n, n, n), 0.05)
    zz, yy, xx = np.ogrid[0:n, 0:n, 0:n]
    for _ in range(60):   # spheres, some across the periodic faces
        c = rng.integers(0, n, 3)
        r = rng.uniform(3.0, 7.0)
        dz = np.minimum(np.abs(zz - c[0]), n - np.abs(zz - c[0]))
        dy = np.minimum(np.abs(yy - c[1]), n - np.abs(yy - c[1]))
        dx = np.minimum(np.abs(xx - c[2]), n - np.abs(xx - c[2]))
        z[dz * dz + dy * dy + dx * dx <= r * r] = 0.8
    x = np.full_like(z, 0.1)
    y = np.full_like(z, 0.1)
    phi = np.zeros_like(z)
    subprocess.run([tools["harness"], "modelb", "0", str(n), str(n), str(n), raw(tmp_path / "x.bin", x),
                    raw(tmp_path / "y.bin", y), raw(tmp_path / "z.bin", z), raw(tmp_path / "phi.bin", phi)],
                   cwd=tmp_path, check=True)
    for f in ("x.bin", "y.bin", "z.bin", "phi.bin"):
        os.remove(tmp_path / f)

    def big_stack():   # the tool's flood fill recurses once per droplet cell
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    r = subprocess.run([tools["coarsening_analysis"]], cwd=tmp_path, capture_output=True, text=True,
                       preexec_fn=big_stack, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Read output_0.vtk success" in r.stdout
    zf = vtk_values(tmp_path / "output_0.vtk", 20).reshape(n, n, n)
    assert zf.size == n ** 3
    avg, vol, sizes = host_droplets(zf)
    assert len(sizes) > 20
    got_sizes = [int(v) for v in open(tmp_path / "size_dist.txt").readline().split()]
    assert got_sizes == sizes
    line0 = open(tmp_path / "avg_radius.txt").readline().strip()
    i, rad = line0.split(",")
    assert int(i) == 0
    assert float(rad) == pytest.approx(float(np.cbrt(vol * 0.75 / np.pi / len(sizes))), rel=1e-5)
