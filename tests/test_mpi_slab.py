"""The MPI slab restatement of the CH regular-mode loop (oracle/mpi_slab.c, the
CPU baseline's all-core leg) equals the serial oracle bit for bit, for rank
counts that split the rows evenly and unevenly (CPU; skipped without MPICH)."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from dba_amd import abi, configs
from oracle.oracle import OracleSim

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_build", "mpi_slab")


def mpirun():
    return shutil.which("mpirun") or ("/opt/conda/bin/mpirun" if os.path.exists("/opt/conda/bin/mpirun") else None)


pytestmark = pytest.mark.skipif(mpirun() is None or not os.path.exists(EXE), reason="MPICH or mpi_slab not built")


@pytest.mark.parametrize("ranks,stencil", [(1, 9), (3, 9), (4, 5)])
def test_mpi_slab_equals_serial_oracle(tmp_path, ranks, stencil):
    nx, ny, steps = 96, 70, 25
    p = configs.ch_params(n=nx, ny=ny, tile=1, mode=abi.MODE_REGULAR, stencil=stencil, react_gamma=3e-3, cbar=0.1)
    c0 = configs.ch_ic(nx, ny, cbar=0.1, seed=77)
    (tmp_path / "p.bin").write_bytes(bytes(p))
    c0.astype(np.float64).tofile(tmp_path / "ic.bin")
    r = subprocess.run([mpirun(), "-np", str(ranks), EXE, str(tmp_path / "p.bin"), str(tmp_path / "ic.bin"),
                        str(steps), str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rep = json.loads(r.stdout.strip().splitlines()[-1])
    assert rep["ranks"] == ranks and rep["steps"] == steps
    got = np.fromfile(tmp_path / "out.bin").reshape(ny, nx)
    o = OracleSim(p)
    o.set_field(0, c0)
    o.step(steps)
    ref = o.get_field(0)
    o.close()
    assert np.array_equal(got, ref)
    assert float(np.max(np.abs(ref - c0))) > 1e-6   # it did evolve
