"""GPU: every request of a served stream (the serving workload's request
distributions) against the oracle run of the same request alone.

The stream mixes Kobayashi dendrites (delta U[0.01, 0.05], j in {4, 6},
K U[1.2, 2.0], 1-3 seeds) and reactive Cahn-Hilliard solves (Gamma
log-uniform in [1e-4, 1e-2], cbar U[-0.2, 0.2]), sides log-uniform in
[256, 1024] (the serving range's lower part, where the oracle finishes in
seconds), under a budget of three large requests, so requests join and leave
the resident set at different ticks and share launches.

Bar: CH fields bit-exact; dendrite fields within 1e-10 absolute (the device
atan differs from glibc's by a few ulp per call); the device count of active
cell-updates of a CH request equals the cells of the tiles the oracle ran,
summed over its steps (partial last tiles count their own cells).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from dba_amd import serve
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu

TILES = {"dendrite": 16, "ch": 32}


def tile_cells(p):
    """Cells of every tile (partial last tiles included), in tile order."""
    wx = np.minimum(p.tile_x, p.nx - p.tile_x * np.arange(-(-p.nx // p.tile_x)))
    wy = np.minimum(p.tile_y, p.ny - p.tile_y * np.arange(-(-p.ny // p.tile_y)))
    return np.outer(wy, wx).reshape(-1)


def oracle_run(r):
    """The request alone in the oracle; for CH also the cells of the tiles that
    ran, summed over the steps."""
    p = r.params(TILES, {})
    o = OracleSim(p)
    for f, a in enumerate(r.fields()):
        o.set_field(f, a)
    cells = tile_cells(p)
    updates = 0
    for _ in range(r.steps):   # ctypes releases the GIL: the requests run in parallel
        if r.kind == "ch":
            updates += int(cells[o.awake(0) == 1].sum())
        o.step(1)
    out = [o.get_field(f) for f in range(2 if r.kind == "dendrite" else 1)]
    o.close()
    return out, updates


@pytest.mark.parametrize("modes", [{"ch": 2, "dendrite": 1}, None])
def test_served_requests_match_oracle(ctx, modes):
    rs = serve.make_requests(8, master_seed=16, side_lo=256, side_hi=1024, steps_lo=50, steps_hi=200)
    assert {r.kind for r in rs} == {"ch", "dendrite"}
    assert max(r.side for r in rs if r.kind == "ch") >= 1000
    assert any(r.side % 32 for r in rs if r.kind == "ch") and any(r.side % 16 for r in rs if r.kind == "dendrite")
    assert {r.aniso_j for r in rs if r.kind == "dendrite"} == {4, 6}
    for k, r in enumerate(rs):
        r.arrival = 0.003 * k
    be = serve.LibraryBackend(ctx)
    cap = 3 * max(ctx.footprint(r.params(TILES, {})) for r in rs)
    srv = serve.Server(be, cap, chunk=25, tiles=TILES, keep_fields=True, modes=modes)
    done = srv.run(rs)
    assert sorted(c.rid for c in done) == list(range(len(rs)))
    assert srv.max_used <= cap and srv.ticks > 1
    with ThreadPoolExecutor(max_workers=8) as ex:
        refs = dict(zip([r.rid for r in rs], ex.map(oracle_run, rs)))
    by = {r.rid: r for r in rs}
    for c in done:
        r = by[c.rid]
        ref, updates = refs[c.rid]
        assert len(c.fields) == len(ref)
        for f, (g, o) in enumerate(zip(c.fields, ref)):
            if r.kind == "ch":
                assert np.array_equal(g.reshape(o.shape), o), (c.rid, f)
            else:
                err = float(np.max(np.abs(g.reshape(o.shape) - o)))
                assert err <= 1e-10, (c.rid, f, err)
        if r.kind == "ch":
            assert c.active_updates == updates, c.rid
