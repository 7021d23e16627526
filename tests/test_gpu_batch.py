"""GPU: batched sweeps and the serving loop through libdba_hip.so.

Bar: an instance's fields and awake sets are bit-identical whether it is
stepped alone or in a merged-worklist batch / serving mix (instances share
launches, not data); the workgroup size never changes results; a batched
dendrite matches the oracle to 1e-10 (the Kobayashi-93 parity bar: atan and the
anisotropy's sqrt/division are correctly rounded on neither side)."""
import numpy as np
import pytest

from dba_amd import batch, serve
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu


def small_sweep(n=6):
    return batch.make_sweep(n=n, master_seed=5, side_lo=64, side_hi=256, multiple=32)


def fields_of(inst):
    return [inst.download(0), inst.download(1)]


@pytest.mark.parametrize("n", [6, 36])
def test_batched_equals_alone(ctx, n):
    """36 instances: the finalize merge takes three workgroups of instance
    reductions (16 instances each) beside the offsets scan."""
    items = small_sweep(n)
    st, insts = batch.run_batched(ctx, items, 300, tile=16, chunk=100, keep=True)
    assert st.active_updates > 0
    infos = [inst.info() for inst in insts]
    assert sum(i.steps_done for i in infos) == 300 * n
    for it, inst in zip(items, insts):
        (alone,) = batch._create(ctx, [it], 16, 0)
        alone.step(300, graph_chunk=100)
        for a, b in zip(fields_of(inst), fields_of(alone)):
            assert np.array_equal(a, b), it.index
        assert np.array_equal(inst.active(0)[1], alone.active(0)[1])
        assert np.array_equal(inst.active(1)[1], alone.active(1)[1])
        ai = alone.info()
        bi = inst.info()
        assert (bi.steps_done, bi.cell_updates, bi.sim_time) == (ai.steps_done, ai.cell_updates, ai.sim_time)
        assert np.array_equal(inst.history(), alone.history())
        alone.destroy()
        inst.destroy()


def test_batched_many_segments_equals_alone(ctx):
    """300 small instances, one segment each: a merged list of more than 257
    segments, whose lookup counts the segment offsets in two rounds of loads and
    reads the item's segment offset back from memory, and a merge of 19
    workgroups of instance reductions."""
    items = batch.make_sweep(n=300, master_seed=7, side_lo=48, side_hi=96, multiple=16)
    st, insts = batch.run_batched(ctx, items, 60, tile=16, chunk=30, keep=True)
    assert st.active_updates > 0
    for it, inst in zip(items, insts):
        (alone,) = batch._create(ctx, [it], 16, 0)
        alone.step(60, graph_chunk=30)
        for a, b in zip(fields_of(inst), fields_of(alone)):
            assert np.array_equal(a, b), it.index
        assert np.array_equal(inst.active(0)[1], alone.active(0)[1])
        assert np.array_equal(inst.active(1)[1], alone.active(1)[1])
        assert inst.info().cell_updates == alone.info().cell_updates
        alone.destroy()
        inst.destroy()


def test_batched_matches_oracle(ctx):
    items = [it for it in small_sweep(10) if it.aniso_j == 6][:1] + \
            [it for it in small_sweep(10) if it.aniso_j == 4][:1]
    assert len(items) == 2
    st, insts = batch.run_batched(ctx, items, 150, tile=8, threads=64, chunk=50, keep=True)
    for it, inst in zip(items, insts):
        o = OracleSim(it.params(tile=8, threads=64))
        phi, T = it.fields()
        o.set_field(0, phi)
        o.set_field(1, T)
        o.step(150)
        for f in (0, 1):
            assert np.max(np.abs(inst.download(f) - o.get_field(f))) <= 1e-10, (it.index, f)
        assert np.array_equal(inst.active(0)[1], o.awake(0))
        assert np.array_equal(inst.active(1)[1], o.awake(1))
        o.close()
        inst.destroy()


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_threads_do_not_change_results(ctx, tile):
    items = small_sweep(4)
    ref = None
    for nt in batch.THREADS:
        _, insts = batch.run_batched(ctx, items, 120, tile=tile, threads=nt, chunk=60, keep=True)
        got = [fields_of(i) for i in insts]
        for i in insts:
            i.destroy()
        if ref is None:
            ref = got
        else:
            for a, b in zip(ref, got):
                assert all(np.array_equal(x, y) for x, y in zip(a, b)), (tile, nt)


def test_autotune_picks_a_timed_candidate(ctx):
    items = small_sweep(4)
    best, table = batch.autotune(ctx, items, trial_steps=20, warm_steps=5, chunk=10)
    assert len(table) == 18
    assert best in [(t, n, m) for t, n, m, _, _ in table]
    assert all(ms > 0 for _, _, _, ms, _ in table)


def run_alone(ctx, r, tiles):
    inst = ctx.create(r.params(tiles, {}))
    flds = r.fields()
    for f, a in enumerate(flds):
        inst.upload(f, a)
    inst.step(r.steps, graph_chunk=100)
    out = [inst.download(f) for f in range(len(flds))]
    upd = inst.info().cell_updates
    inst.destroy()
    return out, upd


@pytest.mark.parametrize("chunk,modes", [(100, None), (7, None), (13, {"ch": 2, "dendrite": 1})])
def test_serving_result_same_alone_and_in_mix(ctx, chunk, modes):
    """Staggered arrivals under a tight budget: requests join and leave the
    resident set at different ticks (chunk 7 also drives the graph and group
    caches through eviction; the per-kind modes step the CH batch with one
    launch per step); each result equals its solo run bit for bit."""
    rs = serve.make_requests(10, master_seed=21, side_lo=64, side_hi=256, steps_lo=30, steps_hi=400)
    for k, r in enumerate(rs):
        r.arrival = 0.002 * k
    tiles = {"dendrite": 16, "ch": 32}
    be = serve.LibraryBackend(ctx)
    cap = 3 * max(ctx.footprint(r.params(tiles, {})) for r in rs)
    srv = serve.Server(be, cap, chunk=chunk, tiles=tiles, keep_fields=True, modes=modes)
    done = srv.run(rs)
    assert sorted(c.rid for c in done) == list(range(10))
    assert srv.max_used <= cap
    assert srv.ticks > 1
    by = {r.rid: r for r in rs}
    for c in done:
        alone, upd = run_alone(ctx, by[c.rid], tiles)
        for a, b in zip(c.fields, alone):
            assert np.array_equal(a, b), c.rid
        assert c.active_updates == upd
