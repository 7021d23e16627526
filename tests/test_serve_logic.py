"""Host logic of batching and serving (no GPU): request and sweep draws,
admission against the memory budget, exact per-request step counts, latency
bookkeeping and percentiles, with a fake device that only counts steps."""
import math

import numpy as np
import pytest

from dba_amd import batch, serve


class FakeTime:
    def __init__(self):
        self.t = 0.0

    def clock(self):
        return self.t

    def sleep(self, s):
        self.t += s


class FakeBackend:
    """Counts steps per handle; device time = 1 ns per cell-step."""

    def __init__(self, tm, bytes_per_cell=40):
        self.tm = tm
        self.bpc = bytes_per_cell
        self.steps = {}
        self.calls = []
        self.modes = []
        self.live = set()
        self.next = 0

    def footprint(self, p):
        return p.nx * p.ny * self.bpc

    def admit(self, p, fields):
        assert all(f.shape == (p.ny, p.nx) for f in fields)
        self.next += 1
        h = (self.next, p.nx * p.ny)
        self.steps[h] = 0
        self.live.add(h)
        return h

    def step(self, handles, n, chunk, mode=None):
        assert n >= 1 and len(set(handles)) == len(handles)
        assert all(h in self.live for h in handles)
        self.calls.append((tuple(handles), n))
        self.modes.append(mode)
        for h in handles:
            self.steps[h] += n
        self.tm.t += 1e-9 * n * sum(h[1] for h in handles)

    def sync(self):
        pass

    def updates(self, h):
        return self.steps[h] * h[1]

    def download(self, h, nfields):
        return [np.full(4, self.steps[h])] * nfields

    def release(self, h):
        self.live.remove(h)


def small_requests(n, seed=3):
    return serve.make_requests(n, master_seed=seed, side_lo=64, side_hi=512, steps_lo=10, steps_hi=500)


def test_request_draws():
    rs = serve.make_requests(2000)
    kinds = [r.kind for r in rs]
    assert 0.45 < kinds.count("dendrite") / len(rs) < 0.55
    sides = np.array([r.side for r in rs])
    assert sides.min() >= 256 and sides.max() <= 4096 and np.all(sides % 8 != 1)
    # not rounded to the tile: every residue the library accepts occurs
    assert len(set((sides % 32).tolist())) >= 28
    steps = np.array([r.steps for r in rs])
    assert steps.min() >= 1000 and steps.max() <= 50000
    # log-uniform: the median of log(steps) sits at the middle of [log 1e3, log 5e4]
    assert abs(np.median(np.log(steps)) - 0.5 * (math.log(1e3) + math.log(5e4))) < 0.15
    for r in rs:
        if r.kind == "dendrite":
            assert 0.01 <= r.delta <= 0.05 and r.aniso_j in (4, 6) and 1.2 <= r.latent_k <= 2.0
            assert 1 <= len(r.seeds) <= 3
            assert all(16 <= c < r.side - 16 for s in r.seeds for c in s)
        else:
            assert 1e-4 <= r.react_gamma <= 1e-2 and -0.2 <= r.cbar <= 0.2
    # deterministic
    assert [(r.kind, r.side, r.steps) for r in serve.make_requests(50)] == \
        [(r.kind, r.side, r.steps) for r in serve.make_requests(50)]


def test_sweep_draws():
    items = batch.make_sweep()
    assert len(items) == 64
    for it in items:
        assert 256 <= it.side <= 2048 and it.side % 8 != 1
        assert 0.01 <= it.delta <= 0.05 and it.aniso_j in (4, 6) and 1.2 <= it.latent_k <= 2.0
        assert 1 <= len(it.seeds) <= 3
        assert all(16 <= c < it.side - 16 for s in it.seeds for c in s)
    phi, T = items[0].fields()
    assert phi.shape == (items[0].side,) * 2 and T.sum() == 0.0
    assert phi.sum() > 0
    s = batch.summary(items)
    assert s["instances"] == 64 and s["cells"] == sum(it.cells for it in items)


def test_log_uniform_side_rounding():
    assert batch.log_uniform_side(0.0, 256, 2048, 32) == 256
    assert batch.log_uniform_side(1.0, 256, 2048, 32) == 2048
    assert batch.log_uniform_side(0.5, 256, 2048, 32) == 736  # sqrt(256*2048) = 724.1 -> 736
    assert batch.log_uniform_side(0.5, 256, 2048) == 724
    # 8 k + 1 cells would leave a partial tile one cell wide: one cell more
    u = math.log(265 / 256) / math.log(2048 / 256)
    assert batch.log_uniform_side(u, 256, 2048) == 266


def test_percentile_nearest_rank():
    v = list(range(1, 101))
    assert serve.percentile(v, 50) == 50
    assert serve.percentile(v, 99) == 99
    assert serve.percentile(v, 100) == 100
    assert serve.percentile([3.0], 99) == 3.0
    assert serve.percentile([5, 1, 4, 2, 3], 50) == 3


@pytest.mark.parametrize("chunk", [1, 7, 100])
def test_every_request_runs_its_exact_steps_within_budget(chunk):
    tm = FakeTime()
    be = FakeBackend(tm)
    rs = small_requests(40)
    for k, r in enumerate(rs):
        r.arrival = 0.0005 * k
    cap = 3 * max(r.cells for r in rs) * 40
    srv = serve.Server(be, cap, chunk=chunk, keep_fields=True)
    done = srv.run(rs, clock=tm.clock, sleep=tm.sleep)
    assert sorted(c.rid for c in done) == [r.rid for r in rs]
    assert srv.max_used <= cap
    assert not be.live and srv.used == 0
    by = {r.rid: r for r in rs}
    for c in done:
        assert c.steps == by[c.rid].steps
        assert c.active_updates == c.steps * c.cells       # the fake counts every cell
        assert c.fields[0][0] == c.steps
        assert c.arrival <= c.admitted <= c.done
    # FIFO admission: requests are admitted in arrival order
    adm = sorted(done, key=lambda c: c.rid)
    assert all(a.admitted <= b.admitted for a, b in zip(adm, adm[1:]))
    # every tick is at most `chunk` steps
    assert all(n <= chunk for _, n in be.calls)


def test_continuous_batching_overlaps_requests():
    """Requests that arrive while others run join them: some tick steps more than one."""
    tm = FakeTime()
    be = FakeBackend(tm)
    rs = small_requests(12)
    for r in rs:
        r.arrival = 0.0
    srv = serve.Server(be, 10 ** 12, chunk=50)
    srv.run(rs, clock=tm.clock, sleep=tm.sleep)
    assert max(len(h) for h, _ in be.calls) == 12


def test_modes_step_each_kind_in_its_own_call():
    """With per-kind single-launch modes each tick steps the dendrites and the CH
    requests in separate calls, in their modes; every request still runs exactly
    its steps."""
    tm = FakeTime()
    be = FakeBackend(tm)
    rs = small_requests(16)
    for r in rs:
        r.arrival = 0.0
    srv = serve.Server(be, 10 ** 12, chunk=30, modes={"ch": 2, "dendrite": 1})
    done = srv.run(rs, clock=tm.clock, sleep=tm.sleep)
    by = {r.rid: r for r in rs}
    assert all(c.steps == by[c.rid].steps for c in done) and len(done) == len(rs)
    assert set(be.modes) == {1, 2}
    assert max(len(h) for h, _ in be.calls) < len(rs)   # never both kinds in one call


def test_request_larger_than_budget_is_rejected():
    tm = FakeTime()
    be = FakeBackend(tm)
    rs = small_requests(3)
    srv = serve.Server(be, 100, chunk=10)
    with pytest.raises(ValueError, match="budget"):
        srv.run(rs, clock=tm.clock, sleep=tm.sleep)


def test_rate_steering_raises_rate_when_memory_is_idle():
    tm = FakeTime()
    be = FakeBackend(tm)
    rs = small_requests(60)
    cap = 40 * max(r.cells for r in rs) * 40
    srv = serve.Server(be, cap, chunk=20, target_occupancy=0.8, gain=5.0)
    done = srv.run(rs, rate=20.0, clock=tm.clock, sleep=tm.sleep)
    assert len(done) == 60
    assert srv.rate_trace and max(r for _, r in srv.rate_trace) > 20.0
    assert all(20.0 / 8 - 1e-9 <= r <= 8 * 20.0 + 1e-9 for _, r in srv.rate_trace)
    arr = [r.arrival for r in rs]
    assert all(a < b for a, b in zip(arr, arr[1:]))


def test_report_fields():
    c = [serve.Completion(i, "ch", 0.0, 0.0, float(i + 1), 10, 100, 500) for i in range(10)]
    rep = serve.report(c, 10.0, [(1.0, 0.5), (2.0, 0.9)])
    assert rep["completed"] == 10
    assert rep["p50_latency_s"] == 5.0 and rep["p99_latency_s"] == 10.0
    assert rep["active_rate"] == 500.0 and rep["full_rate"] == 1000.0
    assert abs(rep["occupancy_mean"] - 0.7) < 1e-12


def test_spread_order_covers_and_spreads():
    for n in (1, 2, 7, 300, 1024):
        o = serve.spread_order(n)
        assert sorted(o) == list(range(n))
    o = serve.spread_order(300)
    # any prefix of k entries reaches across the stream: its largest gap is about 2 n / k
    for k in (8, 16, 50):
        pre = sorted(o[:k])
        gaps = np.diff([0] + pre + [300])
        assert gaps.max() <= 2 * 300 // k + 2, (k, gaps.max())
