"""Serving loop on the CPU: the native admission policy (pure function of the
C ABI, no device) and the open-loop driver's latency accounting against a
synthetic-clock server whose schedule is known in closed form."""
import math

import numpy as np
import pytest

from bvh_amd import serving as S


# ---- admission policy -----------------------------------------------------------

def test_form_group_fifo_prefix_that_fits():
    # FIFO prefix stops at the first request that does not fit (no skipping).
    assert S.form_group([10, 10, 50, 1], [1, 1, 1, 1], [1, 1, 1, 1], 30, 64, 1 << 24) == 2


def test_form_group_limits():
    assert S.form_group([1] * 10, [1] * 10, [1] * 10, 100, 4, 1 << 24) == 4
    assert S.form_group([1] * 5, [100, 100, 100, 100, 100], [1] * 5, 100, 64, 250) == 2
    # an oversized head runs alone
    assert S.form_group([1, 1], [1000, 1], [1, 1], 100, 64, 250) == 1


def test_form_group_waits():
    assert S.form_group([1, 1], [1, 1], [0, 1], 100, 64, 1 << 24) == 0  # head not runnable
    assert S.form_group([200, 1], [1, 1], [1, 1], 100, 64, 1 << 24) == 0  # head does not fit
    assert S.form_group([1, 1, 1], [1, 1, 1], [1, 0, 1], 100, 64, 1 << 24) == 1


def test_footprint_grows_with_size():
    t1, r1 = S.footprint(1000, 100, S.ALGO_LBVH)
    t2, r2 = S.footprint(2000, 100, S.ALGO_LBVH)
    t3, _ = S.footprint(1000, 100, S.ALGO_PLOC)
    assert t2 > t1 and r2 > r1 and t3 > t1
    # resident >= mesh (SoA + AoS copy) + BVH nodes + indices + 64-byte traversal records
    n = 1000
    exact = ((n + 63) // 64 * 64) * 36 + 48 * n + (48 + 64) * (2 * n - 1) + 4 * n
    assert exact <= r1 <= exact * 1.2


# ---- driver + statistics on a synthetic clock ---------------------------------------

class FakeServer:
    """Single FIFO server with deterministic service times on a synthetic
    clock: done_i = max(arrival_i, done_{i-1}) + service(i)."""

    def __init__(self, service):
        self.t = 0.0
        self.service = service
        self.free_at = 0.0
        self.pending = []

    def now(self):
        return self.t

    def wait_until(self, t):
        self.t = max(self.t, t)

    def submit(self, rid, mesh, rays, key=0, arrival=None):
        start = max(arrival, self.free_at)
        self.free_at = start + self.service(rid)
        self.pending.append({"id": rid, "status": 0, "arrival": arrival, "admitted": start, "done": self.free_at})
        return None, None

    def poll(self, wait_ms=0):
        if wait_ms:
            self.t = max(self.t, max(c["done"] for c in self.pending) if self.pending else self.t)
        out = [c for c in self.pending if c["done"] <= self.t]
        self.pending = [c for c in self.pending if c["done"] > self.t]
        return out


def test_open_loop_latency_accounting():
    srv = FakeServer(lambda rid: 2.0)
    reqs = [{"id": i, "mesh": None, "rays": None, "key": 0} for i in range(4)]
    done, _, t0 = S.run_open_loop(srv, reqs, [0.0, 1.0, 1.5, 4.0], lead=0.0)
    lat = [c["done"] - c["arrival"] for c in done]
    assert lat == pytest.approx([2.0, 3.0, 4.5, 4.0])
    st = S.latency_stats(done, t0)
    assert st["p50_ms"] == pytest.approx(3000.0)  # nearest rank: 2nd of 4
    assert st["p99_ms"] == pytest.approx(4500.0)
    assert st["throughput_rps"] == pytest.approx(4 / 8.0)


def test_open_loop_md1_queue_matches_lindley():
    rng = np.random.default_rng(5)
    svc = rng.uniform(0.001, 0.003, 300)
    arr = S.poisson_arrivals(400.0, 300, 23)
    srv = FakeServer(lambda rid: float(svc[rid]))
    reqs = [{"id": i, "mesh": None, "rays": None, "key": 0} for i in range(300)]
    done, _, t0 = S.run_open_loop(srv, reqs, arr, lead=0.0)
    d = 0.0
    for i in range(300):
        d = max(arr[i], d) + svc[i]
        assert done[i]["done"] - t0 == pytest.approx(d)


def test_poisson_arrivals_rate():
    a = S.poisson_arrivals(100.0, 20000, 23)
    assert a[0] == 0.0 and np.all(np.diff(a) >= 0)
    assert 20000 / a[-1] == pytest.approx(100.0, rel=0.03)


def test_log_uniform_sizes():
    v = S.log_uniform_ints(np.random.default_rng(1), 50000, 10_000, 2_000_000)
    assert v.min() >= 10_000 and v.max() < 2_000_000
    # median of a log-uniform is the geometric mean of the bounds
    assert np.median(v) == pytest.approx(math.sqrt(10_000 * 2_000_000), rel=0.05)


def test_percentile_nearest_rank():
    v = list(range(1, 101))
    assert S.percentile_nearest_rank(v, 0.5) == 50
    assert S.percentile_nearest_rank(v, 0.99) == 99
    assert S.percentile_nearest_rank([7.0], 0.99) == 7.0


def _footprint(algo, n, k=1, rays=0):
    import ctypes as C
    from bvh_amd import _native as N
    f = N.Footprint()
    assert N.hip().bvh_footprint_estimate(algo, n, k, rays, C.byref(f)) == 0
    return f


@pytest.mark.parametrize("n,k", [(1, 1), (1000, 1), (65537, 1), (1_000_000, 1), (4_000_000, 1), (200_000, 7),
                                 (20_000_000, 512)])
def test_footprint_covers_the_reserve_sizes(n, k):
    """The estimate is at least the buffers a build allocates, counted by hand
    from their element sizes: LBVH boxes 32 + keys 2x8 + handoff 8 + hierarchy
    units 2x48 B per triangle; PLOC boxes 32 + keys 16 + handoff 8 + cluster
    boxes 2x32 + ids 2x4 + neighbour 4; LBVH+ adds 60-bit payloads 2x4, parents
    8, ranges 8, counters 4.  Trees: 2n-1 nodes of 48 B + 4 B indices + 64 B
    traversal records; meshes 36 B SoA + 48 B AoS."""
    from bvh_amd import ALGO_LBVH, ALGO_LBVH_PLUS, ALGO_PLOC
    lo = {ALGO_LBVH: 32 + 16 + 8 + 96, ALGO_PLOC: 32 + 16 + 8 + 64 + 8 + 4, ALGO_LBVH_PLUS: 32 + 16 + 8 + 96 + 28}
    for algo, per in lo.items():
        if algo == ALGO_LBVH_PLUS and k > 1:
            continue
        f = _footprint(algo, n, k)
        assert f.build_workspace >= per * n
        assert f.build_workspace <= per * n * 1.2 + (8 << 20) + k * 64 * 1024   # slack, chunk rounding, tables
        assert f.tree >= (2 * n - 1) * (48 + 64) + 4 * n
        assert f.mesh >= 84 * n
    # Monotone in the triangle count.
    assert _footprint(ALGO_LBVH, n + 5000, k).build_workspace >= _footprint(ALGO_LBVH, n, k).build_workspace


@pytest.mark.parametrize("rays", [1, 1000, 16383, 16384, 1_048_576])
def test_footprint_covers_trace_buffers(rays):
    from bvh_amd import ALGO_LBVH
    f = _footprint(ALGO_LBVH, 0, 1, rays)
    # rays 24 + t 4 + prim 4 + counters 12 + overflow list 4 (+ two 8-byte sort keys from 16384 rays on)
    per = 48 + (16 if rays >= 16384 else 0)
    assert f.trace_workspace >= per * rays


def test_server_rejects_an_unknown_trace_order():
    # Checked before any device is touched.
    from bvh_amd import NativeError
    with pytest.raises(NativeError) as ei:
        S.Server(algo=S.ALGO_LBVH, trace_order=5)
    assert ei.value.status == 1
