"""Serving: continuous batching, the matrix cache's budget, and equality of a
request served alone and inside a mixed batch.

CPU tests drive the scheduler with a stand-in lane (numpy MLEM on a small
dense matrix per geometry) to check its invariants: every request gets
exactly its iteration count, slots are reused, queued requests join at chunk
boundaries, idle lanes are evicted LRU, oversize lanes are rejected.  GPU
tests run real lanes and compare images bit for bit.
"""
import numpy as np
import pytest

from raptr_amd import geometry as G
from raptr_amd import serving


# ------------------------------------------------------------ CPU stand-in
class _Stream:
    handle = None

    def sync(self):
        pass


class FakeLane(serving.Lane):
    """Same slot/queue logic as Lane; iterations are FP64 numpy MLEM."""
    log = []

    def __init__(self, geometry, frames, stream=None, autotune=True):
        rng = np.random.default_rng(abs(hash(geometry)) % 2**32)
        self.geometry = geometry
        self.frames = frames
        self.stream = _Stream()
        self.m, self.n = geometry.nchannels, geometry.nvox
        self.A = rng.random((self.m, self.n)) * (rng.random((self.m, self.n)) < 0.3)
        self.s = self.A.sum(0)
        self.Y = np.zeros((self.m, frames))
        self.X = np.zeros((self.n, frames))
        self.iters = [0] * frames
        self.slots = [None] * frames
        self.remaining = [0] * frames
        self.admitted = [0.0] * frames
        from collections import deque
        self.queue = deque()
        self.bytes = serving.estimate_lane_bytes(geometry, frames, 1000)
        self.closed = False
        self.partial = frames == serving.GROUPED_FRAMES
        self.active_log = []
        FakeLane.log.append(("build", geometry))

    def load(self, slot, req, now):
        serving.check_request(req)
        self.Y[:, slot] = 0
        self.Y[req.channels, slot] = req.counts
        self.X[:, slot] = 1.0 / self.n if req.x0 is None else req.x0
        self.iters[slot] = 0
        self.slots[slot] = req
        self.remaining[slot] = req.niter
        self.admitted[slot] = now

    def launch(self, k):
        a = self.active()
        # every busy slot lies in the frames the pass computes
        assert all(r is None for r in self.slots[a:])
        self.active_log.append(a)
        for _ in range(k):
            for f in range(a):
                self.X[:, f] = mlem_step(self.A, self.s, self.Y[:, f], self.X[:, f])
                if self.slots[f] is not None:
                    self.iters[f] += 1
        return a

    def _move(self, src, dst):
        for arr in (self.Y, self.X):
            arr[:, dst] = arr[:, src]
            arr[:, src] = 0
        self.iters[dst], self.iters[src] = self.iters[src], 0

    def retire(self, k, now):
        out = []
        for s, req in enumerate(self.slots):
            if req is None:
                continue
            self.remaining[s] -= k
            if self.remaining[s] > 0:
                continue
            assert self.iters[s] == req.niter
            out.append(serving.Result(req.rid, self.X[:, s].copy(), req.arrival,
                                      self.admitted[s], now, req.niter))
            self.slots[s] = None
            self.Y[:, s] = 0
            self.X[:, s] = 0
        return out

    def close(self):
        self.closed = True
        FakeLane.log.append(("evict", self.geometry))


def mlem_step(A, s, y, x):
    yt = A @ x
    e = np.divide(y, yt, out=np.zeros_like(y), where=(yt != 0) & (y != 0))
    c = A.T @ e
    d = s + (s == 0)
    xn = x * c / d
    tot = xn.sum()
    return xn / tot if tot != 0 else np.zeros_like(xn)


def _geometry(npix, na, nvox):
    setup = G.two_head_setup(npix, na, 36.0)
    grid = G.centred_grid(nvox, 2.0)
    return serving.Geometry(setup, grid, 4)


def _requests(geoms, n, seed=0, iters=(1, 9)):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        g = geoms[rng.integers(len(geoms))]
        ch = np.sort(rng.choice(g.nchannels, size=g.nchannels // 2, replace=False)).astype(np.int32)
        out.append(serving.Request(i, g, ch, rng.poisson(20, len(ch)).astype(np.float32),
                                   int(rng.integers(iters[0], iters[1] + 1)),
                                   arrival=float(i) * 1e-3))
    return out


@pytest.fixture
def fake(monkeypatch):
    FakeLane.log = []
    monkeypatch.setattr(serving, "Lane", FakeLane)
    monkeypatch.setattr(serving, "estimate_nnz", lambda g, **k: 1000)
    return FakeLane


def test_each_request_gets_its_iterations_alone_or_mixed(fake):
    geoms = [_geometry(2, 2, 3), _geometry(2, 3, 2)]
    reqs = _requests(geoms, 23)
    cache = serving.MatrixCache(1 << 40, frames=4)
    stats = serving.Server(cache, max_chunk=3).serve(reqs, realtime=False)
    assert len(stats.results) == 23 and all(r.error is None for r in stats.results)
    got = {r.rid: r.image for r in stats.results}
    for req in reqs:
        lane = cache.get(req.geometry)
        x = np.full(lane.n, 1.0 / lane.n)
        y = np.zeros(lane.m)
        y[req.channels] = req.counts
        for _ in range(req.niter):
            x = mlem_step(lane.A, lane.s, y, x)
        np.testing.assert_array_equal(got[req.rid], x)
    # 23 requests through 4 slots per lane: slots were reused and requests
    # joined running batches
    assert stats.chunks > 23 // 8
    s = stats.summary()
    assert s["completed"] == 23 and s["p99_ms"] >= s["p50_ms"]


def test_64_slot_lanes_shrink_to_16_or_32_frames(fake):
    """64-slot lanes: 40 requests at once run 64-frame passes, then, as
    requests finish, the remaining ones move into the lowest slots and the
    lane runs 32- and 16-frame passes; every request still gets exactly its
    iterations and the image it gets alone."""
    g = _geometry(3, 2, 4)
    reqs = _requests([g], 40, seed=3, iters=(1, 12))
    srv = serving.Server(serving.MatrixCache(1 << 40, 64), max_chunk=2)
    st = srv.serve(reqs, realtime=False)
    mixed = {r.rid: r for r in st.results}
    log = srv.cache.get(g).active_log
    assert log[0] == 64 and 32 in log and log[-1] == 16
    assert st.moves > 0 and st.summary()["chunks_by_active_frames"][16] > 0
    for req in reqs[:6]:
        alone = srv.serve([req], realtime=False).results[0]
        np.testing.assert_allclose(mixed[req.rid].image, alone.image, rtol=1e-12)


def test_invalid_requests_are_rejected(fake):
    g = _geometry(2, 2, 3)
    good = _requests([g], 2)
    bad = serving.Request(99, g, np.array([0, g.nchannels], np.int32),
                          np.ones(2, np.float32), 3)
    zero = serving.Request(98, g, np.array([0], np.int32), np.ones(1, np.float32), 0)
    stats = serving.Server(serving.MatrixCache(1 << 40, 4)).serve(good + [bad, zero],
                                                                  realtime=False)
    err = {r.rid: r.error for r in stats.results}
    assert err[0] is None and err[1] is None
    assert "outside" in err[99] and "niter" in err[98]


def test_budget_evicts_idle_lanes_lru_and_rejects_oversize(fake):
    g1, g2, g3 = _geometry(2, 2, 3), _geometry(2, 3, 2), _geometry(2, 4, 2)
    one = serving.estimate_lane_bytes(g1, 4, 1000)
    cache = serving.MatrixCache(int(2.5 * one), frames=4)
    srv = serving.Server(cache)
    srv.serve(_requests([g1], 3), realtime=False)
    srv.serve(_requests([g2], 3), realtime=False)
    assert set(cache.lanes) == {g1, g2}
    srv.serve(_requests([g1], 2), realtime=False)  # g1 most recently used
    srv.serve(_requests([g3], 2), realtime=False)  # needs room: evicts g2 (LRU)
    assert set(cache.lanes) == {g1, g3} and cache.evicted == 1
    assert ("evict", g2) in FakeLane.log
    tiny = serving.MatrixCache(one // 2, frames=4)
    st = serving.Server(tiny).serve(_requests([g1], 2), realtime=False)
    assert all(r.error and "budget" in r.error for r in st.results)


def test_realtime_arrivals_wait_for_their_time(fake):
    g = _geometry(2, 2, 3)
    reqs = _requests([g], 4)
    for i, r in enumerate(reqs):
        r.arrival = 0.02 * i
    stats = serving.Server(serving.MatrixCache(1 << 40, 4)).serve(reqs, realtime=True)
    for r in stats.results:
        assert r.admitted >= r.arrival - 1e-9 and r.finished >= r.admitted


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("frames", [4, 64])
def test_gpu_alone_equals_mixed_and_matches_oracle(gpu, oracle, frames):
    """Requests on the 20,480-channel geometry: each image from a mixed,
    continuously batched run equals the image of the request served alone
    (bit for bit), and one matches the FP64 oracle (1e-4).  64 slots: lanes
    with grouped layouts only, on the matrix cores."""
    from conftest import oracle_geom
    setup, grid, nrays = G.config_small()
    geo = serving.Geometry(setup, grid, nrays)
    from raptr_amd import synthetic
    s, g = oracle_geom(oracle, setup, grid)
    allc = np.arange(setup.nchannels, dtype=np.int32)
    rp, col, val, v64 = oracle.build_csr(s, g, allc, nrays, with_f64=True)
    xt = synthetic.spheres_phantom(grid)
    ybar = oracle.csrmv_f64(0, len(allc), grid.nvox, 1.0, v64, rp, col, xt, 0.0,
                            np.zeros(len(allc)))
    reqs = serving.synthetic_requests([geo], 100.0, 7, seed=4, events=(1e4, 1e6),
                                      iters=(3, 12), ybar={geo: ybar})
    cache = serving.MatrixCache(40 << 30, frames=frames)
    srv = serving.Server(cache, max_chunk=4)
    mixed = {r.rid: r for r in srv.serve(reqs, realtime=False).results}
    assert all(r.error is None for r in mixed.values())
    for req in reqs:
        alone = srv.serve([req], realtime=False).results[0]
        np.testing.assert_array_equal(mixed[req.rid].image, alone.image)
    req = reqs[0]
    y = np.zeros(len(allc))
    y[req.channels] = req.counts
    sens64 = oracle.sensitivity_f64(rp, col, v64, grid.nvox)
    x64 = oracle.mlem_f64(rp, col, v64, y, sens64, np.full(grid.nvox, 1.0 / grid.nvox),
                          req.niter)
    rel = np.linalg.norm(mixed[req.rid].image - x64) / np.linalg.norm(x64)
    assert rel <= 1e-4


@pytest.mark.gpu
def test_gpu_lane_shrinks_to_active_frames(gpu):
    """A 64-slot lane on the 20,480-channel geometry with 40 requests at
    once: it runs 64-, 32- and 16-frame passes as requests finish and the
    rest move into the lowest slots, and every image equals the image of
    the request served alone (a 16-frame pass) bit for bit."""
    from raptr_amd import synthetic, sysmat
    setup, grid, nrays = G.config_small()
    geo = serving.Geometry(setup, grid, nrays)
    img, _ = synthetic.random_ellipsoids(grid, n=4, seed=3)
    ybar = sysmat.forward_project(setup, grid, img.astype(np.float32), nrays)
    reqs = serving.synthetic_requests([geo], 1000.0, 40, seed=6, events=(1e4, 1e6),
                                      iters=(2, 14), ybar={geo: ybar})
    srv = serving.Server(serving.MatrixCache(40 << 30, frames=64), max_chunk=2)
    st = srv.serve(reqs, realtime=False)
    summ = st.summary()
    print(summ["chunks_by_active_frames"], "moves", summ["slot_moves"])
    assert set(summ["chunks_by_active_frames"]) == {16, 32, 64} and summ["slot_moves"] > 0
    mixed = {r.rid: r for r in st.results}
    assert all(r.error is None for r in mixed.values())
    for req in reqs[::5]:
        alone = srv.serve([req], realtime=False).results[0]
        np.testing.assert_array_equal(mixed[req.rid].image, alone.image)


@pytest.mark.gpu
def test_gpu_geometry_mix_alone_equals_mixed(gpu, oracle):
    """The serving workload's three geometries (8x8 heads / 32^3, 13x13 /
    52^3, 16x16 / 96^3; 2 mm voxels; angle counts cut to 12, 6 and 4 to keep
    the test short), 64-slot lanes on the matrix cores, requests of 1e4-1e6
    events and 3-12 iterations arriving together: every image equals the
    image of the same request served alone, bit for bit, across geometries;
    the chunks of different lanes overlap on the device (HIP event
    timestamps on the lanes' streams); one request of the smallest geometry
    matches the FP64 oracle (1e-4)."""
    from conftest import oracle_geom
    from raptr_amd import synthetic, sysmat
    geos = [serving.Geometry(G.two_head_setup(p, na, 180.0 / na), G.centred_grid(v, 2.0), 10)
            for p, v, na in ((8, 32, 12), (13, 52, 6), (16, 96, 4))]
    ybar = {}
    for g in geos:
        img, _ = synthetic.random_ellipsoids(g.grid, n=4, seed=5)
        ybar[g] = sysmat.forward_project(g.setup, g.grid, img.astype(np.float32), g.nrays)
    reqs = serving.synthetic_requests(geos, 100.0, 9, seed=8, events=(1e4, 1e6), iters=(3, 12),
                                      ybar=ybar)
    assert len({r.geometry for r in reqs}) == 3
    cache = serving.MatrixCache(200 << 30, frames=64)
    srv = serving.Server(cache, max_chunk=3, trace=True)
    st = srv.serve(reqs, realtime=False)
    mixed = {r.rid: r for r in st.results}
    assert all(r.error is None for r in mixed.values())
    # overlap: some chunk of one lane starts before a chunk of another ends
    ov = serving.overlapping_chunks(st.trace)
    print(f"chunks {len(st.trace)}, overlapping pairs of different lanes {ov}")
    assert ov > 0
    for req in reqs:
        alone = srv.serve([req], realtime=False).results[0]
        np.testing.assert_array_equal(mixed[req.rid].image, alone.image)
    # FP64 oracle for a request of the 8x8 / 32^3 geometry
    req = next(r for r in reqs if r.geometry is geos[0])
    setup, grid, nrays = geos[0].setup, geos[0].grid, geos[0].nrays
    s, g = oracle_geom(oracle, setup, grid)
    allc = np.arange(setup.nchannels, dtype=np.int32)
    rp, col, val, v64 = oracle.build_csr(s, g, allc, nrays, with_f64=True)
    y = np.zeros(len(allc))
    y[req.channels] = req.counts
    x64 = oracle.mlem_f64(rp, col, v64, y, oracle.sensitivity_f64(rp, col, v64, grid.nvox),
                          np.full(grid.nvox, 1.0 / grid.nvox), req.niter)
    rel = np.linalg.norm(mixed[req.rid].image - x64) / np.linalg.norm(x64)
    assert rel <= 1e-4
    for geo in list(cache.lanes):
        cache.lanes.pop(geo).close()


@pytest.mark.gpu
def test_gpu_cache_budget_and_rejection(gpu):
    tiny = serving.Geometry(*G.config_tiny())
    small = serving.Geometry(*G.config_small())
    cache = serving.MatrixCache(1 << 30, frames=2)
    lane = cache.admit(tiny)
    assert lane.bytes < 1 << 30
    with pytest.raises(MemoryError):
        serving.MatrixCache(1 << 20, frames=2).admit(small)
