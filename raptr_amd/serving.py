"""Serving of MLEM reconstruction requests on one GPU: continuous batching
by geometry, a device matrix cache under an HBM budget, and latency
figures.

The reference runs one reconstruction per process (reco.cu:230-321) and
rebuilds the system matrix every iteration.  Here each geometry (setup,
grid, rays, value type) gets a *lane*: its matrix over all channels, built
once and kept in HBM as a batched reconstruction plan with F frame slots.
A request occupies one slot.  Its counts go into column `slot` of y
([channels][F]), its start image into column `slot` of x ([voxels][F]), and
the lane advances all occupied slots together in chunks of iterations (a
HIP graph replay).  Frames never interact: every per-frame quantity
(forward sums, ratio, back-projection, FP64 normalisation) is computed with
the same arithmetic whatever the other slots hold, so a request gives the
same image alone or in a mixed batch.  A chunk never runs past the fewest
remaining iterations of the lane's requests, so each request gets exactly
its iteration count.  Requests join free slots at chunk boundaries
(continuous batching) and leave when done.  A 64-slot lane keeps its busy
slots in the lowest frames (a request moves down when a lower slot frees
and that lets the lane shrink) and runs each chunk over the first 16, 32
or 64 frames only, the fewest that hold every busy slot
(raptr_recon_mlem_active): a pass costs about in proportion to the frames
it gathers, so a lightly loaded lane iterates faster.

The cache admits a lane only if its estimated device bytes fit the budget.
It evicts idle lanes (no request in a slot or queued) least recently used
first.  A request whose lane cannot fit even after eviction is rejected with
an error result.  Lanes run on their own streams, so chunks of different
geometries overlap on the GPU.
"""
import ctypes as C
import heapq
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field

import numpy as np

from . import _capi as capi
from . import sysmat
from .device import DeviceArray, Event, Stream
from .geometry import MeasurementSetup, VoxelGrid
from .recon import Reconstructor

# device bytes per matrix entry of a plan: forward value + column, transposed
# value + row (FP32), and the build's CSR while the plan is made from it
_PLAN_BYTES_PER_NNZ = {"f32": 16, "u16": 12, "f16": 12, "bf16": 12}
# 64-slot lanes keep only the grouped layouts (16 rows / 4x1x4 voxel bricks
# per union entry; ~0.16 and ~0.26 union entries per matrix entry, each a
# 4-byte key and 16 values)
_GROUPED_BYTES_PER_NNZ = {"f32": 30, "u16": 16, "f16": 16, "bf16": 16}
_BUILD_BYTES_PER_NNZ = 8
GROUPED_FRAMES = 64


@dataclass(frozen=True)
class Geometry:
    setup: MeasurementSetup
    grid: VoxelGrid
    nrays: int
    seed: int = 1
    raygen: int = capi.RAYGEN_RAND
    value_type: str = "f32"

    @property
    def nchannels(self):
        return self.setup.nchannels

    @property
    def nvox(self):
        return self.grid.nvox


@dataclass
class Request:
    """counts[i] events in channel channels[i] (measured channels, any
    order), niter MLEM iterations from x0 (uniform 1/N if None)."""
    rid: int
    geometry: Geometry
    channels: np.ndarray
    counts: np.ndarray
    niter: int
    arrival: float = 0.0
    x0: np.ndarray = None


@dataclass
class Result:
    rid: int
    image: np.ndarray
    arrival: float
    admitted: float
    finished: float
    niter: int
    error: str = None

    @property
    def latency(self):
        return self.finished - self.arrival


def check_request(req):
    g = req.geometry
    if req.niter < 1:
        raise ValueError(f"request {req.rid}: niter must be >= 1")
    ch = np.asarray(req.channels)
    if len(ch) != len(req.counts):
        raise ValueError(f"request {req.rid}: {len(ch)} channels but {len(req.counts)} counts")
    if len(ch) and (ch.min() < 0 or ch.max() >= g.nchannels):
        raise ValueError(f"request {req.rid}: channel ids outside [0, {g.nchannels})")
    if req.x0 is not None and np.asarray(req.x0).size != g.nvox:
        raise ValueError(f"request {req.rid}: x0 has {np.asarray(req.x0).size} voxels, "
                         f"grid has {g.nvox}")


def estimate_nnz(geometry, sample_every=64, stream=None):
    """Non-zero count of the all-channel matrix from a count pass over every
    sample_every-th channel."""
    g = geometry
    sample = np.arange(0, g.nchannels, sample_every, dtype=np.int32)
    d_s = DeviceArray.from_numpy(sample)
    s, gr = g.setup.struct(), g.grid.struct()
    L = capi.lib()
    ws = DeviceArray(max(L.raptr_sm_build_workspace_size(C.byref(gr), g.nrays, 1 << 20), 1),
                     np.uint8)
    rp = DeviceArray(len(sample) + 1, np.int64)
    nnz = C.c_int64()
    capi.call("raptr_sm_build_csr", C.byref(s), C.byref(gr), d_s.ptr, len(sample), g.nrays,
              g.seed, g.raygen, rp.ptr, None, None, 0, C.byref(nnz), ws.ptr, ws.nbytes, stream)
    return int(nnz.value * g.nchannels / max(len(sample), 1))


def estimate_lane_bytes(geometry, frames, nnz):
    g = geometry
    if frames == GROUPED_FRAMES:
        per = _GROUPED_BYTES_PER_NNZ[g.value_type] + _BUILD_BYTES_PER_NNZ
        extra = 4 * 8 * frames * g.nvox  # per-XCD partial images
    else:
        per = _PLAN_BYTES_PER_NNZ[g.value_type] + _BUILD_BYTES_PER_NNZ
        extra = 0
    return int(per * nnz + 8 * 3 * (g.nchannels + 1) + 4 * frames * (3 * g.nchannels + 4 * g.nvox)
               + 16 * g.nvox + extra)


class Lane:
    """One geometry on the device: plan with `frames` slots and its y/x."""

    def __init__(self, geometry, frames, stream=None, autotune=True):
        g = geometry
        self.geometry = g
        self.frames = frames
        self.stream = stream or Stream()
        st = self.stream.handle
        t = time.perf_counter()
        allc = np.arange(g.nchannels, dtype=np.int32)
        A = sysmat.build_csr(g.setup, g.grid, allc, g.nrays, seed=g.seed, raygen=g.raygen,
                             stream=st, capacity="auto")
        sens = sysmat.sensitivity(g.setup, g.grid, g.nrays, seed=g.seed, raygen=g.raygen,
                                  stream=st)
        if frames == GROUPED_FRAMES:
            # 64 slots: rows in locality order, grouped layouts only (16 rows
            # / 4x1x4 voxel bricks on the matrix cores)
            self.plan = Reconstructor(A, sens, g.value_type, frames, stream=st,
                                      row_order=g.setup, grid=g.grid, group=16,
                                      grouped_only=True)
        else:
            self.plan = Reconstructor(A, sens, g.value_type, frames, stream=st)
        self.nnz = A.nnz
        del A, sens
        self.m, self.n = g.nchannels, g.nvox
        self.y = DeviceArray.zeros(self.m * frames, np.float32, stream=st)
        self.x = DeviceArray.zeros(self.n * frames, np.float32, stream=st)
        self.bad = DeviceArray.zeros(1, np.int32, stream=st)
        self.slots = [None] * frames      # request per slot
        self.remaining = [0] * frames
        self.admitted = [0.0] * frames
        self.queue = deque()
        if autotune:
            self.config, _ = self.plan.autotune(self.y, self.x, frames, reps=2, stream=st)
        self.stream.sync()
        self.build_s = time.perf_counter() - t
        self.bytes = self.plan.info()["device_bytes"] + 4 * frames * (self.m + self.n)
        gi = self.plan.grouped() if frames == GROUPED_FRAMES else {}
        # passes over 16 / 32 of the 64 frames: grouped layouts of 16, kind 3
        self.partial = bool(gi) and gi["on"] == 1 and gi["group"] == 16 and gi["kind"] == 3

    # ---------------------------------------------------------- slots
    def free_slots(self):
        return [k for k, r in enumerate(self.slots) if r is None]

    def busy(self):
        return any(r is not None for r in self.slots) or bool(self.queue)

    def active(self):
        """Frames the next chunk computes: the fewest of 16, 32, 64 holding
        every busy slot (all frames on lanes without partial passes)."""
        if not self.partial:
            return self.frames
        hi = max((s + 1 for s, r in enumerate(self.slots) if r is not None), default=0)
        return 16 if hi <= 16 else (32 if hi <= 32 else 64)

    def compact(self):
        """Move requests from the highest busy slots into free lower ones
        until the busy slots fit the fewest active frames their count
        allows (a column copy of y and x; each request's frame computes the
        same whichever slot holds it)."""
        if not self.partial:
            return 0
        busy = [s for s, r in enumerate(self.slots) if r is not None]
        target = 16 if len(busy) <= 16 else (32 if len(busy) <= 32 else 64)
        high = [s for s in reversed(busy) if s >= target]
        free = [s for s in range(target) if self.slots[s] is None]
        for src, dst in zip(high, free):
            self._move(src, dst)
            self.slots[dst], self.slots[src] = self.slots[src], None
            self.remaining[dst], self.remaining[src] = self.remaining[src], 0
            self.admitted[dst] = self.admitted[src]
        return min(len(high), len(free))

    def _move(self, src, dst):
        """Columns src of y and x to dst; src cleared."""
        st = self.stream.handle
        for arr, rows in ((self.y, self.m), (self.x, self.n)):
            capi.call("raptr_frame_copy", arr.ptr, rows, self.frames, src, dst, st)
            capi.call("raptr_frame_fill", arr.ptr, rows, self.frames, src, 0.0, st)

    def load(self, slot, req, now):
        st = self.stream.handle
        ch = np.ascontiguousarray(req.channels, np.int32)
        ct = np.ascontiguousarray(req.counts, np.float32)
        check_request(req)
        capi.call("raptr_frame_fill", self.y.ptr, self.m, self.frames, slot, 0.0, st)
        if len(ch):
            d_i = DeviceArray.from_numpy(ch, stream=st)
            d_v = DeviceArray.from_numpy(ct, stream=st)
            capi.call("raptr_frame_scatter", self.y.ptr, self.m, self.frames, slot, d_i.ptr,
                      d_v.ptr, len(ch), self.bad.ptr, st)
            self.stream.sync()  # the staging arrays are freed on return
        if req.x0 is None:
            capi.call("raptr_frame_fill", self.x.ptr, self.n, self.frames, slot,
                      1.0 / self.n, st)
        else:
            x0 = np.ascontiguousarray(req.x0, np.float32).reshape(-1)
            d_v = DeviceArray.from_numpy(x0, stream=st)
            d_i = DeviceArray.from_numpy(np.arange(self.n, dtype=np.int32), stream=st)
            capi.call("raptr_frame_scatter", self.x.ptr, self.n, self.frames, slot, d_i.ptr,
                      d_v.ptr, self.n, self.bad.ptr, st)
            self.stream.sync()
        self.slots[slot] = req
        self.remaining[slot] = req.niter
        self.admitted[slot] = now

    def chunk(self, max_chunk):
        """Iterations the next chunk may run: the fewest remaining (so no
        request runs past its count), at most max_chunk."""
        rem = [r for r, q in zip(self.remaining, self.slots) if q is not None]
        return min(min(rem), max_chunk) if rem else 0

    def launch(self, k):
        """k iterations over the active frames; returns their count."""
        a = self.active()
        self.plan.run(self.y, self.x, k, self.frames, use_graph=True, stream=self.stream.handle,
                      active=a)
        return a

    def retire(self, k, now):
        """After a chunk of k iterations: results of the slots that are done."""
        st = self.stream.handle
        out = []
        for s, req in enumerate(self.slots):
            if req is None:
                continue
            self.remaining[s] -= k
            if self.remaining[s] > 0:
                continue
            img = DeviceArray(self.n, np.float32)
            capi.call("raptr_frame_gather", self.x.ptr, self.n, self.frames, s, img.ptr, st)
            capi.call("raptr_frame_fill", self.y.ptr, self.m, self.frames, s, 0.0, st)
            capi.call("raptr_frame_fill", self.x.ptr, self.n, self.frames, s, 0.0, st)
            image = img.numpy(stream=st)
            out.append(Result(req.rid, image, req.arrival, self.admitted[s], now, req.niter))
            self.slots[s] = None
        return out

    def close(self):
        self.plan.close()
        self.y.free()
        self.x.free()


class MatrixCache:
    """Lanes by geometry under a device-byte budget, idle lanes evicted
    least recently used first."""

    def __init__(self, budget_bytes, frames=16):
        self.budget = int(budget_bytes)
        self.frames = frames
        self.lanes = OrderedDict()
        self.built = 0
        self.evicted = 0

    @property
    def bytes(self):
        return sum(l.bytes for l in self.lanes.values())

    def get(self, geometry):
        lane = self.lanes.get(geometry)
        if lane is not None:
            self.lanes.move_to_end(geometry)
        return lane

    def admit(self, geometry, nnz=None):
        """The lane of `geometry`, built if absent.  Raises MemoryError when
        it cannot fit the budget even with every idle lane evicted."""
        lane = self.get(geometry)
        if lane is not None:
            return lane
        need = estimate_lane_bytes(geometry, self.frames,
                                   nnz if nnz is not None else estimate_nnz(geometry))
        idle = [g for g, l in self.lanes.items() if not l.busy()]
        freeable = sum(self.lanes[g].bytes for g in idle)
        if self.bytes - freeable + need > self.budget:
            raise MemoryError(f"lane needs ~{need / 1e9:.1f} GB; budget {self.budget / 1e9:.1f} GB "
                              f"with {self.bytes / 1e9:.1f} GB held, "
                              f"{freeable / 1e9:.1f} GB evictable")
        for g in idle:
            if self.bytes + need <= self.budget:
                break
            self.lanes.pop(g).close()
            self.evicted += 1
        lane = Lane(geometry, self.frames)
        self.lanes[geometry] = lane
        self.built += 1
        return lane


@dataclass
class ServeStats:
    results: list = field(default_factory=list)
    chunks: int = 0
    wall_s: float = 0.0
    # with Server(trace=True): (lane number, start ms, end ms) per chunk, from
    # HIP events recorded on the lane's stream around the chunk's launch
    # (times relative to the first chunk's start)
    trace: list = field(default_factory=list)
    # chunks per active frame count (16 / 32 / 64-slot passes; F otherwise)
    active_frames: dict = field(default_factory=dict)
    moves: int = 0

    def summary(self):
        ok = [r for r in self.results if r.error is None]
        lat = np.array([r.latency for r in ok]) if ok else np.zeros(0)
        q = (lambda p: float(np.percentile(lat, p)) * 1e3 if len(lat) else None)
        frames_it = sum(r.niter for r in ok)
        return {"requests": len(self.results), "completed": len(ok),
                "rejected": len(self.results) - len(ok),
                "p50_ms": q(50), "p99_ms": q(99), "max_ms": q(100),
                "mean_ms": float(lat.mean()) * 1e3 if len(lat) else None,
                "requests_per_s": len(ok) / self.wall_s if self.wall_s else None,
                "frame_iterations_per_s": frames_it / self.wall_s if self.wall_s else None,
                "chunks": self.chunks, "wall_s": self.wall_s,
                "chunks_by_active_frames": dict(sorted(self.active_frames.items())),
                "slot_moves": self.moves}


class Server:
    """Continuous-batching loop over a MatrixCache."""

    def __init__(self, cache, max_chunk=10, trace=False):
        self.cache = cache
        self.max_chunk = max_chunk
        self.trace = trace

    def _admit_queued(self, lane, now):
        for s in lane.free_slots():
            if not lane.queue:
                break
            lane.load(s, lane.queue.popleft(), now)

    def serve(self, requests, clock=None, realtime=True):
        """Serve `requests` (arrival times in seconds from the start).  With
        realtime, arrivals follow the wall clock; otherwise every request is
        available at once (throughput mode).  Returns ServeStats."""
        clock = clock or time.perf_counter
        pending = [(r.arrival, r.rid, r) for r in requests]
        heapq.heapify(pending)
        stats = ServeStats()
        t0 = clock()
        now = lambda: clock() - t0  # noqa: E731
        events = []  # (lane number, start event, end event)
        lane_no = {}
        while True:
            t = now() if realtime else float("inf")
            while pending and pending[0][0] <= t:
                _, _, req = heapq.heappop(pending)
                try:
                    check_request(req)
                    lane = self.cache.admit(req.geometry)
                except (MemoryError, ValueError) as e:
                    stats.results.append(Result(req.rid, None, req.arrival, now(), now(),
                                                req.niter, error=str(e)))
                    continue
                lane.queue.append(req)
            active = []
            for lane in list(self.cache.lanes.values()):
                self._admit_queued(lane, max(now(), 0.0))
                stats.moves += lane.compact()
                k = lane.chunk(self.max_chunk)
                if k > 0:
                    if self.trace:
                        e0, e1 = Event(), Event()
                        e0.record(lane.stream.handle)
                    a = lane.launch(k)
                    stats.active_frames[a] = stats.active_frames.get(a, 0) + 1
                    if self.trace:
                        e1.record(lane.stream.handle)
                        events.append((lane_no.setdefault(id(lane), len(lane_no)), e0, e1))
                    active.append((lane, k))
            for lane, k in active:
                lane.stream.sync()
                stats.chunks += 1
                stats.results.extend(lane.retire(k, now()))
            if not active:
                if not pending:
                    break
                if realtime:
                    wait = pending[0][0] - now()
                    if wait > 0:
                        time.sleep(min(wait, 0.005))
        stats.wall_s = now()
        if events:
            ref = events[0][1]
            stats.trace = [(ln, ref.elapsed_ms(a), ref.elapsed_ms(b)) for ln, a, b in events]
        return stats


def overlapping_chunks(trace):
    """Pairs of chunks of different lanes whose device time intervals
    overlap (ServeStats.trace)."""
    n = 0
    for i, (li, a0, a1) in enumerate(trace):
        for lj, b0, b1 in trace[i + 1:]:
            if li != lj and a0 < b1 and b0 < a1:
                n += 1
    return n


def synthetic_requests(geometries, rate_hz, n, seed=0, events=(1e4, 5e6), iters=(10, 50),
                       ybar=None):
    """Poisson arrivals at rate_hz of n requests over `geometries` (uniform
    choice), counts log-uniform in `events`, iterations uniform in `iters`.
    ybar[g]: FP64 projection of a phantom for geometry g (all channels)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate_hz)
        g = geometries[rng.integers(len(geometries))]
        total = float(np.exp(rng.uniform(np.log(events[0]), np.log(events[1]))))
        lam = ybar[g] * (total / ybar[g].sum())
        y = np.random.Generator(np.random.Philox(key=seed * 100003 + i)).poisson(lam)
        ch = np.nonzero(y)[0].astype(np.int32)
        out.append(Request(i, g, ch, y[ch].astype(np.float32),
                           int(rng.integers(iters[0], iters[1] + 1)), arrival=t))
    return out
