"""GPU tests of the reconstruction plan: plain FP32 MLEM against the FP64
host recurrence, independence of the row-block cut, batched frames against
single-frame runs, mixed-precision error bounds, graph replay."""
import numpy as np
import pytest

from conftest import oracle_geom
from raptr_amd import geometry as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_system(gpu, oracle):
    """20,480-channel / 32^3 workload: FP64 oracle matrix over all channels,
    two-sphere phantom, 2e5 Poisson events, measured rows, GPU matrix."""
    from raptr_amd import synthetic, sysmat
    setup, grid, nrays = G.config_small()
    s, g = oracle_geom(oracle, setup, grid)
    allc = np.arange(setup.nchannels, dtype=np.int32)
    rp, col, val, v64 = oracle.build_csr(s, g, allc, nrays, with_f64=True)
    xt = synthetic.spheres_phantom(grid)
    ybar = oracle.csrmv_f64(0, len(allc), grid.nvox, 1.0, v64, rp, col, xt, 0.0,
                            np.zeros(len(allc)))
    y = synthetic.poisson_counts(ybar, 2e5, seed=1).astype(np.float64)
    meas = np.nonzero(y > 0)[0].astype(np.int32)
    sens64 = oracle.sensitivity_f64(rp, col, v64, grid.nvox)
    mrp = np.concatenate([[0], np.cumsum(np.diff(rp)[meas])])
    msel = np.concatenate([np.arange(rp[i], rp[i + 1]) for i in meas])
    A = sysmat.build_csr(setup, grid, meas, nrays)
    sens = sysmat.sensitivity(setup, grid, nrays)
    return dict(setup=setup, grid=grid, nrays=nrays, meas=meas, y=y[meas], A=A, sens=sens,
                orp=mrp, ocol=col[msel], oval=v64[msel], sens64=sens64, ybar=ybar)


def _oracle_mlem(oracle, S, y, niter):
    n = S["grid"].nvox
    x0 = np.full(n, 1.0 / n)
    return oracle.mlem_f64(S["orp"], S["ocol"], S["oval"], y, S["sens64"], x0, niter)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("block_rows", [0, 1777])
def test_recon_f32_vs_fp64_oracle(gpu, oracle, small_system, block_rows):
    """Plain FP32 MLEM (one and several row blocks) against the FP64 host
    recurrence."""
    from raptr_amd.device import Stream
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    x64 = _oracle_mlem(oracle, S, S["y"], 5)
    R = Reconstructor(S["A"], S["sens"], "f32", 1, block_rows=block_rows)
    if block_rows:
        assert R.info()["row_blocks"] > 5
    x0 = np.full(n, 1.0 / n, np.float32)
    st = Stream()
    for graph in (False, True):
        x = R.mlem(S["y"], x0, 5, use_graph=graph, stream=st.handle if graph else None)
        assert _rel(x, x64) <= 1e-4
        assert abs(float(x.astype(np.float64).sum()) - 1.0) < 1e-5


@pytest.mark.parametrize("block_rows", [0, 800, 5000])
def test_recon_backward_per_xcd_row_ranges(gpu, oracle, small_system, block_rows):
    """Single-frame backward mode 3 (each XCD takes an eighth of the row
    blocks, per-XCD partial column sums added in a fixed order; 16, 32 or 64
    lanes per column), with fewer blocks than XCDs (block_rows 0 and 5000)
    and many (800): against the FP64 oracle, the block-by-block backward
    (1e-6) and itself (bit for bit)."""
    from raptr_amd import RaptrError
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    x0 = np.full(n, 1.0 / n, np.float32)
    x64 = _oracle_mlem(oracle, S, S["y"], 5)
    R = Reconstructor(S["A"], S["sens"], "f32", 1, block_rows=block_rows, layout="blocked",
                      row_order=S["setup"])
    R.set_fused(False)  # the separate forward and backward passes
    R.set_backward(0)
    b = R.mlem(S["y"], x0, 5)
    from raptr_amd.device import Stream
    st = Stream()
    for unroll, lpc in ((1, 64), (2, 64), (4, 64), (1, 16), (2, 32), (4, 16)):
        R.set_config(16, 1, unroll)
        R.set_backward(3, lpc)
        a1 = R.mlem(S["y"], x0, 5, use_graph=False)
        for _ in range(2):  # graph replays (eager work between them)
            a2 = R.mlem(S["y"], x0, 5, use_graph=True, stream=st.handle)
            np.testing.assert_array_equal(a2, a1)
        assert _rel(a1, x64) <= 1e-4
        assert _rel(a1, b) <= 1e-6
        # 16-bit transpose rows / forward columns: the same sums
        if R.info()["compact_rows"] >= 0:
            R.set_compact_rows(False)
            a3 = R.mlem(S["y"], x0, 5, use_graph=False)
            R.set_compact_rows(True)
            np.testing.assert_array_equal(a3, a1)
    # 16-bit rows exist for blocks of <= 65,536 rows (the single-frame
    # default); 16-bit forward columns for <= 262,144 voxels (32^3 here)
    assert R.info()["compact_rows"] == ((1 if block_rows <= 65536 else 0) | 2)
    with pytest.raises(RaptrError):
        Reconstructor(S["A"], S["sens"], "f32", 2).set_backward(3)


def test_recon_backward_modes(gpu, small_system):
    """Fewer than 8 row blocks: the block-by-block backward is the default;
    the per-XCD pass is still allowed there.  Only modes 0 and 3 exist, the
    per-XCD pass is single-frame only, and lanes per column are 8-64."""
    from raptr_amd import RaptrError
    from raptr_amd.recon import Reconstructor
    S = small_system
    R = Reconstructor(S["A"], S["sens"], "f32", 1, block_rows=30000)
    assert R.info()["row_blocks"] < 8
    assert R.info()["bwd_staged"] == 0  # per-XCD row ranges need >= 8 blocks
    R.set_backward(True)
    assert R.info()["bwd_staged"] == 3
    for bad in ((1, 16), (2, 16), (4, 16), (3, 12)):
        with pytest.raises(RaptrError):
            R.set_backward(*bad)
    Rb = Reconstructor(S["A"], S["sens"], "f32", 2)
    with pytest.raises(RaptrError):
        Rb.set_backward(True)


def test_recon_row_block_cut_is_immaterial(gpu, small_system):
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    x0 = np.full(n, 1.0 / n, np.float32)
    Ra = Reconstructor(S["A"], S["sens"], "f32", 1, block_rows=0, layout="blocked")
    Ra.set_fused(False)  # the separate passes: the backward walks the row blocks
    a = Ra.mlem(S["y"], x0, 5)
    Rb = Reconstructor(S["A"], S["sens"], "f32", 1, block_rows=999, layout="blocked")
    Rb.set_fused(False)
    b = Rb.mlem(S["y"], x0, 5)
    assert _rel(a, b) <= 1e-6
    # several row blocks per backward launch (sums carried in registers)
    assert Rb.info()["row_blocks"] > 8
    for bpl in (3, 64):
        Rb.set_blocks_per_launch(bpl)
        assert _rel(Rb.mlem(S["y"], x0, 5), a) <= 1e-6


@pytest.mark.parametrize("F", [2, 8, 16, 64])
def test_recon_batched_frames_match_single(gpu, oracle, small_system, F):
    """F frames with different count totals reconstructed together equal
    each frame reconstructed alone (1e-6 relative L2) and the FP64 oracle
    (1e-4).  F = 2 uses 2-frame lane vectors, F >= 4 4-frame vectors with
    64/(F/4) entry slots per wave."""
    from raptr_amd import synthetic
    from raptr_amd.recon import Reconstructor
    S = small_system
    n, m = S["grid"].nvox, S["A"].m
    ybar = S["ybar"][S["meas"]]
    Y = np.stack([synthetic.poisson_counts(ybar, tot, seed=10 + f)
                  for f, tot in enumerate(np.geomspace(1e4, 2e6, F))], axis=1).astype(np.float32)
    Rb = Reconstructor(S["A"], S["sens"], "f32", F)
    X0 = np.full((n, F), 1.0 / n, np.float32)
    X = Rb.mlem(Y, X0, 5)
    R1 = Reconstructor(S["A"], S["sens"], "f32", 1)
    for f in range(F):
        x1 = R1.mlem(Y[:, f], X0[:, f], 5)
        assert _rel(X[:, f], x1) <= 1e-6, f
        if f in (0, F - 1):
            assert _rel(X[:, f], _oracle_mlem(oracle, S, Y[:, f].astype(np.float64), 5)) <= 1e-4
    # 64 frames: the grouped layouts are the default; the general kernels
    # with and without in-row broadcast of values and indices agree
    if F == 64:
        assert Rb.grouped()["on"] == 1
        Rb.set_grouped(False)
        for bc in (True, False):
            Rb.set_broadcast(bc)
            Xc = Rb.mlem(Y, X0, 5)
            assert _rel(Xc, X) <= 1e-6
        Rb.set_grouped(True)
    # fewer frames than the plan's maximum, in the same plan
    if F >= 4:
        Xh = Rb.mlem(Y[:, :F // 2].copy(), X0[:, :F // 2].copy(), 5)
        assert _rel(Xh, X[:, :F // 2]) <= 1e-6
    # frames with all-zero counts stay finite (scale 0), others unaffected
    Y[:, F // 2] = 0
    X = Rb.mlem(Y, X0, 3)
    assert np.all(np.isfinite(X))


@pytest.mark.parametrize("vt,tol", [("u16", 1e-4), ("f16", 4e-4), ("bf16", 5e-3)])
def test_recon_mixed_precision_error(gpu, oracle, small_system, vt, tol):
    """16-bit matrix values with per-row FP32 scale, FP32 accumulation:
    relative L2 against the FP64 oracle after 5 iterations.  Stated bounds:
    UNORM16 1e-4 (measured ~1e-5), FP16 4e-4 (measured 2.0e-4), BF16 5e-3
    (measured 2.6e-3)."""
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    x64 = _oracle_mlem(oracle, S, S["y"], 5)
    R = Reconstructor(S["A"], S["sens"], vt, 1)
    x = R.mlem(S["y"], np.full(n, 1.0 / n, np.float32), 5)
    rel = _rel(x, x64)
    mx = float(np.max(np.abs(x - x64)) / np.max(np.abs(x64)))
    print(f"{vt}: rel L2 {rel:.2e}, max voxel {mx:.2e}")
    assert rel <= tol


def test_recon_graph_replays_match_eager(gpu, small_system):
    """Graph replays (with eager work and syncs between them) equal the eager
    result bit for bit every time."""
    from raptr_amd.device import DeviceArray, Stream
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    R = Reconstructor(S["A"], S["sens"], "f32", 1)
    x0 = np.full(n, 1.0 / n, np.float32)
    eager = R.mlem(S["y"], x0, 4, use_graph=False)
    st = Stream()
    dy = DeviceArray.from_numpy(S["y"].astype(np.float32))
    dx = DeviceArray.from_numpy(x0)
    for rep in range(3):
        dx.copy_from(x0, stream=st.handle)
        R.run(dy, dx, 4, 1, use_graph=True, stream=st.handle)
        st.sync()
        np.testing.assert_array_equal(dx.numpy(), eager)
        R.mlem(S["y"], x0, 1, use_graph=False)  # eager work between replays


def test_recon_graph_cache_exact_keys_and_bound(gpu, small_system):
    """Each distinct (frames, y, x) buffer pair gets its own captured graph
    (exact keys, no aliasing); the cache holds at most 16, least recently
    used evicted; every replay, including a recaptured evicted pair, equals
    the eager result bit for bit."""
    from raptr_amd import _capi as capi
    from raptr_amd.device import DeviceArray, Stream
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    R = Reconstructor(S["A"], S["sens"], "f32", 1)
    count = lambda: capi.lib().raptr_recon_graph_count(R.handle)  # noqa: E731
    x0 = np.full(n, 1.0 / n, np.float32)
    y = S["y"].astype(np.float32)
    eager = R.mlem(y, x0, 3, use_graph=False)
    st = Stream()
    pairs = [(DeviceArray.from_numpy(y), DeviceArray.from_numpy(x0)) for _ in range(18)]
    assert count() == 0
    for k, (dy, dx) in enumerate(pairs):
        R.run(dy, dx, 3, 1, use_graph=True, stream=st.handle)
        st.sync()
        assert count() == min(k + 1, 16)
        np.testing.assert_array_equal(dx.numpy(), eager)
    for k in (17, 0):  # cached (no new capture), then evicted (recaptured)
        dy, dx = pairs[k]
        dx.copy_from(x0, stream=st.handle)
        R.run(dy, dx, 3, 1, use_graph=True, stream=st.handle)
        st.sync()
        assert count() == 16
        np.testing.assert_array_equal(dx.numpy(), eager)
    inf = R.info()  # a shape change drops every graph
    R.set_config(inf["fwd_lanes"], 1 if inf["fwd_unroll"] != 1 else 2, inf["bwd_unroll"])
    assert count() == 0


@pytest.mark.parametrize("cap", [0, 3000])
def test_recon_forward_ratio(gpu, small_system, cap):
    """The forward pass alone: e = y / (A x) against an FP64 host product."""
    from raptr_amd import _capi as capi
    from raptr_amd.device import DeviceArray
    from raptr_amd.recon import Reconstructor
    S = small_system
    A, n = S["A"], S["grid"].nvox
    rp, col, val = A.to_host()
    x = np.random.default_rng(3).random(n).astype(np.float32)
    yt = np.array([np.dot(val[rp[i]:rp[i + 1]].astype(np.float64), x[col[rp[i]:rp[i + 1]]])
                   for i in range(A.m)])
    R = Reconstructor(A, S["sens"], "f32", 1, block_rows=cap)
    dx, dy = DeviceArray.from_numpy(x), DeviceArray.from_numpy(S["y"].astype(np.float32))
    de = DeviceArray(A.m, np.float32)
    capi.call("raptr_recon_forward", R.handle, 1, dx.ptr, dy.ptr, de.ptr, None)
    np.testing.assert_allclose(de.numpy(), S["y"] / yt, rtol=1e-5)


def test_recon_rejects_bad_arguments(gpu, small_system):
    from raptr_amd import RaptrError
    from raptr_amd.recon import Reconstructor
    S = small_system
    with pytest.raises(RaptrError):
        Reconstructor(S["A"], S["sens"], "f32", 3)
    with pytest.raises(KeyError):
        Reconstructor(S["A"], S["sens"], "f32", 2, layout="tiled")
    R = Reconstructor(S["A"], S["sens"], "f32", 2)
    with pytest.raises(RaptrError):
        R.mlem(np.zeros((S["A"].m, 4), np.float32), np.zeros((S["grid"].nvox, 4), np.float32), 1)


def _locality_order(setup, ch, T, kind):
    """The locality orders as a stable lexicographic sort of independently
    computed key columns (most significant first)."""
    ida, z0, y0, z1, y1 = (np.asarray(v, np.int64) for v in setup.split(ch))
    zm, ym = z0 + z1, y0 + y1
    dz, dy = z1 - z0, y1 - y0
    if kind == "y":
        cols = [ym, ida, zm // T, zm % T, dz, dy]
    elif kind == "angle":
        cols = [ida, zm // T, ym // T, zm % T, ym % T, dz, dy]
    else:
        # slopes in 4 x 4 tiles counted from the smallest slope
        sz, sy = dz + setup.n0z - 1, dy + setup.n0y - 1
        head = [ida // 4] if kind == "angle_slope" else [ida]
        tail = [ida % 4] if kind == "angle_slope" else []
        cols = head + [zm // T, ym // T, zm % T, ym % T, sz // 4, sy // 4, sz % 4,
                       sy % 4] + tail
    return np.lexsort(cols[::-1])


@pytest.mark.parametrize("kind,tile", [("angle", 1), ("angle", 3), ("angle", 8), ("y", 8),
                                       ("slope", 8), ("angle_slope", 8), ("angle_slope", 3)])
def test_locality_order_is_a_stable_sort(gpu, small_system, kind, tile):
    """raptr_locality_order_ex: a permutation of the rows, sorted by the
    kind's keys computed independently here, ties in list order."""
    from raptr_amd import sysmat
    S = small_system
    order = sysmat.locality_order(S["setup"], S["meas"], tile=tile, kind=kind).numpy()
    assert np.array_equal(np.sort(order), np.arange(len(S["meas"])))
    np.testing.assert_array_equal(order, _locality_order(S["setup"], S["meas"], tile, kind))


@pytest.mark.parametrize("F", [1, 4])
def test_recon_row_ordered_plan(gpu, oracle, small_system, F):
    """A plan in locality row order: the forward's ratios come back in the
    caller's row order bit for bit (no row's sum changes), and MLEM matches
    the caller-ordered plan to FP32 rounding and the FP64 oracle to 1e-4."""
    from raptr_amd import _capi as capi
    from raptr_amd.device import DeviceArray
    from raptr_amd.recon import Reconstructor
    S = small_system
    A, n, m = S["A"], S["grid"].nvox, S["A"].m
    R0 = Reconstructor(A, S["sens"], "f32", F)
    R1 = Reconstructor(A, S["sens"], "f32", F, row_order=S["setup"])
    assert R1.row_ordered and not R0.row_ordered
    rng = np.random.default_rng(5)
    x = rng.random((n, F)).astype(np.float32)
    y = np.repeat(S["y"].astype(np.float32)[:, None], F, 1) * rng.random((1, F)).astype(np.float32)
    dx, dy = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y)
    es = []
    for R in (R0, R1):
        de = DeviceArray(m * F, np.float32)
        capi.call("raptr_recon_forward", R.handle, F, dx.ptr, dy.ptr, de.ptr, None)
        es.append(de.numpy())
    np.testing.assert_array_equal(es[1], es[0])
    x0 = np.full((n, F), 1.0 / n, np.float32)
    a = R0.mlem(y if F > 1 else y[:, 0], x0 if F > 1 else x0[:, 0], 5)
    b = R1.mlem(y if F > 1 else y[:, 0], x0 if F > 1 else x0[:, 0], 5)
    assert _rel(b, a) <= 1e-6
    x64 = _oracle_mlem(oracle, S, y[:, 0].astype(np.float64), 5)
    assert _rel((b if F == 1 else b[:, 0]).astype(np.float64), x64) <= 1e-4


def test_recon_row_order_must_be_a_permutation(gpu, small_system):
    from raptr_amd import RaptrError
    from raptr_amd.recon import Reconstructor
    S = small_system
    m = S["A"].m
    bad = np.arange(m, dtype=np.int32)
    bad[1] = 0
    with pytest.raises(RaptrError, match="permutation"):
        Reconstructor(S["A"], S["sens"], "f32", 1, row_order=bad)
    with pytest.raises(RaptrError, match="permutation"):
        Reconstructor(S["A"], S["sens"], "f32", 1, row_order=np.arange(1, m + 1, dtype=np.int32))


@pytest.mark.parametrize("vt,tol", [("f32", 1e-6), ("f16", 1e-5), ("bf16", 1e-5), ("u16", 1e-4)])
@pytest.mark.parametrize("group,bricks", [(16, True), (16, False), (32, True)])
def test_recon_grouped64(gpu, oracle, small_system, vt, tol, group, bricks):
    """64-frame grouped layouts (G rows / G voxels per union list; voxel
    groups as bricks of the grid or runs of column ids) against the
    general 64-frame kernels of the same plan, for every value type and
    kernel shape, with several row blocks per launch, in the caller's and
    in the locality row order; graph replays equal eager runs bit for
    bit."""
    from raptr_amd import synthetic
    from raptr_amd.device import DeviceArray, Stream
    from raptr_amd.recon import Reconstructor
    S = small_system
    n, m = S["grid"].nvox, S["A"].m
    ybar = S["ybar"][S["meas"]]
    Y = np.stack([synthetic.poisson_counts(ybar, tot, seed=40 + f)
                  for f, tot in enumerate(np.geomspace(1e4, 2e6, 64))], axis=1).astype(np.float32)
    X0 = np.full((n, 64), 1.0 / n, np.float32)
    for order in (None, S["setup"]):
        R = Reconstructor(S["A"], S["sens"], vt, 64, block_rows=1500, row_order=order,
                          grid=S["grid"] if bricks else None, group=group)
        gi = R.grouped()
        assert gi["on"] == 1 and gi["group"] == group
        # union lists are no longer than the member lists together
        assert 0 < gi["fwd_entries"] <= S["A"].nnz and 0 < gi["bwd_entries"] <= S["A"].nnz
        R.set_grouped(False)
        ref = R.mlem(Y, X0, 5)
        for gu, gp in Reconstructor.GROUP_SHAPES[group]:
            for bpl in (1, 3):
                R.set_grouped(True, gu, gp)
                R.set_blocks_per_launch(bpl)
                # kind 3: also its backward split over launches of 1 and 2
                # row blocks per XCD (partial images added between launches)
                R.set_xcd_blocks(bpl - 1 if gp == 3 else 0)
                X = R.mlem(Y, X0, 5)
                rel = _rel(X, ref)
                print(f"{vt} G={group} bricks={bricks} order={order is not None} unroll {gu} "
                      f"kind {gp} bpl {bpl} xcd-blocks {bpl - 1 if gp == 3 else 0}: "
                      f"rel {rel:.2e} (union fwd {gi['fwd_entries']} "
                      f"bwd {gi['bwd_entries']} of {S['A'].nnz})")
                assert rel <= tol
        if vt == "f32":
            x64 = _oracle_mlem(oracle, S, Y[:, 7].astype(np.float64), 5)
            assert _rel(X[:, 7].astype(np.float64), x64) <= 1e-4
            # graph replays of every shape equal its eager run bit for bit
            st = Stream()
            dy, dx = DeviceArray.from_numpy(Y.reshape(-1)), DeviceArray.from_numpy(X0.reshape(-1))
            for gu, gp in Reconstructor.GROUP_SHAPES[group]:
                R.set_grouped(True, gu, gp)
                R.set_blocks_per_launch(2)
                dx.copy_from(X0.reshape(-1), stream=st.handle)
                R.run(dy, dx, 3, 64, use_graph=False, stream=st.handle)
                eager = dx.numpy(stream=st.handle)
                for _ in range(2):
                    dx.copy_from(X0.reshape(-1), stream=st.handle)
                    R.run(dy, dx, 3, 64, use_graph=True, stream=st.handle)
                    st.sync()
                    np.testing.assert_array_equal(dx.numpy(stream=st.handle), eager)
                    R.mlem(Y[:, :1].copy(), X0[:, :1].copy(), 1, use_graph=False)
        R.close()


def test_recon_grouped64_rejects_bad_shapes(gpu, small_system):
    from raptr_amd import RaptrError
    from raptr_amd.recon import Reconstructor
    S = small_system
    for g in (6, 4, 8):  # groups of 4 and 8 ran the removed scalar-register kernels
        with pytest.raises(RaptrError, match="group"):
            Reconstructor(S["A"], S["sens"], "f32", 64, group=g)
    R = Reconstructor(S["A"], S["sens"], "f32", 64, grid=S["grid"], group=32)
    with pytest.raises(RaptrError, match="unroll"):
        R.set_grouped(True, 8, 2)  # kind 2 needs groups of 16
    R16 = Reconstructor(S["A"], S["sens"], "f32", 64, grid=S["grid"], group=16)
    for gu, kind in ((2, 1), (4, 0), (4, 3), (8, 4)):
        with pytest.raises(RaptrError, match="unroll"):
            R16.set_grouped(True, gu, kind)
    R1 = Reconstructor(S["A"], S["sens"], "f32", 16)
    assert R1.grouped()["group"] == 0
    with pytest.raises(RaptrError, match="no grouped"):
        R1.set_grouped(True, 8, 3)


@pytest.mark.parametrize("vt,active", [("f32", 16), ("f32", 32), ("u16", 16)])
def test_recon_mlem_active_frames(gpu, small_system, vt, active):
    """raptr_recon_mlem_active: a 64-frame grouped pass over the first 16 or
    32 frames gives those frames the images of the full 64-frame pass bit
    for bit (eager and graph replays) and leaves frames active .. 63 as they
    were; kind 2 plans, other counts and non-grouped plans are refused."""
    from raptr_amd import RaptrError, synthetic
    from raptr_amd.device import DeviceArray, Stream
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    ybar = S["ybar"][S["meas"]]
    Y = np.stack([synthetic.poisson_counts(ybar, tot, seed=90 + f)
                  for f, tot in enumerate(np.geomspace(1e4, 2e6, 64))], axis=1).astype(np.float32)
    rng = np.random.default_rng(1)
    X0 = (rng.random((n, 64)) + 0.5).astype(np.float32) / n
    R = Reconstructor(S["A"], S["sens"], vt, 64, block_rows=1500, row_order=S["setup"],
                      grid=S["grid"], group=16, grouped_only=True)
    assert R.grouped()["kind"] == 3
    st = Stream()
    dy = DeviceArray.from_numpy(Y.reshape(-1))
    dx = DeviceArray.from_numpy(X0.reshape(-1))
    R.run(dy, dx, 4, 64, use_graph=False, stream=st.handle)
    full = dx.numpy(stream=st.handle).reshape(n, 64)
    for graph in (False, True, True):
        dx.copy_from(X0.reshape(-1), stream=st.handle)
        R.run(dy, dx, 4, 64, use_graph=graph, stream=st.handle, active=active)
        got = dx.numpy(stream=st.handle).reshape(n, 64)
        np.testing.assert_array_equal(got[:, :active], full[:, :active])
        np.testing.assert_array_equal(got[:, active:], X0[:, active:])
    with pytest.raises(RaptrError):
        R.run(dy, dx, 1, 64, use_graph=False, active=8)
    R.set_grouped(True, 8, 2)
    with pytest.raises(RaptrError, match="kind 3"):
        R.run(dy, dx, 1, 64, use_graph=False, active=active)
    R2 = Reconstructor(S["A"], S["sens"], vt, 16)
    with pytest.raises(RaptrError):
        R2.run(DeviceArray.from_numpy(Y[:, :16].copy().reshape(-1)),
               DeviceArray.from_numpy(X0[:, :16].copy().reshape(-1)), 1, 16, use_graph=False,
               active=8)


@pytest.mark.parametrize("vt", ["f32", "f16"])
def test_recon_grouped_only_plan(gpu, small_system, vt):
    """A grouped-only plan (no padded CSR / transpose) gives the images of
    the same grouped kernels in a full plan bit for bit, uses less device
    memory, and refuses passes that need the general layouts."""
    from raptr_amd import RaptrError, synthetic
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    ybar = S["ybar"][S["meas"]]
    Y = np.stack([synthetic.poisson_counts(ybar, tot, seed=70 + f)
                  for f, tot in enumerate(np.geomspace(1e4, 2e6, 64))], axis=1).astype(np.float32)
    X0 = np.full((n, 64), 1.0 / n, np.float32)
    kw = dict(block_rows=1500, row_order=S["setup"], grid=S["grid"], group=16)
    full = Reconstructor(S["A"], S["sens"], vt, 64, **kw)
    only = Reconstructor(S["A"], S["sens"], vt, 64, grouped_only=True, **kw)
    # the general layouts it drops: value + 4-byte index per padded CSR and
    # padded transpose entry (the grouped lists and per-XCD images remain)
    fi = full.info()
    bv = 4 if vt == "f32" else 2
    assert (fi["device_bytes"] - only.info()["device_bytes"] >=
            (bv + 4) * (fi["padded_nnz"] + fi["transpose_nnz"]))
    assert only.info()["padded_nnz"] == 0 and only.info()["transpose_nnz"] == 0
    for gu, kind in ((16, 3), (8, 2), (32, 3)):
        full.set_grouped(True, gu, kind)
        only.set_grouped(True, gu, kind)
        np.testing.assert_array_equal(only.mlem(Y, X0, 3), full.mlem(Y, X0, 3))
    with pytest.raises(RaptrError, match="grouped-only"):
        only.mlem(Y[:, :16].copy(), X0[:, :16].copy(), 1)
    with pytest.raises(RaptrError, match="grouped-only"):
        only.set_grouped(False)
    with pytest.raises(RaptrError, match="grouped-only"):
        Reconstructor(S["A"], S["sens"], vt, 16, grouped_only=True)


def _runs(rng, n, L):
    """L ascending columns in runs of 1..9 consecutive voxels, the last run
    ending at voxel n - 1."""
    out, c = [], n
    while len(out) < L:
        r = int(rng.integers(1, 10))
        c -= r + int(rng.integers(1, 60))
        out.extend(range(max(c, 0), max(c, 0) + r))
        if c <= 0:
            break
    c = np.unique(np.array(out[:L], np.int64))
    if len(c):
        c = c + (n - 1 - c.max())
    return c


def _wide_csr(n, m=2500, seed=5, unsorted_row=False):
    """Host CSR over n voxels whose rows span several 65,536-voxel blocks:
    lengths 0..700 plus a few rows of 3,000-9,000 entries (several 128-entry
    interleave chunks and a ragged last chunk), ascending columns; every
    third row made of runs of 1..9 consecutive voxels ending at voxel n - 1."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 700, m)
    lens[rng.choice(m, 6, replace=False)] = rng.integers(3000, 9000, 6)
    lens[:3] = 0
    cols = []
    for r, L in enumerate(lens):
        if r % 3 == 1:
            c = _runs(rng, n, L)
            lens[r] = len(c)
            cols.append(c)
            continue
        c0 = int(rng.integers(0, n))
        span = int(rng.integers(L + 1, min(n, 40 * L + 2) + 1))
        c = np.sort(rng.choice(span, L, replace=False)) + c0
        cols.append(np.sort(c % n))
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate(cols).astype(np.int32)
    if unsorted_row:  # one long row with descending columns over the whole grid
        i = int(np.argmax(lens))
        a, b = rp[i], rp[i + 1]
        col[a:b] = np.sort(rng.choice(n, b - a, replace=False))[::-1]
    val = rng.random(len(col)).astype(np.float32) + 0.01
    return rp, col, val


def _local_csr(n, m=1500, seed=9, window=8000):
    """Host CSR whose groups of 256 consecutive rows draw their columns from
    one window of `window` voxels placed anywhere in the grid (the column
    unions of the LDS-staged forward fit): lengths 0..700, ascending."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 700, m)
    lens[:2] = 0
    cols = []
    for r, L in enumerate(lens):
        if r % 256 == 0:
            w0 = int(rng.integers(0, n - window + 1)) if r < m - 256 else n - window
        cols.append(np.sort(rng.choice(window, L, replace=False)) + w0)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate(cols).astype(np.int32)
    val = rng.random(len(col)).astype(np.float32) + 0.01
    return rp, col, val


def _device_csr(rp, col, val, n):
    from raptr_amd.device import DeviceArray
    from raptr_amd.sysmat import DeviceCSR
    return DeviceCSR(len(rp) - 1, n, len(col), DeviceArray.from_numpy(rp),
                     DeviceArray.from_numpy(col), DeviceArray.from_numpy(val))


@pytest.mark.parametrize("vt", ["f32", "f16", "bf16", "u16"])
def test_recon_forward_compact_columns(gpu, vt):
    """Single-frame forward on a 262,139-voxel matrix whose rows cross the
    65,536-voxel blocks (and end at the last voxel): 16-bit columns (low
    parts + per-row split points), plain and software-pipelined, give the
    same ratios, bit for bit, as 32-bit columns for every lanes/unroll
    shape; in FP32 they match an FP64 host product."""
    from raptr_amd import _capi as capi
    from raptr_amd.device import DeviceArray
    from raptr_amd.recon import Reconstructor
    n = 4 * 65536 - 5
    rp, col, val = _wide_csr(n)
    m = len(rp) - 1
    assert (col >> 16).max() == 3 and col.max() == n - 1
    A = _device_csr(rp, col, val, n)
    sens = DeviceArray.from_numpy(np.ones(n))
    rng = np.random.default_rng(7)
    x = rng.random(n).astype(np.float32)
    y = rng.random(m).astype(np.float32) + 0.5
    R = Reconstructor(A, sens, vt, 1, layout="blocked")
    assert R.info()["compact_rows"] & 2
    R.set_forward_kind(0)  # rows of global columns, not the LDS-staged unions
    assert R.forward_kind()[0] == 0
    dx, dy, de = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y), DeviceArray(m, np.float32)
    yt = np.array([np.dot(val[rp[i]:rp[i + 1]].astype(np.float64), x[col[rp[i]:rp[i + 1]]])
                   for i in range(m)])
    ref = np.where(yt != 0, y / np.where(yt != 0, yt, 1), 0)

    def fwd():
        capi.call("raptr_recon_forward", R.handle, 1, dx.ptr, dy.ptr, de.ptr, None)
        return de.numpy()

    for lanes, unroll in [(16, 1), (32, 2), (64, 4), (16, 4), (64, 1), (32, 4)]:
        R.set_config(lanes, unroll, 1)
        out = []
        for on in (True, False):
            R.set_compact_rows(on)
            assert bool(R.info()["compact_rows"] & 2) == on
            out.append(fwd())
        np.testing.assert_array_equal(out[0], out[1])
        if vt == "f32":
            np.testing.assert_allclose(out[0], ref, rtol=1e-5)


@pytest.mark.parametrize("case", ["descending", "too_many_voxels"])
def test_recon_forward_compact_columns_absent(gpu, case):
    """No 16-bit columns when a row's columns step down across 65,536-voxel
    blocks or the grid has more than 262,144 voxels: the plan keeps its
    32-bit columns and the forward still matches the host product."""
    from raptr_amd import RaptrError
    from raptr_amd import _capi as capi
    from raptr_amd.device import DeviceArray
    from raptr_amd.recon import Reconstructor
    n = 4 * 65536 - 5 if case == "descending" else 4 * 65536 + 1
    rp, col, val = _wide_csr(n, m=800, unsorted_row=case == "descending")
    m = len(rp) - 1
    A = _device_csr(rp, col, val, n)
    R = Reconstructor(A, DeviceArray.from_numpy(np.ones(n)), "f32", 1, layout="blocked",
                      block_rows=512)
    assert R.info()["compact_rows"] & 2 == 0
    R.set_union_rows(64)  # groups of 64 wide rows touch more than 16,384 voxels
    assert R.forward_kind() == (0, [0])
    with pytest.raises(RaptrError):
        R.set_forward_kind(2)
    x = np.random.default_rng(2).random(n).astype(np.float32)
    y = np.ones(m, np.float32)
    dx, dy, de = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y), DeviceArray(m, np.float32)
    capi.call("raptr_recon_forward", R.handle, 1, dx.ptr, dy.ptr, de.ptr, None)
    yt = np.array([np.dot(val[rp[i]:rp[i + 1]].astype(np.float64), x[col[rp[i]:rp[i + 1]]])
                   for i in range(m)])
    ref = np.where(yt != 0, y / np.where(yt != 0, yt, 1), 0)
    np.testing.assert_allclose(de.numpy(), ref, rtol=1e-5)


@pytest.mark.parametrize("vt", ["f32", "bf16", "u16"])
def test_recon_forward_lds_staged(gpu, vt):
    """LDS-staged forward (kind 2: x over each 64-row group's column union
    in LDS, 16-bit ranks) on a 258,000-voxel matrix whose groups sit in
    8,000-voxel windows across the grid: the same ratios, bit for bit, as
    kind 0 over 16-bit and 32-bit columns for every lanes/unroll shape; FP32
    against an FP64 host product."""
    from raptr_amd import _capi as capi
    from raptr_amd.device import DeviceArray
    from raptr_amd.recon import Reconstructor
    n = 258000
    rp, col, val = _local_csr(n)
    m = len(rp) - 1
    assert col.max() == n - 1 and (col >> 16).max() == 3
    A = _device_csr(rp, col, val, n)
    R = Reconstructor(A, DeviceArray.from_numpy(np.ones(n)), vt, 1, layout="blocked")
    assert R.forward_kind() == (2, [0, 2])
    rng = np.random.default_rng(4)
    x = rng.random(n).astype(np.float32)
    y = rng.random(m).astype(np.float32) + 0.5
    dx, dy, de = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y), DeviceArray(m, np.float32)

    def fwd():
        capi.call("raptr_recon_forward", R.handle, 1, dx.ptr, dy.ptr, de.ptr, None)
        return de.numpy()

    for lanes, unroll in [(16, 1), (32, 2), (64, 4), (16, 4), (64, 1)]:
        R.set_config(lanes, unroll, 1)
        out = []
        for on, k in ((True, 2), (True, 0), (False, 0)):
            R.set_compact_rows(on)
            R.set_forward_kind(k)
            out.append(fwd())
        for o in out[1:]:
            np.testing.assert_array_equal(out[0], o)
    if vt == "f32":
        yt = np.array([np.dot(val[rp[i]:rp[i + 1]].astype(np.float64), x[col[rp[i]:rp[i + 1]]])
                       for i in range(m)])
        np.testing.assert_allclose(out[0], np.where(yt != 0, y / np.where(yt != 0, yt, 1), 0),
                                   rtol=1e-5)


def _straddling_csr(n, m=1200, seed=13, window=6000, win=256000):
    """Host CSR whose groups of 64 consecutive rows draw their columns from
    one window of `window` voxels; every other group's window straddles a
    multiple of `win` (the union builder's voxel windows), the last group's
    ends at voxel n - 1.  Lengths 0..500, ascending."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 500, m)
    cols = []
    nb = (n - 1) // win
    for r, L in enumerate(lens):
        if r % 64 == 0:
            g = r // 64
            if r >= m - 64:
                w0 = n - window
            elif g % 2 == 0 and nb > 0:
                w0 = win * int(rng.integers(1, nb + 1)) - window // 2
            else:
                w0 = int(rng.integers(0, n - window + 1))
        c = rng.choice(window, L, replace=False)
        if r == m - 1:  # the last row ends at voxel n - 1
            c = np.union1d(c, [window - 1])
            lens[r] = len(c)
        cols.append(np.sort(c) + w0)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate(cols).astype(np.int32)
    val = rng.random(len(col)).astype(np.float32) + 0.01
    return rp, col, val


@pytest.mark.parametrize("n", [600001, 884736])
def test_recon_forward_lds_staged_large_grid(gpu, n):
    """Grids of more than one 256,000-voxel window of the union builder
    (600,001 voxels; 96^3): the LDS-staged forward is built, groups whose
    columns straddle a window edge included, and gives the ratios of the
    32-bit per-entry forward bit for bit and of an FP64 host product."""
    from raptr_amd import _capi as capi
    from raptr_amd.device import DeviceArray
    from raptr_amd.recon import Reconstructor
    rp, col, val = _straddling_csr(n)
    m = len(rp) - 1
    assert col.max() == n - 1
    R = Reconstructor(_device_csr(rp, col, val, n), DeviceArray.from_numpy(np.ones(n)), "f32", 1)
    assert R.forward_kind() == (2, [0, 2])
    assert R.info()["compact_rows"] & 2 == 0  # no 16-bit columns beyond 262,144 voxels
    rng = np.random.default_rng(6)
    x = rng.random(n).astype(np.float32)
    y = rng.random(m).astype(np.float32) + 0.5
    dx, dy, de = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y), DeviceArray(m, np.float32)
    for lanes, unroll in [(16, 4), (32, 2), (64, 1)]:
        R.set_config(lanes, unroll, 1)
        out = []
        for k in (2, 0):
            R.set_forward_kind(k)
            capi.call("raptr_recon_forward", R.handle, 1, dx.ptr, dy.ptr, de.ptr, None)
            out.append(de.numpy())
        np.testing.assert_array_equal(out[0], out[1])
    yt = np.array([np.dot(val[rp[i]:rp[i + 1]].astype(np.float64), x[col[rp[i]:rp[i + 1]]])
                   for i in range(m)])
    np.testing.assert_allclose(out[0], np.where(yt != 0, y / np.where(yt != 0, yt, 1), 0),
                               rtol=1e-5)


def test_recon_forward_lds_staged_absent(gpu):
    """No column unions when a 64-row group touches more than 16,384
    voxels (rows of random columns over the grid, one window of the union
    builder or several; groups of 64 rows set explicitly, since the plan's
    own choice for 600 rows is the smallest group); kind 2 is then
    refused."""
    from raptr_amd import RaptrError
    from raptr_amd.device import DeviceArray
    from raptr_amd.recon import Reconstructor
    for n, (rp, col, val) in ((200000, _wide_csr(200000, m=600)),
                              (700000, _wide_csr(700000, m=600))):
        R = Reconstructor(_device_csr(rp, col, val, n), DeviceArray.from_numpy(np.ones(n)), "f32",
                          1, layout="blocked")
        R.set_union_rows(64)
        assert R.unions()["rows"] == 0 and R.forward_kind() == (0, [0])
        with pytest.raises(RaptrError):
            R.set_forward_kind(2)


def test_recon_forward_lds_staged_small_system(gpu, oracle, small_system):
    """The small workload in locality row order builds column unions (kind 2
    is the default); MLEM with kinds 2 and 0 is identical bit for bit and
    within 1e-4 of the FP64 oracle."""
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    x0 = np.full(n, 1.0 / n, np.float32)
    R = Reconstructor(S["A"], S["sens"], "f32", 1, layout="blocked", row_order=S["setup"])
    assert R.forward_kind() == (2, [0, 2])
    R.set_fused(False)  # the separate passes (the fused iteration has its own forward)
    a = R.mlem(S["y"], x0, 5)
    R.set_forward_kind(0)
    b = R.mlem(S["y"], x0, 5)
    np.testing.assert_array_equal(a, b)
    assert _rel(a, _oracle_mlem(oracle, S, S["y"], 5)) <= 1e-4


@pytest.mark.parametrize("case", ["empty_group", "one_row"])
def test_recon_forward_lds_staged_edges(gpu, case):
    """LDS-staged forward edge cases: 256 rows (4 groups) with no entries
    and a ragged last group (m = 785), and a one-row plan: the same ratios,
    bit for bit, as the per-entry forward, and the FP64 host product."""
    from raptr_amd import _capi as capi
    from raptr_amd.device import DeviceArray
    from raptr_amd.recon import Reconstructor
    n = 5000
    rng = np.random.default_rng(11)
    m = 785 if case == "empty_group" else 1
    lens = rng.integers(1, 300, m)
    if case == "empty_group":
        lens[256:512] = 0
    cols = [np.sort(rng.choice(n, L, replace=False)) for L in lens]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate(cols).astype(np.int32)
    val = rng.random(len(col)).astype(np.float32) + 0.01
    R = Reconstructor(_device_csr(rp, col, val, n), DeviceArray.from_numpy(np.ones(n)), "f32", 1,
                      layout="blocked")
    assert R.forward_kind()[0] == 2
    x = rng.random(n).astype(np.float32)
    y = rng.random(m).astype(np.float32) + 0.5
    dx, dy, de = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y), DeviceArray(m, np.float32)
    out = []
    for k in (2, 0):
        R.set_forward_kind(k)
        capi.call("raptr_recon_forward", R.handle, 1, dx.ptr, dy.ptr, de.ptr, None)
        out.append(de.numpy())
    np.testing.assert_array_equal(out[0], out[1])
    yt = np.array([np.dot(val[rp[i]:rp[i + 1]].astype(np.float64), x[col[rp[i]:rp[i + 1]]])
                   for i in range(m)])
    np.testing.assert_allclose(out[0], np.where(yt != 0, y / np.where(yt != 0, yt, 1), 0),
                               rtol=1e-5)
    if case == "empty_group":
        assert np.all(out[0][256:512] == 0)
