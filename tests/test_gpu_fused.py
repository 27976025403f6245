"""GPU tests of the fused single-frame iteration (raptr_recon_set_fused):
one pass over the matrix per iteration (forward, ratio and backward per
group of 64 plan rows with per-wave column sums in LDS, then a fixed-order
reduce of the groups' partial sums).  Checked against the FP64 oracle
recurrence, against the separate forward and backward passes, and against
itself run to run (bit for bit); edge cases: empty row groups, a ragged last
group, a one-row plan, rows longer than the two register-held steps, zero
counts, unions too large for LDS."""
import numpy as np
import pytest

from test_gpu_recon import _device_csr, _oracle_mlem, _rel, small_system  # noqa: F401

pytestmark = pytest.mark.gpu

SHAPES = [(o, w, r, l) for o in (1, 2) for w in ((2, 4) if o == 1 else (2, 4, 8))
          for r in (16, 32, 64) for l in ((16, 32, 64) if r == 32 else (64,))]


def test_fused_small_system(gpu, oracle, small_system):
    """Small workload in locality row order: the fused pass is built; every
    shape (a column-sum region per wave or one shared in turn; waves, lanes
    per row, reduce lanes per column) is within 1e-4 of the FP64 oracle
    after 5 iterations, within 1e-6 of the separate passes, and repeats bit
    for bit; graph replay equals eager."""
    from raptr_amd.device import Stream
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    x0 = np.full(n, 1.0 / n, np.float32)
    x64 = _oracle_mlem(oracle, S, S["y"], 5)
    R = Reconstructor(S["A"], S["sens"], "f32", 1, layout="blocked", row_order=S["setup"])
    # the default shape: one shared region, 4 waves where five such
    # workgroups fit a CU's LDS (else 8), 64 lanes per row with 2
    # register-held steps (rows of about 140 entries), the column-tile
    # reduce with 64 lanes per segment slot where its table gives 512
    # workgroups, else the per-XCD reduce with 16 lanes per column
    u = R.unions()
    lds = 8 * (u["rows"] + 1) + 4 * (2 * u["largest"] + 1)
    tiled = R.fused_tiles()["workgroups"] >= 512
    assert R.fused() == {"available": 1, "on": 2, "waves": 4 if 5 * lds <= 160 * 1024 else 8,
                         "lanes_per_row": 64, "lanes_per_column": 64 if tiled else 16,
                         "per_xcd": 2 if tiled else 1, "steps": 2}
    R.set_fused(False)
    sep = R.mlem(S["y"], x0, 5)
    assert _rel(sep, x64) <= 1e-4
    st = Stream()
    for o, w, lr, rw in SHAPES:
        R.set_fused(o, w, lr, rw)
        for xr in (2, 1, 0):
            R.set_fused_reduce(xr)
            assert R.fused() == {"available": 1, "on": o, "waves": w, "lanes_per_row": lr,
                                 "lanes_per_column": rw, "per_xcd": xr, "steps": 2}
            a = R.mlem(S["y"], x0, 5)
            b = R.mlem(S["y"], x0, 5)
            np.testing.assert_array_equal(a, b)
            assert _rel(a, x64) <= 1e-4, (o, w, lr, rw, xr)
            assert _rel(a, sep) <= 1e-6, (o, w, lr, rw, xr)
            assert abs(float(a.astype(np.float64).sum()) - 1.0) < 1e-5
            g = R.mlem(S["y"], x0, 5, use_graph=True, stream=st.handle)
            np.testing.assert_array_equal(a, g)
    for stp in (3, 4, 2):  # register-held steps of the 64-lane kernel
        R.set_fused(2, 4, 64, 16)
        R.set_fused_steps(stp)
        a = R.mlem(S["y"], x0, 5)
        np.testing.assert_array_equal(a, R.mlem(S["y"], x0, 5))
        assert _rel(a, sep) <= 1e-6, stp
    R.set_fused(False)
    np.testing.assert_array_equal(R.mlem(S["y"], x0, 5), sep)


def test_fused_tile_reduce(gpu, oracle, small_system):
    """The column-tile reduce (mode 2): for tiles of 128-2,048 columns and
    16, 32 or 64 lanes per segment, 5 iterations within 1e-4 of the FP64
    oracle and 1e-6 of the separate passes, bit for bit run to run and in a
    graph replay.  The segment count lies between the number of row groups
    and the number of union entries and falls as the tiles grow; bad tile
    sizes and tables shared with clones are refused."""
    from raptr_amd import _capi as capi
    from raptr_amd.device import Stream
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    x0 = np.full(n, 1.0 / n, np.float32)
    x64 = _oracle_mlem(oracle, S, S["y"], 5)
    R = Reconstructor(S["A"], S["sens"], "f32", 1, layout="blocked", row_order=S["setup"])
    tz = R.fused_tiles()
    assert tz["columns"] == 1024 and tz["segments_per_workgroup"] == 512 and tz["segments"] > 0
    R.set_fused(False)
    sep = R.mlem(S["y"], x0, 5)
    nu = R.forward_union_entries()
    ngroups = (R.m + 63) // 64
    st = Stream()
    last = None
    for tiles, spw in ((128, 64), (128, 1024), (256, 100), (512, 4096), (1024, 64),
                       (2048, 1 << 20)):
        R.set_fused_tiles(tiles, spw)
        tz = R.fused_tiles()
        nseg = tz["segments"]
        assert tz["columns"] == tiles and tz["segments_per_workgroup"] == spw
        assert ngroups <= nseg <= nu
        # every tile in ceil(segments / spw) parts, at least one
        ntiles = -(-n // tiles)
        assert max(ntiles, -(-nseg // spw)) <= tz["workgroups"] <= ntiles + nseg // spw + ntiles
        assert 1 <= tz["max_parts"] <= max(1, -(-nseg // spw))
        if last is not None and last[0] == tiles - tiles // 2:
            assert nseg <= last[1]
        last = (tiles, nseg)
        for lpc in (16, 32, 64):
            R.set_fused(2, 4, 64, lpc)
            R.set_fused_reduce(2)
            assert R.fused()["per_xcd"] == 2
            a = R.mlem(S["y"], x0, 5)
            np.testing.assert_array_equal(a, R.mlem(S["y"], x0, 5))
            assert _rel(a, x64) <= 1e-4, (tiles, lpc)
            assert _rel(a, sep) <= 1e-6, (tiles, lpc)
            g = R.mlem(S["y"], x0, 5, use_graph=True, stream=st.handle)
            np.testing.assert_array_equal(a, g)
    for bad, spw in ((0, 1024), (64, 1024), (100, 1024), (4096, 1024), (512, 63),
                     (512, (1 << 20) + 1)):
        with pytest.raises(capi.RaptrError):
            R.set_fused_tiles(bad, spw)
    assert R.fused_tiles()["columns"] == 2048  # unchanged by the refusals
    C = R.clone()
    with pytest.raises(capi.RaptrError):
        R.set_fused_tiles(512)
    with pytest.raises(capi.RaptrError):
        C.set_fused_tiles(512)
    np.testing.assert_array_equal(C.mlem(S["y"], x0, 5), R.mlem(S["y"], x0, 5))
    C.close()
    R.set_fused_tiles(1024, 512)


def test_fused_union_rows(gpu, oracle, small_system):
    """Plan rows per union group: the plan's choice is one of 16-256 rows;
    for groups of 16-512 rows the union entries do not grow with the group
    size and the fused iteration stays within 1e-4 of the FP64 oracle and
    1e-6 of the separate passes, bit for bit run to run; bad sizes and
    unions shared with clones are refused."""
    from raptr_amd import _capi as capi
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    x0 = np.full(n, 1.0 / n, np.float32)
    x64 = _oracle_mlem(oracle, S, S["y"], 5)
    R = Reconstructor(S["A"], S["sens"], "f32", 1, layout="blocked", row_order=S["setup"])
    u0 = R.unions()
    assert u0["rows"] in (16, 32, 64, 128, 256) and 0 < u0["largest"] <= u0["entries"]
    R.set_fused(False)
    sep = R.mlem(S["y"], x0, 5)
    last = None
    for rows in (16, 32, 64, 128, 256, 512):
        R.set_union_rows(rows)
        u = R.unions()
        assert u["rows"] == rows and R.fused()["on"] == 2
        if last is not None:
            assert u["entries"] <= last
        last = u["entries"]
        a = R.mlem(S["y"], x0, 5)
        np.testing.assert_array_equal(a, R.mlem(S["y"], x0, 5))
        assert _rel(a, x64) <= 1e-4, rows
        assert _rel(a, sep) <= 1e-6, rows
        # groups of several batches of rows: every shape, in particular
        # waves holding 2 or 4 rows whose column sums go in turn
        for shape in ((2, 8, 32, 64), (2, 8, 16, 64), (2, 4, 32, 64), (2, 2, 16, 64),
                      (1, 2, 32, 64), (1, 4, 16, 64), (2, 8, 64, 64)):
            try:
                R.set_fused(*shape)
            except capi.RaptrError:  # a region per wave beyond LDS
                continue
            b = R.mlem(S["y"], x0, 5)
            assert _rel(b, sep) <= 1e-6, (rows, shape)
            np.testing.assert_array_equal(b, R.mlem(S["y"], x0, 5))
        R.set_fused(2, 4, 64, 64)
        R.set_fused(False)  # the LDS-staged forward over the same unions
        assert R.forward_kind()[0] == 2
        np.testing.assert_array_equal(R.mlem(S["y"], x0, 5), sep)
    for bad in (8, 24, 1024, -64):
        with pytest.raises(capi.RaptrError):
            R.set_union_rows(bad)
    assert R.unions()["rows"] == 512
    R.set_union_rows(0)
    assert R.unions() == u0
    C = R.clone()
    with pytest.raises(capi.RaptrError):
        R.set_union_rows(64)
    C.close()


@pytest.mark.parametrize("vt", ["f16", "bf16", "u16"])
def test_fused_16bit_values(gpu, small_system, vt):
    """16-bit values with per-row scales: the fused pass computes the same
    recurrence as the separate passes on the same stored values (within
    1e-5 after 5 iterations) and repeats bit for bit."""
    from raptr_amd.recon import Reconstructor
    S = small_system
    n = S["grid"].nvox
    x0 = np.full(n, 1.0 / n, np.float32)
    R = Reconstructor(S["A"], S["sens"], vt, 1, layout="blocked", row_order=S["setup"])
    R.set_fused(False)
    sep = R.mlem(S["y"], x0, 5)
    for o in (1, 2):
        R.set_fused(o, 4, 32, 64)
        a = R.mlem(S["y"], x0, 5)
        np.testing.assert_array_equal(a, R.mlem(S["y"], x0, 5))
        assert _rel(a, sep) <= 1e-5


def _long_row_csr(n, m, seed, long_rows=True, empty=None, window=5000, lmax=500):
    """Host CSR whose groups of 64 consecutive rows draw their columns from
    one window of `window` voxels; lengths 1..lmax, with rows of 513..2,000
    entries (past the register-held steps of every shape) where long_rows;
    rows in `empty` have no entries."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(1, lmax, m)
    if long_rows:
        k = max(1, m // 10)
        lens[rng.choice(m, k, replace=False)] = rng.integers(513, 2000, k)
        lens[m - 1] = 1999  # the last row of the ragged last group
    if empty is not None:
        lens[empty] = 0
    cols = []
    for r, L in enumerate(lens):
        if r % 64 == 0:
            w0 = int(rng.integers(0, n - window + 1))
        cols.append(np.sort(rng.choice(window, L, replace=False)) + w0)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = np.concatenate(cols).astype(np.int32)
    val = rng.random(len(col)).astype(np.float32) + 0.01
    return rp, col, val


@pytest.mark.parametrize("case", ["long_rows", "empty_group", "one_row", "long_mean"])
def test_fused_edges(gpu, oracle, case):
    """Rows longer than the register-held steps (up to 2,000 entries) and a
    ragged last group (m = 785); four empty row groups; a one-row plan;
    zero counts on some rows.  Every shape: 3 iterations within 1e-5 of the
    FP64 oracle recurrence on the same FP32 matrix, bit for bit run to run,
    and within 1e-6 of the separate passes."""
    from raptr_amd.device import DeviceArray
    from raptr_amd.recon import Reconstructor
    n = 20000
    if case == "long_rows":
        rp, col, val = _long_row_csr(n, 785, 21)
    elif case == "empty_group":
        rp, col, val = _long_row_csr(n, 785, 22, empty=slice(256, 512))
    elif case == "one_row":
        rp, col, val = _long_row_csr(n, 1, 23)
    else:  # rows of 1..1,300 entries: a mean of about 650 + padding -> 3 steps
        rp, col, val = _long_row_csr(n, 785, 26, long_rows=False, lmax=1300)
    m = len(rp) - 1
    rng = np.random.default_rng(7)
    y = (rng.random(m) * 50).astype(np.float32)
    y[rng.random(m) < 0.1] = 0.0
    if m == 1:
        y[0] = 5.0
    R = Reconstructor(_device_csr(rp, col, val, n), DeviceArray.from_numpy(np.ones(n)), "f32", 1,
                      layout="blocked")
    assert R.fused()["available"] == 1
    x0 = np.full(n, 1.0 / n, np.float32)
    x64 = oracle.mlem_f64(rp, col, val.astype(np.float64), y, np.ones(n), x0, 3)
    assert R.fused()["on"] == 2  # the default
    if case == "long_mean":
        assert R.fused()["steps"] == 3
    np.testing.assert_array_equal(R.mlem(y, x0, 3), R.mlem(y, x0, 3))
    R.set_fused(False)
    sep = R.mlem(y, x0, 3)
    assert _rel(sep, x64) <= 1e-5
    if case == "long_rows":  # tiles of 128 columns in parts of 64 segments
        R.set_fused_tiles(128, 64)
    elif case == "empty_group":  # whole tiles of 2,048 columns
        R.set_fused_tiles(2048, 1 << 20)
    for o, w, lr, rw, xr, stp in [(1, 2, 16, 16, 1, 2), (1, 4, 16, 64, 0, 2),
                                  (1, 2, 32, 32, 1, 2), (1, 4, 64, 64, 1, 3), (2, 4, 32, 16, 0, 2),
                                  (2, 2, 16, 64, 1, 2), (2, 4, 64, 32, 1, 2), (2, 8, 64, 64, 0, 4),
                                  (2, 4, 64, 16, 1, 3), (2, 8, 64, 16, 1, 2), (2, 4, 64, 16, 2, 2),
                                  (2, 8, 64, 64, 2, 4), (1, 2, 32, 32, 2, 2)]:
        R.set_fused(o, w, lr, rw)
        R.set_fused_reduce(xr)
        R.set_fused_steps(stp)
        a = R.mlem(y, x0, 3)
        np.testing.assert_array_equal(a, R.mlem(y, x0, 3))
        assert _rel(a, x64) <= 1e-5, (o, w, lr, rw)
        assert _rel(a, sep) <= 1e-6, (o, w, lr, rw)
        if case == "one_row":  # voxels of no row keep x * 0 = 0
            assert np.all(a[np.setdiff1d(np.arange(n), col)] == 0)


def test_fused_lds_limits(gpu):
    """Unions of about 9,900 voxels: with a column-sum region per wave 2
    waves per group fit 160 KiB of LDS (520 + 3 x 4 bytes per union voxel),
    4 waves do not and are refused; one shared region fits with 4 waves;
    bad shapes are refused and leave the plan's shape."""
    from raptr_amd import RaptrError
    from raptr_amd.device import DeviceArray
    from raptr_amd.recon import Reconstructor
    n = 20000
    rp, col, val = _long_row_csr(n, 640, 24, window=10000)
    R = Reconstructor(_device_csr(rp, col, val, n), DeviceArray.from_numpy(np.ones(n)), "f32", 1,
                      layout="blocked")
    R.set_union_rows(64)  # one column window per group
    assert R.unions()["rows"] == 64 and R.fused()["available"] == 1
    # the default: a shared region (79.7 KB of LDS: two workgroups per CU,
    # so 8 waves per group)
    assert R.fused()["on"] == 2 and R.fused()["waves"] == 8
    R.set_fused(2, 4, 16, 64)
    R.set_fused(2, 8, 64, 64)
    R.set_fused(True, 2, 32, 64)
    with pytest.raises(RaptrError):
        R.set_fused(True, 4, 16, 64)
    for o, w, lr, rw in [(1, 3, 16, 64), (1, 8, 16, 64), (1, 2, 8, 64), (1, 2, 128, 64),
                         (1, 2, 16, 8), (1, 2, 16, 128), (1, 0, 16, 64), (3, 2, 16, 64),
                         (-1, 2, 16, 64)]:
        with pytest.raises(RaptrError):
            R.set_fused(o, w, lr, rw)
    # ten groups: too few column-tile workgroups, the per-XCD reduce stays
    assert R.fused_tiles()["workgroups"] < 512
    assert R.fused() == {"available": 1, "on": 1, "waves": 2, "lanes_per_row": 32,
                         "lanes_per_column": 64, "per_xcd": 1, "steps": 2}
    for bad in (1, 5, 0):
        with pytest.raises(RaptrError):
            R.set_fused_steps(bad)
    R.set_fused(2, 8, 64, 64)
    R.set_fused(False)  # keeps the shape
    with pytest.raises(RaptrError):
        R.set_fused(True)  # 8 waves need the shared region
    assert R.fused()["on"] == 0
    x0 = np.full(n, 1.0 / n, np.float32)
    y = np.full(len(rp) - 1, 3.0, np.float32)
    a = R.mlem(y, x0, 2)
    R.set_fused(False)
    assert _rel(a, R.mlem(y, x0, 2)) <= 1e-6


def test_fused_absent(gpu, small_system):
    """No fused pass without column unions (64-frame plans, and
    single-frame plans whose groups touch more than 16,384 voxels) or when
    a row repeats a column (its column sums add each entry with one
    unguarded read-add-write): set_fused(on) is refused, off is accepted;
    the separate passes still add the repeated entries."""
    from raptr_amd import RaptrError
    from raptr_amd.device import DeviceArray
    from raptr_amd.recon import Reconstructor
    from test_gpu_recon import _wide_csr
    S = small_system
    R64 = Reconstructor(S["A"], S["sens"], "f32", 64, row_order=S["setup"], grid=S["grid"],
                        group=16)
    assert R64.fused()["available"] == 0
    with pytest.raises(RaptrError):
        R64.set_fused(True, 4, 16, 64)
    R64.set_fused(False, 4, 16, 64)
    n = 200000
    rp, col, val = _wide_csr(n, m=600)
    R = Reconstructor(_device_csr(rp, col, val, n), DeviceArray.from_numpy(np.ones(n)), "f32", 1,
                      layout="blocked")
    R.set_union_rows(64)  # groups of 64 wide rows touch more than 16,384 voxels
    assert R.unions()["rows"] == 0 and R.fused()["available"] == 0
    with pytest.raises(RaptrError):
        R.set_fused(True, 2, 16, 64)
    n = 20000
    rp, col, val = _long_row_csr(n, 300, 25)
    col = col.copy()
    col[rp[100] + 1] = col[rp[100]]  # row 100 holds its first column twice
    R = Reconstructor(_device_csr(rp, col, val, n), DeviceArray.from_numpy(np.ones(n)), "f32", 1,
                      layout="blocked")
    assert R.forward_kind()[0] == 2 and R.fused()["available"] == 0
    with pytest.raises(RaptrError):
        R.set_fused(True, 2, 64, 64)
    from raptr_amd import _capi as capi
    x = np.random.default_rng(3).random(n).astype(np.float32)
    m = len(rp) - 1
    y = np.full(m, 2.0, np.float32)
    dx, dy, de = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y), DeviceArray(m, np.float32)
    capi.call("raptr_recon_forward", R.handle, 1, dx.ptr, dy.ptr, de.ptr, None)
    yt = np.array([np.dot(val[rp[i]:rp[i + 1]].astype(np.float64), x[col[rp[i]:rp[i + 1]]])
                   for i in range(m)])
    np.testing.assert_allclose(de.numpy(), y / yt, rtol=1e-5)
