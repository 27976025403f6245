"""GPU parity of the HIP path (through the C ABI) against the oracle and the
reference's golden vectors.  Integer/index results and matrix values are
compared bit for bit; reductions within the stated tolerances."""
import hashlib

import numpy as np
import pytest

from conftest import load_golden, oracle_geom
from raptr_amd import geometry as G

pytestmark = pytest.mark.gpu


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, np.int32).tobytes())
    return h.hexdigest()


# ---------------------------------------------------------------- workqueue
def test_workqueue_small_matches_reference(gpu):
    """All 57,119,676 (channel, voxel) pairs of the 20,480-channel workload
    equal the reference's getWorkqueue output (hash + per-channel counts)."""
    from raptr_amd import sysmat
    gold = load_golden("wq_small.npz")
    setup, grid, _ = G.config_small()
    lst = np.arange(setup.nchannels, dtype=np.int32)
    counts = sysmat.workqueue_counts(setup, grid, lst)
    np.testing.assert_array_equal(counts, gold["counts"])
    c, v, li, vi = sysmat.workqueue(setup, grid, lst)
    assert len(c) == int(gold["total"])
    assert (li, vi) == (len(lst), 0)
    assert sha(c, v) == str(gold["sha256"])
    np.testing.assert_array_equal(c[:4096], gold["head_cnl"])
    np.testing.assert_array_equal(v[-4096:], gold["tail_vxl"])


def test_workqueue_full_angles_matches_reference(gpu):
    """The 180-angle, 52^3 geometry of the full-size workload: one channel
    per angle, every (channel, voxel) pair equal to the reference's
    getWorkqueue (its per-call float sinf/cosf against the shared table)."""
    from raptr_amd import sysmat
    from test_oracle import _full_angle_fixture
    gold, cnl, vxl = _full_angle_fixture()
    setup, grid, _ = G.config_full()
    counts = sysmat.workqueue_counts(setup, grid, gold["channels"])
    np.testing.assert_array_equal(counts, gold["counts"])
    c, v, li, vi = sysmat.workqueue(setup, grid, gold["channels"])
    assert (li, vi) == (len(gold["channels"]), 0)
    assert sha(c, v) == str(gold["sha256"])
    np.testing.assert_array_equal(v, vxl)


def test_pixel_centres_full_angles_match_reference(gpu):
    """Pixel-centre rays (REGU) of the 180-angle setup equal the reference
    trafo at the unit-cube centre for both heads at every angle."""
    from raptr_amd import sysmat, RAYGEN_REGU
    gold = load_golden("wq_full_angles.npz")
    setup, _, _ = G.config_full()
    pix, out = gold["trafo_pix"], gold["trafo_out"]
    assert np.array_equal(gold["trafo_pts"][0], [0.5, 0.5, 0.5])
    chans, want = [], []
    for a in range(setup.na):
        for (iz0, iy0) in pix:
            for (iz1, iy1) in pix:
                chans.append(setup.lin_id(iz0, iy0, iz1, iy1, a))
                j0 = [tuple(p) for p in pix].index((iz0, iy0))
                j1 = [tuple(p) for p in pix].index((iz1, iy1))
                want.append(np.concatenate([out[0, a, j0, 0], out[1, a, j1, 0]]))
    r = sysmat.generate_rays(setup, np.asarray(chans, np.int32), 1, raygen=RAYGEN_REGU)
    np.testing.assert_array_equal(r[:, 0], np.asarray(want, np.float32))


def test_workqueue_bounded_cursor_matches_reference(gpu):
    from raptr_amd import sysmat
    gold = load_golden("wq_small.npz")
    setup, grid, _ = G.config_small()
    lst = np.arange(setup.nchannels, dtype=np.int32)
    c, v, li, vi = sysmat.workqueue(setup, grid, lst, list_id=3, vxl_id=17, n=100000)
    assert len(c) == int(gold["bounded_n"])
    assert sha(c, v) == str(gold["bounded_sha256"])
    assert [li, vi] == list(gold["bounded_cursor"])


def test_workqueue_prefix_resume_tiny(gpu, oracle):
    """Bounded calls are prefixes of the unbounded queue and resume from the
    returned cursor (test_getWorkqueue.cu:124-219), edge cases included."""
    from raptr_amd import sysmat
    gold = load_golden("wq_tiny.npz")
    setup, grid, _ = G.config_tiny()
    lst = np.arange(setup.nchannels, dtype=np.int32)
    c, v, li, vi = sysmat.workqueue(setup, grid, lst)
    np.testing.assert_array_equal(c, gold["cnl"])
    np.testing.assert_array_equal(v, gold["vxl"])
    s, g = oracle_geom(oracle, setup, grid)
    for (l0, v0, n) in [(0, 0, 1), (0, 0, 4097), (5, 4095, 10), (10, 123, 0), (767, 4000, 50),
                        (700, 0, 10 ** 6)]:
        gc, gv, gli, gvi = sysmat.workqueue(setup, grid, lst, list_id=l0, vxl_id=v0, n=n)
        oc, ov, oli, ovi = oracle.workqueue(s, g, lst, list_id=l0, vxl_id=v0, n=n)
        np.testing.assert_array_equal(gc, oc)
        np.testing.assert_array_equal(gv, ov)
        assert (gli, gvi) == (oli, ovi)
    # empty list and cursor at the end
    gc, gv, gli, gvi = sysmat.workqueue(setup, grid, lst, list_id=len(lst), vxl_id=0)
    assert len(gc) == 0 and (gli, gvi) == (len(lst), 0)


def test_workqueue_rejects_bad_cursor(gpu):
    from raptr_amd import sysmat, RaptrError
    setup, grid, _ = G.config_tiny()
    with pytest.raises(RaptrError):
        sysmat.workqueue(setup, grid, np.arange(10, dtype=np.int32), list_id=0,
                         vxl_id=grid.nvox)


# ------------------------------------------------------------------- rays
def test_rays_match_oracle(gpu, oracle):
    from raptr_amd import sysmat, RAYGEN_REGU
    setup, grid, nrays = G.config_small()
    s, g = oracle_geom(oracle, setup, grid)
    chans = np.array([0, 1, 777, 20479, 12345], np.int32)
    r = sysmat.generate_rays(setup, chans, nrays, seed=1)
    for i, c in enumerate(chans):
        for k in range(nrays):
            np.testing.assert_array_equal(r[i, k], oracle.ray(s, int(c), k, seed=1))
    r = sysmat.generate_rays(setup, chans, 2, seed=1, raygen=RAYGEN_REGU)
    np.testing.assert_array_equal(r[3, 1], oracle.ray(s, 20479, 1, raygen=oracle.RAYGEN_REGU))


def test_backprojection_matches_oracle(gpu, oracle):
    """Matrix-free summed back-projection A^T y against the oracle's FP64
    transposed product on the oracle's FP32 matrix values."""
    from raptr_amd import sysmat
    setup, grid, nrays = G.config_small()
    s, g = oracle_geom(oracle, setup, grid)
    rng = np.random.default_rng(5)
    chans = np.sort(rng.choice(setup.nchannels, 3000, replace=False)).astype(np.int32)
    y = rng.poisson(7.0, len(chans)).astype(np.float32)
    rp, col, val = oracle.build_csr(s, g, chans, nrays)[:3]
    want = oracle.csrmv_f64(1, len(chans), grid.nvox, 1.0, val.astype(np.float64), rp, col,
                            y.astype(np.float64), 0.0, np.zeros(grid.nvox))
    got = sysmat.backproject(setup, grid, nrays, chans, y, normalise=False)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-9)
    xn = sysmat.backproject(setup, grid, nrays, chans, y)
    assert abs(xn.sum() - 1.0) < 1e-12


# ------------------------------------------------------------ Siddon chords
@pytest.mark.parametrize("cfg", ["tiny", "small"])
def test_get_chords_bit_exact(gpu, oracle, cfg):
    """GPU getChords equals the host restatement bit for bit: sampled
    channel rays, axis-parallel and diagonal rays, a miss, a ray starting
    inside the grid, and a capacity smaller than the traversal."""
    from raptr_amd import sysmat
    setup, grid, nrays = G.config_tiny() if cfg == "tiny" else G.config_small()
    s, g = oracle_geom(oracle, setup, grid)
    rays = [oracle.ray(s, int(c), k) for c in range(0, setup.nchannels, 97)
            for k in range(nrays)]
    lo = np.array(grid.o, np.float32)
    hi = lo + np.array(grid.d, np.float32) * np.array(grid.n, np.float32)
    mid = 0.5 * (lo + hi)
    ext = [
        [lo[0] - 10, mid[1] + 0.3, mid[2] - 0.7, hi[0] + 10, mid[1] + 0.3, mid[2] - 0.7],
        [*(lo - 5), *(hi + 5)],
        [lo[0] - 10, hi[1] + 50, mid[2], hi[0] + 10, hi[1] + 50, mid[2]],  # miss
        [*mid, *(hi + 5)],  # starts inside
        [hi[0] + 7, lo[1] - 3, mid[2] + 1.1, lo[0] - 9, hi[1] + 4, mid[2] - 2.3],
    ]
    rays = np.array(rays + ext, np.float32)
    for cap in (sum(grid.n) + 4, 3):
        ch, ids, cnt = sysmat.get_chords(rays, grid, cap=cap)
        for r, ray in enumerate(rays):
            wc, wi = oracle.get_chords(ray, g, cap=cap)
            assert cnt[r] == len(wc), r
            np.testing.assert_array_equal(ch[r, :cnt[r]], wc)
            np.testing.assert_array_equal(ids[r, :cnt[r]], wi)
    # the closed forms of the extra rays
    ch, ids, cnt = sysmat.get_chords(rays[-5:], grid)
    assert cnt[2] == 0 and cnt[3] == 0
    assert cnt[0] == grid.n[0] and np.allclose(ch[0, :cnt[0]], grid.d[0], rtol=1e-4)
    assert ch[1, :cnt[1]].sum() == pytest.approx(
        float(np.linalg.norm(hi - lo)), rel=1e-4)


# ------------------------------------------------------- matrix from workqueue
@pytest.mark.parametrize("cfg", ["tiny", "small"])
def test_sm_from_wq_bit_exact(gpu, oracle, cfg):
    from raptr_amd import sysmat
    setup, grid, nrays = G.config_tiny() if cfg == "tiny" else G.config_small()
    s, g = oracle_geom(oracle, setup, grid)
    step = 5 if cfg == "tiny" else 97
    chans = np.arange(0, setup.nchannels, step, dtype=np.int32)
    c, v, _, _ = oracle.workqueue(s, g, chans)
    want = oracle.sm_from_wq(s, g, c, v, nrays)
    got = sysmat.sm_from_wq(setup, grid, c, v, nrays)
    np.testing.assert_array_equal(got, want)
    # FP64 oracle on identical rays: 1e-5 relative (absolute floor for
    # elements decided by a near-tangent ray)
    a64 = oracle.sm_from_wq(s, g, c, v, nrays, f64=True)
    assert np.all(np.abs(got - a64) <= 1e-5 * np.abs(a64) + 1e-5 * grid.d[0])


def test_sm_from_wq_invalid_ids_give_zero(gpu):
    from raptr_amd import sysmat
    setup, grid, nrays = G.config_tiny()
    got = sysmat.sm_from_wq(setup, grid, [-1, setup.nchannels, 3], [0, 0, grid.nvox], nrays)
    np.testing.assert_array_equal(got, [0, 0, 0])


# ------------------------------------------------------------ fused build
@pytest.mark.parametrize("cfg", ["tiny", "small"])
def test_build_csr_bit_exact(gpu, oracle, cfg):
    """Fused cull+sample+compact equals the oracle's matrix of the measured
    rows (= two-stage values with exact zeros dropped), bit for bit."""
    from raptr_amd import sysmat
    setup, grid, nrays = G.config_tiny() if cfg == "tiny" else G.config_small()
    s, g = oracle_geom(oracle, setup, grid)
    rng = np.random.default_rng(11)
    n_rows = 400 if cfg == "tiny" else 600
    lst = np.sort(rng.choice(setup.nchannels, n_rows, replace=False)).astype(np.int32)
    rp, col, val, _ = oracle.build_csr(s, g, lst, nrays)
    A = sysmat.build_csr(setup, grid, lst, nrays, staging=1 << 16)
    grp, gcol, gval = A.to_host()
    np.testing.assert_array_equal(grp, rp)
    np.testing.assert_array_equal(gcol, col)
    np.testing.assert_array_equal(gval, val)
    # capacity-based single pass gives the same matrix
    B = sysmat.build_csr(setup, grid, lst, nrays, capacity=A.nnz + 10)
    np.testing.assert_array_equal(B.to_host()[2], val)


@pytest.mark.parametrize("cfg", ["tiny", "small"])
def test_build_csr_k2_modes_bit_exact(gpu, oracle, cfg):
    """The fused kernel's options: zero-valued candidates kept
    (getSystemMatrixDeviceOnly.cu:507-509) and per-element ray streams
    (:396), alone and together, equal the oracle bit for bit; with zeros kept
    the pairs are exactly the workqueue's."""
    from raptr_amd import sysmat
    setup, grid, nrays = G.config_tiny() if cfg == "tiny" else G.config_small()
    s, g = oracle_geom(oracle, setup, grid)
    rng = np.random.default_rng(5)
    lst = np.sort(rng.choice(setup.nchannels, 300, replace=False)).astype(np.int32)
    cw, vw, _, _ = oracle.workqueue(s, g, lst)
    for raygen, keep in ((oracle.RAYGEN_RAND, True), (oracle.RAYGEN_RAND_ELEMENT, False),
                         (oracle.RAYGEN_RAND_ELEMENT, True), (oracle.RAYGEN_REGU, True)):
        rp, col, val, _ = oracle.build_csr(s, g, lst, nrays, raygen=raygen,
                                           flags=oracle.SM_KEEP_ZEROS if keep else 0)
        A = sysmat.build_csr(setup, grid, lst, nrays, raygen=raygen, keep_zeros=keep,
                             staging=1 << 16)
        grp, gcol, gval = A.to_host()
        np.testing.assert_array_equal(grp, rp)
        np.testing.assert_array_equal(gcol, col)
        np.testing.assert_array_equal(gval, val)
        if keep:
            assert A.nnz == len(vw)
            np.testing.assert_array_equal(gcol, vw)
    # two-stage path in element mode
    got = sysmat.sm_from_wq(setup, grid, cw, vw, nrays, raygen=oracle.RAYGEN_RAND_ELEMENT)
    want = oracle.sm_from_wq(s, g, cw, vw, nrays, raygen=oracle.RAYGEN_RAND_ELEMENT)
    np.testing.assert_array_equal(got, want)


def test_build_keep_zeros_small_is_reference_workqueue(gpu):
    """All 20,480 channels with zeros kept: nnz and the SHA-256 of the
    (row channel, column) pairs equal the reference's getWorkqueue output
    (57,119,676 pairs), i.e. the reference's nnz."""
    from raptr_amd import sysmat
    gold = load_golden("wq_small.npz")
    setup, grid, nrays = G.config_small()
    lst = np.arange(setup.nchannels, dtype=np.int32)
    A = sysmat.build_csr(setup, grid, lst, nrays, keep_zeros=True)
    assert A.nnz == int(gold["total"])
    rp, col, _ = A.to_host()
    np.testing.assert_array_equal(np.diff(rp), gold["counts"])
    assert sha(np.repeat(lst, np.diff(rp)), col) == str(gold["sha256"])


def test_build_csr_chunking_independent(gpu):
    """Small staging areas force many chunks and piece flushes; result equal."""
    from raptr_amd import sysmat
    setup, grid, nrays = G.config_small()
    lst = np.arange(0, setup.nchannels, 7, dtype=np.int32)
    A = sysmat.build_csr(setup, grid, lst, nrays, staging=1 << 22)
    B = sysmat.build_csr(setup, grid, lst, nrays, staging=3072 + 64)
    for a, b in zip(A.to_host(), B.to_host()):
        np.testing.assert_array_equal(a, b)


def test_build_csr_matches_two_stage(gpu):
    """Fused path == workqueue + sm_from_wq with zeros dropped (GPU only)."""
    from raptr_amd import sysmat
    setup, grid, nrays = G.config_small()
    lst = np.arange(100, 20480, 211, dtype=np.int32)
    c, v, _, _ = sysmat.workqueue(setup, grid, lst)
    val = sysmat.sm_from_wq(setup, grid, c, v, nrays)
    nz = val != 0
    rp, col, gval = sysmat.build_csr(setup, grid, lst, nrays).to_host()
    np.testing.assert_array_equal(col, v[nz])
    np.testing.assert_array_equal(gval, val[nz])
    pos = np.searchsorted(lst, c[nz])
    np.testing.assert_array_equal(rp[:-1], np.searchsorted(pos, np.arange(len(lst))))


def test_sensitivity_bit_reproducible_and_exact(gpu, oracle):
    """The FP64 sensitivity of the 20,480-channel workload (all channels)
    accumulates in fixed point: two computations are bit-identical, and both
    equal the exactly summed column sums of the oracle's float matrix to
    1e-13 (FP64 atomics would give order-dependent last bits)."""
    import math
    from raptr_amd import sysmat
    setup, grid, nrays = G.config_small()
    s, g = oracle_geom(oracle, setup, grid)
    a = sysmat.sensitivity(setup, grid, nrays).numpy()
    b = sysmat.sensitivity(setup, grid, nrays).numpy()
    np.testing.assert_array_equal(a, b)
    lst = np.arange(setup.nchannels, dtype=np.int32)
    rp, col, val, _ = oracle.build_csr(s, g, lst, nrays)
    order = np.argsort(col, kind="stable")
    cs, vs = col[order], val[order].astype(np.float64)
    starts = np.searchsorted(cs, np.arange(grid.nvox + 1))
    exact = np.array([math.fsum(vs[starts[j]:starts[j + 1]]) for j in range(grid.nvox)])
    np.testing.assert_allclose(a, exact, rtol=1e-13, atol=1e-13)


def test_sensitivity_and_forward_projection(gpu, oracle):
    from raptr_amd import sysmat
    from raptr_amd.device import DeviceArray
    setup, grid, nrays = G.config_tiny()
    s, g = oracle_geom(oracle, setup, grid)
    lst = np.arange(setup.nchannels, dtype=np.int32)
    rp, col, val, v64 = oracle.build_csr(s, g, lst, nrays, with_f64=True)
    sens_o = np.zeros(grid.nvox)
    np.add.at(sens_o, col, val.astype(np.float64))
    sens = sysmat.sensitivity(setup, grid, nrays).numpy()
    np.testing.assert_allclose(sens, sens_o, rtol=1e-12, atol=1e-12)
    x = np.random.default_rng(2).random(grid.nvox).astype(np.float32)
    fp = sysmat.forward_project(setup, grid, x, nrays)
    want = np.array([np.dot(val[rp[i]:rp[i + 1]].astype(np.float64),
                            x[col[rp[i]:rp[i + 1]]].astype(np.float64)) for i in range(len(lst))])
    np.testing.assert_allclose(fp, want, rtol=1e-12, atol=1e-12)
    sub = lst[::3]
    np.testing.assert_allclose(sysmat.forward_project(setup, grid, DeviceArray.from_numpy(x),
                                                      nrays, cnl_list=sub), want[::3], rtol=1e-12)


# ------------------------------------------------------------ sparse formats
def test_coo_sort_reference_cases(gpu):
    from raptr_amd import sparse
    v, r, c = sparse.coo_sort([4, 1, 3, 2], [1, 0, 1, 0], [1, 0, 0, 1])
    np.testing.assert_array_equal(r, [0, 0, 1, 1])
    np.testing.assert_array_equal(c, [0, 1, 0, 1])
    np.testing.assert_array_equal(v, [1, 2, 3, 4])
    v, r, c = sparse.coo_sort([], [], [])
    assert len(v) == 0
    v, r, c = sparse.coo_sort([7.0], [3], [2])
    assert (v[0], r[0], c[0]) == (7.0, 3, 2)


def test_coo_sort_random_stable(gpu, oracle):
    from raptr_amd import sparse
    rng = np.random.default_rng(9)
    n = 300001
    row = rng.integers(0, 50000, n).astype(np.int32)
    col = rng.integers(0, 3000, n).astype(np.int32)
    val = rng.random(n).astype(np.float32)
    v, r, c = sparse.coo_sort(val, row, col)
    order = np.lexsort((np.arange(n), col, row))  # stable on (row, col)
    np.testing.assert_array_equal(r, row[order])
    np.testing.assert_array_equal(c, col[order])
    np.testing.assert_array_equal(v, val[order])
    ov, orr, oc = oracle.coo_sort(val, row, col)
    np.testing.assert_array_equal(v, ov)


def test_coo2csr_and_ecsr_known_answers(gpu):
    from raptr_amd import sparse
    np.testing.assert_array_equal(sparse.coo2csr([0, 1, 1, 1, 2, 2, 3, 3, 3], 4), [0, 1, 4, 6, 9])
    np.testing.assert_array_equal(sparse.csr2ecsr([2, 4, 6], [0, 0, 0, 3, 3, 3, 3, 5], 7),
                                  [0, 3, 3, 5])
    np.testing.assert_array_equal(sparse.coo2csr([], 3), [0, 0, 0, 0])


def test_transpose_exact(gpu):
    from raptr_amd import sparse, sysmat
    setup, grid, nrays = G.config_small()
    lst = np.arange(0, setup.nchannels, 5, dtype=np.int32)
    A = sysmat.build_csr(setup, grid, lst, nrays)
    At = sparse.transpose(A)
    rp, col, val = A.to_host()
    trp, tcol, tval = At.to_host()
    rows = np.repeat(np.arange(A.m), np.diff(rp))
    order = np.lexsort((rows, col))
    np.testing.assert_array_equal(tcol, rows[order])
    np.testing.assert_array_equal(tval, val[order])
    np.testing.assert_array_equal(trp, np.searchsorted(col[order], np.arange(grid.nvox + 1)))


# ----------------------------------------------------------------- csrmv
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_csrmv_dense(gpu, dtype):
    from raptr_amd import sparse
    rng = np.random.default_rng(3)
    for m, n, dens in [(37, 53, 0.3), (2000, 700, 0.05), (300, 20000, 0.01)]:
        D = rng.random((m, n)) * (rng.random((m, n)) < dens)
        rp = np.concatenate([[0], np.cumsum((D != 0).sum(1))]).astype(np.int32)
        rr, cc = np.nonzero(D)
        val = D[rr, cc]
        x = rng.random(n)
        y0 = rng.random(m)
        tol = 2e-5 if dtype == np.float32 else 1e-12
        got = sparse.csrmv(0, m, n, 1.5, val, rp, cc, x, 0.5, y0, dtype=dtype)
        np.testing.assert_allclose(got, 1.5 * D @ x + 0.5 * y0, rtol=tol)
        got = sparse.csrmv(0, m, n, 1.0, val, rp, cc, x, 0.0, np.full(m, np.nan), dtype=dtype)
        np.testing.assert_allclose(got, D @ x, rtol=tol)
        xt, yt = rng.random(m), rng.random(n)
        got = sparse.csrmv(1, m, n, 2.0, val, rp, cc, xt, 1.0, yt, dtype=dtype)
        np.testing.assert_allclose(got, 2.0 * D.T @ xt + yt, rtol=tol)


# -------------------------------------------------------------- MLEM ops
def test_mlem_operations_known_answers(gpu):
    from raptr_amd import mlem
    i = np.arange(10, dtype=np.float32)
    A, B, Cc = (i + 1) ** 2, np.full(10, 10, np.float32), i + 1
    np.testing.assert_array_equal(mlem.divides(A, Cc), i + 1)
    D = mlem.divides_multiplies(A, B, Cc)
    np.testing.assert_array_equal(D, 10 * (i + 1))
    assert mlem.fsum(D) == 550.0
    np.testing.assert_allclose(mlem.scales(D, np.float32(1.0 / 550.0)), (i + 1) / 55, rtol=1e-6)
    np.testing.assert_array_equal(mlem.divides_multiplies([2.0], [3.0], [0.0]), [6.0])
    x = np.random.default_rng(0).random(1_000_003).astype(np.float32)
    assert mlem.fsum(x) == pytest.approx(float(x.astype(np.float64).sum()), rel=1e-7)


def test_mlem_ops_bit_exact_vs_oracle(gpu, oracle):
    from raptr_amd import mlem
    rng = np.random.default_rng(8)
    a, b, c = (rng.random(100_000).astype(np.float32) + 0.01 for _ in range(3))
    c[::17] = 0
    np.testing.assert_array_equal(mlem.divides(a, b), oracle.divides(a, b))
    np.testing.assert_array_equal(mlem.divides_multiplies(a, b, c),
                                  oracle.divides_multiplies(a, b, c))
    np.testing.assert_array_equal(mlem.scales(a, 0.37), oracle.scales(a, np.float32(0.37)))


# ------------------------------------------------------------------- MLEM
def _phantom_small(grid):
    c = grid.centres()
    x = np.zeros(grid.nvox)
    x[np.linalg.norm(c - [-10, 0, 0], axis=1) < 15] = 1.0
    x[np.linalg.norm(c - [20, 0, 0], axis=1) < 8] = 4.0
    return x


def test_mlem_vs_fp64_oracle_small(gpu, oracle):
    """20,480-channel / 32^3 workload, 2e5 Poisson events, 5 iterations:
    GPU FP32 MLEM within 1e-4 relative L2 of the FP64 host recurrence on the
    same matrix (FP64 values on identical rays) and the same counts."""
    from raptr_amd import mlem, sparse, sysmat
    setup, grid, nrays = G.config_small()
    s, g = oracle_geom(oracle, setup, grid)
    allc = np.arange(setup.nchannels, dtype=np.int32)
    rp, col, val, v64 = oracle.build_csr(s, g, allc, nrays, with_f64=True)
    xt = _phantom_small(grid)
    ybar = oracle.csrmv_f64(0, len(allc), grid.nvox, 1.0, v64, rp, col, xt, 0.0,
                            np.zeros(len(allc)))
    rng = np.random.Generator(np.random.Philox(key=1))
    y = rng.poisson(ybar * 2e5 / ybar.sum()).astype(np.float64)
    meas = np.nonzero(y > 0)[0].astype(np.int32)
    sens64 = oracle.sensitivity_f64(rp, col, v64, grid.nvox)
    mrp = np.concatenate([[0], np.cumsum(np.diff(rp)[meas])])
    msel = np.concatenate([np.arange(rp[i], rp[i + 1]) for i in meas])
    x0 = np.full(grid.nvox, 1.0 / grid.nvox)
    x64 = oracle.mlem_f64(mrp, col[msel], v64[msel], y[meas], sens64, x0, 5)
    A = sysmat.build_csr(setup, grid, meas, nrays)
    At = sparse.transpose(A)
    sens = sysmat.sensitivity(setup, grid, nrays)
    for use_graph in (False, True):
        x = mlem.mlem(A, At, y[meas].astype(np.float32), sens, x0.astype(np.float32), 5,
                      use_graph=use_graph)
        rel = np.linalg.norm(x - x64) / np.linalg.norm(x64)
        assert rel <= 1e-4, rel
        assert abs(float(x.astype(np.float64).sum()) - 1.0) < 1e-5


def test_mlem_graph_replay_deterministic(gpu):
    from raptr_amd import mlem, sparse, sysmat
    setup, grid, nrays = G.config_tiny()
    lst = np.arange(0, setup.nchannels, 2, dtype=np.int32)
    A = sysmat.build_csr(setup, grid, lst, nrays)
    At = sparse.transpose(A)
    sens = sysmat.sensitivity(setup, grid, nrays)
    y = np.random.default_rng(1).poisson(5.0, len(lst)).astype(np.float32)
    x0 = np.full(grid.nvox, 1.0 / grid.nvox, np.float32)
    a = mlem.mlem(A, At, y, sens, x0, 7, use_graph=False)
    b = mlem.mlem(A, At, y, sens, x0, 7, use_graph=True)
    c = mlem.mlem(A, At, y, sens, x0, 7, use_graph=True)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(b, c)
