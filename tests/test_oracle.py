"""CPU checks that pin the oracle (the restatement of the reference) against
the reference's own code, its known-answer tests, and closed forms."""
import hashlib

import numpy as np
import pytest

from conftest import load_golden, oracle_geom
from raptr_amd import geometry as G


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, np.int32).tobytes())
    return h.hexdigest()


# ------------------------------------------------------------- Philox
def test_philox_known_answers(oracle):
    # Random123 known-answer vectors for philox4x32-10
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, want in kat:
        assert list(oracle.philox(ctr, key)) == want


def test_uniform_range(oracle):
    assert oracle.lib().ro_u01(0) == pytest.approx(2.0 ** -24)
    assert oracle.lib().ro_u01(0xffffffff) == 1.0


# ------------------------------------------------------ index maps (A2/A3)
def test_channel_index_round_trip(oracle):
    # test_MeasurementSetupLinIndex.cu:57-66 (round trip of all channel ids)
    setup, grid, _ = G.config_small()
    s, _ = oracle_geom(oracle, setup, grid)
    L = oracle.lib()
    import ctypes as C
    ids = [C.c_int() for _ in range(5)]
    for cnl in list(range(0, setup.nchannels, 97)) + [setup.nchannels - 1]:
        L.ro_cnl_split(C.byref(s), cnl, *[C.byref(i) for i in ids])
        ida, id0z, id0y, id1z, id1y = (i.value for i in ids)
        assert L.ro_cnl_lin(C.byref(s), id0z, id0y, id1z, id1y, ida) == cnl
        assert (ida, id0z, id0y, id1z, id1y) == tuple(int(v) for v in setup.split(cnl))


def test_voxel_index_round_trip(oracle):
    # test_VoxelGridLinIndex.cu:57-64; z fastest (VoxelGridLinIndex.hpp:73-80)
    import ctypes as C
    grid = G.VoxelGrid((0, 0, 0), (1, 1, 1), (5, 7, 3))
    g = oracle.make_grid(grid.o, grid.d, grid.n)
    L = oracle.lib()
    ix, iy, iz = C.c_int(), C.c_int(), C.c_int()
    for v in range(grid.nvox):
        L.ro_vxl_split(C.byref(g), v, C.byref(ix), C.byref(iy), C.byref(iz))
        assert v == (ix.value * 7 + iy.value) * 3 + iz.value
        assert L.ro_vxl_lin(C.byref(g), ix.value, iy.value, iz.value) == v


# ------------------------------------------------------------- trafo (A4)
def test_trafo_matches_reference_build(oracle):
    """Bit-exact against the reference's trafo functors compiled from source."""
    gold = load_golden("trafo_ref.npz")
    sf, si, pts, out = gold["setup_f"], gold["setup_i"], gold["pts"], gold["out"]
    s = oracle.make_setup(sf[0], sf[1], si[0], si[1], si[2], si[3], si[4], sf[2], sf[3], sf[4],
                          sf[5])
    for det in range(2):
        for a in range(int(si[0])):
            for iz in range(4):
                for iy in range(4):
                    for k, p in enumerate(pts):
                        got = oracle.trafo(s, det, a, iz, iy, p)
                        np.testing.assert_array_equal(got, out[det, a, iz, iy, k])


def test_trafo_closed_form(oracle):
    """MeasurementSetupTrafo2CartCoordUT.cu:91-137 (no rotation) and :309-320
    (rotation about y), in double precision with a tolerance."""
    s = oracle.make_setup(-35.0, 35.0, 6, 4, 4, 4, 4, 60.0, 10.0, 10.0, 10.0)
    rng = np.random.default_rng(3)
    for det, pos in ((0, -35.0), (1, 35.0)):
        for _ in range(20):
            iz, iy = rng.integers(0, 4, 2)
            b = rng.random(3).astype(np.float32)
            x = pos + (-0.5 + b[0]) * 10.0
            y = ((-0.5 * 4 + iy) + b[1]) * 10.0
            z = ((-0.5 * 4 + iz) + b[2]) * 10.0
            np.testing.assert_allclose(oracle.trafo(s, det, 0, iz, iy, b), [x, y, z], atol=2e-5)
            for a in range(1, 6):
                th = np.deg2rad(60.0 * a)
                want = [np.sin(th) * z + np.cos(th) * x, y, np.cos(th) * z - np.sin(th) * x]
                np.testing.assert_allclose(oracle.trafo(s, det, a, iz, iy, b), want, atol=5e-5)


def test_angle_table_matches_libm(oracle):
    """The shared sin/cos table (oracle and the HIP library's host code)
    equals the C library's float sinf/cosf of the reference's float argument
    for every angle of the named workloads; the reference-built trafo at all
    180 angles (test_trafo_full_angles_matches_reference_build) pins that
    this is the reference's evaluation."""
    import ctypes as C
    libm = C.CDLL("libm.so.6")
    for f in (libm.sinf, libm.cosf):
        f.restype, f.argtypes = C.c_float, [C.c_float]
    for setup, _, _ in (G.config_small(), G.config_full(), G.config_tiny()):
        s = oracle.make_setup(setup.pos0x, setup.pos1x, setup.na, setup.n0z, setup.n0y,
                              setup.n1z, setup.n1y, setup.da, setup.segx, setup.segy,
                              setup.segz)
        sn, cs = oracle.angle_table(s)
        arg = (np.float32(np.pi) * np.arange(setup.na, dtype=np.float32)) * np.float32(setup.da)
        arg = arg / np.float32(180.0)
        np.testing.assert_array_equal(sn, np.array([libm.sinf(float(a)) for a in arg], np.float32))
        np.testing.assert_array_equal(cs, np.array([libm.cosf(float(a)) for a in arg], np.float32))
        from raptr_amd import _capi as capi
        hs, hc = np.zeros(setup.na, np.float32), np.zeros(setup.na, np.float32)
        capi.call("raptr_angle_table", C.byref(setup.struct()), hs.ctypes.data_as(C.POINTER(C.c_float)),
                  hc.ctypes.data_as(C.POINTER(C.c_float)))
        np.testing.assert_array_equal(hs, sn)
        np.testing.assert_array_equal(hc, cs)


def _full_angle_fixture():
    gold = load_golden("wq_full_angles.npz")
    counts = gold["counts"].astype(np.int64)
    cnl = np.repeat(gold["channels"], counts)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    vxl = np.empty(len(cnl), np.int64)
    for s0, k in zip(starts, counts):
        vxl[s0:s0 + k] = np.cumsum(gold["vxl_delta"][s0:s0 + k].astype(np.int64))
    return gold, cnl.astype(np.int32), vxl.astype(np.int32)


def test_trafo_full_angles_matches_reference_build(oracle):
    """All 180 one-degree angles of the full-size setup: the shared sin/cos
    table gives the reference trafo (per-call float sinf/cosf,
    MeasurementSetupTrafo2CartCoord.hpp:110-111) bit for bit."""
    gold = load_golden("wq_full_angles.npz")
    setup, _, _ = G.config_full()
    s = oracle.make_setup(setup.pos0x, setup.pos1x, setup.na, setup.n0z, setup.n0y,
                          setup.n1z, setup.n1y, setup.da, setup.segx, setup.segy, setup.segz)
    out = gold["trafo_out"]
    assert out.shape[1] == 180
    for det in range(2):
        for a in range(setup.na):
            for j, (iz, iy) in enumerate(gold["trafo_pix"]):
                for k, p in enumerate(gold["trafo_pts"]):
                    np.testing.assert_array_equal(oracle.trafo(s, det, a, int(iz), int(iy), p),
                                                  out[det, a, j, k])


def test_workqueue_full_angles_matches_reference(oracle):
    """One channel per angle of the 180-angle, 52^3 geometry: every (channel,
    voxel) decision equals the reference's getWorkqueue, including the pairs
    within 1e-4 mm of the cylinder radius."""
    gold, cnl, vxl = _full_angle_fixture()
    assert sha(cnl, vxl) == str(gold["sha256"])
    setup, grid, _ = G.config_full()
    s, g = oracle_geom(oracle, setup, grid)
    c, v, li, vi = oracle.workqueue(s, g, gold["channels"], capacity=len(cnl) + 16)
    assert (li, vi) == (len(gold["channels"]), 0)
    np.testing.assert_array_equal(c, cnl)
    np.testing.assert_array_equal(v, vxl)
    key = c.astype(np.int64) * grid.nvox + v
    near = gold["near"]
    got_in = np.isin(near[:, 0].astype(np.int64) * grid.nvox + near[:, 1], key)
    np.testing.assert_array_equal(got_in.astype(np.int8), gold["near_in"])


# ------------------------------------------------------------ distance (A5)
def test_distance_textbook(oracle):
    import ctypes as C
    def d(a, b, p):
        a, b, p = (np.asarray(v, np.float32) for v in (a, b, p))
        return oracle.lib().ro_distance(oracle._p(a, C.c_float), oracle._p(b, C.c_float),
                                        oracle._p(p, C.c_float))
    assert d([0, 0, 0], [1, 0, 0], [5, 3, 4]) == pytest.approx(5.0)
    assert d([0, 0, 0], [0, 2, 0], [7, -3, 0]) == pytest.approx(7.0)
    assert d([1, 1, 1], [2, 2, 2], [3, 3, 3]) == pytest.approx(0.0, abs=1e-6)
    rng = np.random.default_rng(0)
    for _ in range(200):
        a, b, p = rng.normal(size=(3, 3)) * 50
        want = np.linalg.norm(np.cross(b - a, p - a)) / np.linalg.norm(b - a)
        assert d(a, b, p) == pytest.approx(want, rel=1e-4, abs=1e-4)


# ----------------------------------------------------------- workqueue (A6/A7)
def test_workqueue_matches_reference_golden(oracle):
    """Every (channel, voxel) decision of the 20,480-channel workload equals the
    reference's getWorkqueue compiled from its sources."""
    gold = load_golden("wq_small.npz")
    setup, grid, _ = G.config_small()
    s, g = oracle_geom(oracle, setup, grid)
    lst = np.arange(setup.nchannels, dtype=np.int32)
    counts = oracle.workqueue_counts(s, g, lst)
    np.testing.assert_array_equal(counts, gold["counts"])
    assert counts.sum() == int(gold["total"])
    c, v, li, vi = oracle.workqueue(s, g, lst[:96])
    np.testing.assert_array_equal(c[:4096], gold["head_cnl"])
    np.testing.assert_array_equal(v[:4096], gold["head_vxl"])
    tail_c0 = int(gold["tail_cnl"][0])
    c, v, _, _ = oracle.workqueue(s, g, lst[tail_c0:])
    np.testing.assert_array_equal(c[-4096:], gold["tail_cnl"])
    np.testing.assert_array_equal(v[-4096:], gold["tail_vxl"])


def test_workqueue_near_boundary_decisions(oracle):
    """Pairs within 1e-4 mm of the cylinder radius decide as the reference."""
    gold = load_golden("wq_small.npz")
    near, inside = gold["near"], gold["near_in"]
    setup, grid, _ = G.config_small()
    s, g = oracle_geom(oracle, setup, grid)
    chans = np.unique(near[:, 0])
    c, v, _, _ = oracle.workqueue(s, g, chans)
    keys = set((c.astype(np.int64) * grid.nvox + v).tolist())
    got = np.array([(int(a) * grid.nvox + int(b)) in keys for a, b in near], np.int8)
    np.testing.assert_array_equal(got, inside)


def test_workqueue_bounded_prefix_and_cursor(oracle):
    """test_getWorkqueue.cu:124-219: the n-bounded queue is a prefix of the
    unbounded one; resuming from the returned cursor continues it."""
    gold = load_golden("wq_small.npz")
    setup, grid, _ = G.config_small()
    s, g = oracle_geom(oracle, setup, grid)
    lst = np.arange(setup.nchannels, dtype=np.int32)
    c, v, li, vi = oracle.workqueue(s, g, lst, list_id=3, vxl_id=17, n=100000)
    assert sha(c, v) == str(gold["bounded_sha256"])
    assert [li, vi] == list(gold["bounded_cursor"])
    full_c, full_v, _, _ = oracle.workqueue(s, g, lst[:60])
    for n in (1, 7, 1000, 33333):
        pc, pv, li, vi = oracle.workqueue(s, g, lst[:60], n=n)
        np.testing.assert_array_equal(pc, full_c[:n])
        np.testing.assert_array_equal(pv, full_v[:n])
        rc, rv, li2, vi2 = oracle.workqueue(s, g, lst[:60], list_id=li, vxl_id=vi)
        np.testing.assert_array_equal(np.concatenate([pc, rc]), full_c)
        np.testing.assert_array_equal(np.concatenate([pv, rv]), full_v)
        assert (li2, vi2) == (60, 0)


def test_workqueue_tiny_full_list(oracle):
    gold = load_golden("wq_tiny.npz")
    setup, grid, _ = G.config_tiny()
    s, g = oracle_geom(oracle, setup, grid)
    c, v, _, _ = oracle.workqueue(s, g, np.arange(setup.nchannels, dtype=np.int32))
    np.testing.assert_array_equal(c, gold["cnl"])
    np.testing.assert_array_equal(v, gold["vxl"])


def test_workqueue_is_conservative(oracle):
    """Every voxel hit by a sampled ray (non-zero FP64 slab length) is a
    workqueue candidate of that channel."""
    setup, grid, nrays = G.config_tiny()
    s, g = oracle_geom(oracle, setup, grid)
    chans = np.arange(0, setup.nchannels, 37, dtype=np.int32)
    c, v, _, _ = oracle.workqueue(s, g, chans)
    cand = set((c.astype(np.int64) * grid.nvox + v).tolist())
    allc = np.repeat(chans, grid.nvox)
    allv = np.tile(np.arange(grid.nvox, dtype=np.int32), len(chans))
    val = oracle.sm_from_wq(s, g, allc, allv, nrays, f64=True)
    hit = np.nonzero(val > 0)[0]
    assert len(hit) > 0
    for k in hit:
        assert int(allc[k]) * grid.nvox + int(allv[k]) in cand


# ------------------------------- fused-kernel semantics (K2): zeros, streams
def test_keep_zeros_build_is_the_reference_workqueue(oracle):
    """With zero-valued candidates kept (getSystemMatrixDeviceOnly.cu:507-509)
    the matrix's (row, column) pairs are exactly the reference's getWorkqueue
    output (the 768-channel golden list, all 369,968 pairs), so nnz equals the
    workqueue length as in the reference; the non-zero values are those of the
    zero-dropping build."""
    setup, grid, nrays = G.config_tiny()
    s, g = oracle_geom(oracle, setup, grid)
    gold = load_golden("wq_tiny.npz")
    lst = np.arange(setup.nchannels, dtype=np.int32)
    rp, col, val, _ = oracle.build_csr(s, g, lst, nrays, flags=oracle.SM_KEEP_ZEROS)
    assert rp[-1] == len(gold["cnl"])
    np.testing.assert_array_equal(np.repeat(lst, np.diff(rp)), gold["cnl"])
    np.testing.assert_array_equal(col, gold["vxl"])
    rp0, col0, val0, _ = oracle.build_csr(s, g, lst, nrays)
    nz = val != 0
    np.testing.assert_array_equal(col[nz], col0)
    np.testing.assert_array_equal(val[nz], val0)
    assert np.count_nonzero(~nz) > 0


def test_element_rays_are_independent_streams(oracle):
    """RAND_ELEMENT: each (channel, voxel) element draws its own rays
    (Philox counter word 3 = voxel + 1); word 3 = 0 is the per-channel stream,
    and every element ray still lies in the channel's two pixel volumes."""
    setup, grid, nrays = G.config_small()
    s, _ = oracle_geom(oracle, setup, grid)
    for cnl in (0, 777, 20479):
        ida, id0z, id0y, id1z, id1y = (int(v) for v in setup.split(cnl))
        for r in range(3):
            np.testing.assert_array_equal(
                oracle.ray_element(s, cnl, -1, r, raygen=oracle.RAYGEN_RAND_ELEMENT),
                oracle.ray(s, cnl, r))
            a = oracle.ray_element(s, cnl, 5, r)
            b = oracle.ray_element(s, cnl, 6, r)
            assert not np.array_equal(a, b)
            # RAND with an element argument is the per-channel stream
            np.testing.assert_array_equal(
                oracle.ray_element(s, cnl, 5, r, raygen=oracle.RAYGEN_RAND),
                oracle.ray(s, cnl, r))
            # end points inside the pixel boxes: the inverse trafo maps them
            # into the unit cube
            for det, (iz, iy), p in ((0, (id0z, id0y), a[:3]), (1, (id1z, id1y), a[3:])):
                o = oracle.trafo(s, det, ida, iz, iy, np.zeros(3, np.float32))
                ex = oracle.trafo(s, det, ida, iz, iy, np.array([1, 0, 0], np.float32)) - o
                ey = oracle.trafo(s, det, ida, iz, iy, np.array([0, 1, 0], np.float32)) - o
                ez = oracle.trafo(s, det, ida, iz, iy, np.array([0, 0, 1], np.float32)) - o
                u = np.linalg.solve(np.stack([ex, ey, ez], 1).astype(np.float64),
                                    (p - o).astype(np.float64))
                assert np.all(u > -1e-5) and np.all(u < 1 + 1e-5)


def test_element_values_vs_fp64_and_expectation(oracle):
    """RAND_ELEMENT matrix values: within 1e-5 of the FP64 restatement on
    identical rays, equal to the slab mean of the element's own rays, and,
    averaged over many rows, the same row sums as per-channel streams (both
    estimate the mean in-grid path length; per-element sums scatter)."""
    setup, grid, nrays = G.config_small()
    s, g = oracle_geom(oracle, setup, grid)
    chans = np.arange(0, setup.nchannels, 41, dtype=np.int32)
    c, v, _, _ = oracle.workqueue(s, g, chans)
    E = oracle.RAYGEN_RAND_ELEMENT
    a32 = oracle.sm_from_wq(s, g, c, v, nrays, raygen=E)
    a64 = oracle.sm_from_wq(s, g, c, v, nrays, raygen=E, f64=True)
    assert np.all(np.abs(a32 - a64) <= 1e-5 * np.abs(a64) + 1e-5 * grid.d[0])
    for w in (0, 1234, len(c) - 1):
        box = oracle.voxel_box(g, int(v[w]))
        want = np.float32(0)
        for r in range(nrays):
            want = np.float32(want + oracle.intersection_length(
                box, oracle.ray_element(s, int(c[w]), int(v[w]), r)))
        assert a32[w] == np.float32(want / np.float32(nrays))
    a_ch = oracle.sm_from_wq(s, g, c, v, nrays)
    rows_e = np.bincount(np.searchsorted(chans, c), weights=a32)
    rows_c = np.bincount(np.searchsorted(chans, c), weights=a_ch)
    assert not np.allclose(rows_e, rows_c, rtol=1e-3)
    assert rows_e.sum() == pytest.approx(rows_c.sum(), rel=5e-3)
    # the build in element mode equals the two-stage values, zeros dropped
    lst = chans[:60]
    rp, col, val, _ = oracle.build_csr(s, g, lst, nrays, raygen=E)
    cc, vv, _, _ = oracle.workqueue(s, g, lst)
    ae = oracle.sm_from_wq(s, g, cc, vv, nrays, raygen=E)
    np.testing.assert_array_equal(col, vv[ae != 0])
    np.testing.assert_array_equal(val, ae[ae != 0])


# --------------------------------------------------- slab length (A9)
def test_intersection_length_closed_forms(oracle):
    box = np.array([0, 0, 0, 2, 2, 2], np.float32)
    # axis-parallel through the middle: the other two axes are excluded from
    # the slab test (sects false) exactly like the reference
    assert oracle.intersection_length(box, [-5, 1, 1, 5, 1, 1]) == pytest.approx(2.0)
    # body diagonal
    assert oracle.intersection_length(box, [-1, -1, -1, 3, 3, 3]) == pytest.approx(
        2 * np.sqrt(3), rel=1e-6)
    # miss
    assert oracle.intersection_length(box, [-5, 5, 1, 5, 6, 1.5]) == 0.0
    # degenerate ray
    assert oracle.intersection_length(box, [1, 1, 1, 1, 1, 1]) == 0.0
    # reference quirk: an axis with equal end points is not range-checked
    assert oracle.intersection_length(box, [-5, 7, 1, 5, 7, 1]) == pytest.approx(2.0)


def test_intersection_length_brute_force(oracle):
    """Random lines vs numeric integration along the line."""
    rng = np.random.default_rng(7)
    box = np.array([1, -2, 0.5, 3.5, 1, 2.5], np.float64)
    for _ in range(60):
        p = rng.uniform(-6, 6, 3)
        q = rng.uniform(-6, 6, 3)
        r = np.concatenate([p, q])
        got = oracle.intersection_length_f64(box, r)
        t = np.linspace(-3, 4, 400001)
        pts = p[None, :] + t[:, None] * (q - p)[None, :]
        inside = np.all((pts >= box[:3]) & (pts <= box[3:]), axis=1)
        want = inside.sum() * (t[1] - t[0]) * np.linalg.norm(q - p)
        assert got == pytest.approx(want, abs=2e-4 * np.linalg.norm(q - p) + 1e-9)
        g32 = oracle.intersection_length(box.astype(np.float32), r.astype(np.float32))
        assert g32 == pytest.approx(got, abs=1e-4)


# ------------------------------------------------- matrix from workqueue (A10)
def _grid_path_length(r, grid):
    """Length of the infinite line through r inside the grid box (FP64)."""
    lo = np.array(grid.o, np.float64)
    hi = lo + np.array(grid.d) * np.array(grid.n)
    box = np.concatenate([lo, hi])
    return box, r


def test_sm_row_sum_identity(oracle):
    """Sum over all voxels of a channel's elements = mean in-grid length of
    its rays (voxels tile the grid), FP64; the workqueue restriction loses
    nothing (conservative cull)."""
    setup, grid, nrays = G.config_tiny()
    s, g = oracle_geom(oracle, setup, grid)
    lo = np.array(grid.o, np.float64)
    box = np.concatenate([lo, lo + np.array(grid.d) * np.array(grid.n)])
    for cnl in (0, 5, 300, 767):
        c, v, _, _ = oracle.workqueue(s, g, np.array([cnl], np.int32))
        val = oracle.sm_from_wq(s, g, c, v, nrays, f64=True)
        rays = [oracle.ray(s, cnl, r).astype(np.float64) for r in range(nrays)]
        want = np.mean([oracle.intersection_length_f64(box, r) for r in rays])
        assert val.sum() == pytest.approx(want, rel=1e-9, abs=1e-9)


def test_sm_f32_vs_f64(oracle):
    """FP32 restatement vs FP64 on identical rays: <= 1e-5 relative for
    elements above 1e-3 of the row maximum, absolute 1e-5*voxel otherwise."""
    setup, grid, nrays = G.config_tiny()
    s, g = oracle_geom(oracle, setup, grid)
    c, v, _, _ = oracle.workqueue(s, g, np.arange(0, setup.nchannels, 11, dtype=np.int32))
    a32 = oracle.sm_from_wq(s, g, c, v, nrays)
    a64 = oracle.sm_from_wq(s, g, c, v, nrays, f64=True)
    err = np.abs(a32 - a64)
    assert np.all(err <= 1e-5 * np.abs(a64) + 1e-5 * grid.d[0])
    big = a64 > 1e-3 * a64.max()
    assert np.median(err[big] / a64[big]) < 5e-6


# ----------------------------------------------------- Siddon (A11)
def test_get_chords_vs_slab(oracle):
    setup, grid, nrays = G.config_tiny()
    s, g = oracle_geom(oracle, setup, grid)
    lo = np.array(grid.o, np.float64)
    gbox = np.concatenate([lo, lo + np.array(grid.d) * np.array(grid.n)])
    for cnl in (1, 200, 555):
        for r in range(nrays):
            ray = oracle.ray(s, cnl, r)
            ch, ids = oracle.get_chords(ray, g)
            total = oracle.intersection_length_f64(gbox, ray.astype(np.float64))
            assert ch.sum() == pytest.approx(total, rel=1e-4, abs=1e-4)
            for length, (ix, iy, iz) in zip(ch, ids):
                v = (ix * grid.n[1] + iy) * grid.n[2] + iz
                want = oracle.intersection_length(oracle.voxel_box(g, v), ray)
                assert length == pytest.approx(want, abs=2e-3)


# --------------------------------------------------- sparse formats (A14-A16)
def test_coo_sort_reference_cases(oracle):
    """cooSortUT.cu:68-159 (dense 2x2 shuffled), :164-184 (empty), :194-307
    (sparse density 0.7: membership and order)."""
    val, row, col = oracle.coo_sort([4, 1, 3, 2], [1, 0, 1, 0], [1, 0, 0, 1])
    np.testing.assert_array_equal(row, [0, 0, 1, 1])
    np.testing.assert_array_equal(col, [0, 1, 0, 1])
    np.testing.assert_array_equal(val, [1, 2, 3, 4])
    v, r, c = oracle.coo_sort([], [], [])
    assert len(v) == 0
    rng = np.random.default_rng(1)
    dense = rng.random((40, 30)) < 0.7
    rr, cc = np.nonzero(dense)
    perm = rng.permutation(len(rr))
    vals = rng.random(len(rr)).astype(np.float32)
    v, r, c = oracle.coo_sort(vals[perm], rr[perm], cc[perm])
    order = np.lexsort((cc, rr))
    np.testing.assert_array_equal(r, rr[order])
    np.testing.assert_array_equal(c, cc[order])
    key = dict(zip(zip(rr[perm].tolist(), cc[perm].tolist()), vals[perm].tolist()))
    assert all(key[(a, b)] == x for a, b, x in zip(r.tolist(), c.tolist(), v.tolist()))


def test_coo2csr_known_answer(oracle):
    # test_cusparseWrapper.cu:46-75
    np.testing.assert_array_equal(oracle.coo2csr([0, 1, 1, 1, 2, 2, 3, 3, 3], 4), [0, 1, 4, 6, 9])


def test_csr2ecsr_known_answer(oracle):
    # test_convertCsr2Ecsr.cu:27-46
    np.testing.assert_array_equal(oracle.csr2ecsr([2, 4, 6], [0, 0, 0, 3, 3, 3, 3, 5], 7),
                                  [0, 3, 3, 5])


# -------------------------------------------------------------- csrmv (A17)
def test_csrmv_dense_brute_force(oracle):
    rng = np.random.default_rng(2)
    D = rng.random((17, 23)) * (rng.random((17, 23)) < 0.3)
    rp = np.concatenate([[0], np.cumsum((D != 0).sum(1))]).astype(np.int32)
    rr, cc = np.nonzero(D)
    val = D[rr, cc].astype(np.float32)
    x = rng.random(23).astype(np.float32)
    y0 = rng.random(17).astype(np.float32)
    got = oracle.csrmv_f32(0, 17, 23, 1.5, val, rp, cc, x, 0.5, y0)
    np.testing.assert_allclose(got, 1.5 * D @ x + 0.5 * y0, rtol=1e-5)
    xt = rng.random(17).astype(np.float32)
    yt = rng.random(23).astype(np.float32)
    got = oracle.csrmv_f32(1, 17, 23, 2.0, val, rp, cc, xt, 1.0, yt)
    np.testing.assert_allclose(got, 2.0 * D.T @ xt + yt, rtol=1e-5)


# ---------------------------------------------------------- MLEM ops (A18-20)
def test_mlem_operations_known_answers(oracle):
    """test_mlemOperations.cu:59-114: A=(i+1)^2, B=10, C=i+1, N=10."""
    i = np.arange(10, dtype=np.float32)
    A, B, Cc = (i + 1) ** 2, np.full(10, 10, np.float32), i + 1
    np.testing.assert_array_equal(oracle.divides(A, Cc), i + 1)
    D = oracle.divides_multiplies(A, B, Cc)
    np.testing.assert_array_equal(D, 10 * (i + 1))
    assert oracle.fsum(D) == 550.0
    np.testing.assert_allclose(oracle.scales(D, np.float32(1.0 / 550.0)), (i + 1) / 55, rtol=1e-6)
    # zero divisor guard of dividesMultiplies (mlemOperations.hpp:84-86)
    np.testing.assert_array_equal(oracle.divides_multiplies([2.0], [3.0], [0.0]), [6.0])


# ----------------------------------------------------------------- MLEM (A21)
def _tiny_system(oracle, rows_all=True):
    setup, grid, nrays = G.config_tiny()
    s, g = oracle_geom(oracle, setup, grid)
    lst = np.arange(setup.nchannels, dtype=np.int32)
    rp, col, val, v64 = oracle.build_csr(s, g, lst, nrays, with_f64=True)
    return setup, grid, rp, col, v64


def test_mlem_fixed_point_and_normalisation(oracle):
    setup, grid, rp, col, v64 = _tiny_system(oracle)
    n = grid.nvox
    sens = oracle.sensitivity_f64(rp, col, v64, n)
    rng = np.random.default_rng(4)
    xt = rng.random(n) + 0.1
    xt /= xt.sum()
    y = oracle.csrmv_f64(0, len(rp) - 1, n, 1.0, v64, rp, col, xt, 0.0, np.zeros(len(rp) - 1))
    seen = sens > 0
    xt[~seen] = 0
    xt /= xt.sum()
    y = oracle.csrmv_f64(0, len(rp) - 1, n, 1.0, v64, rp, col, xt, 0.0, np.zeros(len(rp) - 1))
    x = oracle.mlem_f64(rp, col, v64, y, sens, xt, 3)
    np.testing.assert_allclose(x, xt, rtol=1e-10, atol=1e-14)
    x0 = np.full(n, 1.0 / n)
    x = oracle.mlem_f64(rp, col, v64, y, sens, x0, 4)
    assert x.sum() == pytest.approx(1.0, abs=1e-12)


def test_mlem_likelihood_monotone(oracle):
    setup, grid, rp, col, v64 = _tiny_system(oracle)
    n, m = grid.nvox, len(rp) - 1
    sens = oracle.sensitivity_f64(rp, col, v64, n)
    rng = np.random.default_rng(5)
    xt = np.zeros(n)
    c = grid.centres()
    xt[np.linalg.norm(c, axis=1) < 25] = 1.0
    ybar = oracle.csrmv_f64(0, m, n, 1.0, v64, rp, col, xt, 0.0, np.zeros(m))
    y = rng.poisson(ybar * 2e5 / ybar.sum()).astype(np.float64)
    x = np.full(n, 1.0 / n)
    prev = -np.inf
    for _ in range(6):
        x = oracle.mlem_f64(rp, col, v64, y, sens, x, 1)
        yt = oracle.csrmv_f64(0, m, n, 1.0, v64, rp, col, x, 0.0, np.zeros(m))
        # Poisson log-likelihood of the scaled model (scale-free in x)
        yt = yt * y.sum() / yt.sum()
        ll = np.sum(y[yt > 0] * np.log(yt[yt > 0])) - yt.sum()
        assert ll >= prev - 1e-9 * abs(ll)
        prev = ll


def test_mlem_f32_restatement(oracle):
    """The FP32 recurrence (the reference's precision, the CPU baseline):
    fixed point for consistent data within FP32 rounding, sum 1 after each
    iteration, and within 1e-5 of the FP64 recurrence on the same values."""
    setup, grid, rp, col, v64 = _tiny_system(oracle)
    n = grid.nvox
    v32 = v64.astype(np.float32)
    sens = oracle.sensitivity_f64(rp, col, v32.astype(np.float64), n)
    rng = np.random.default_rng(5)
    xt = rng.random(n) + 0.1
    xt[sens == 0] = 0
    xt /= xt.sum()
    y = oracle.csrmv_f64(0, len(rp) - 1, n, 1.0, v32.astype(np.float64), rp, col, xt, 0.0,
                         np.zeros(len(rp) - 1))
    x = oracle.mlem_f32(rp, col, v32, y, sens, xt, 3)
    np.testing.assert_allclose(x, xt, rtol=2e-5, atol=1e-9)
    x0 = np.full(n, 1.0 / n)
    yp = rng.poisson(y * 1e5 / y.sum()).astype(np.float64)
    x32 = oracle.mlem_f32(rp, col, v32, yp, sens, x0, 5)
    x64 = oracle.mlem_f64(rp, col, v32.astype(np.float64), yp, sens, x0, 5)
    assert float(np.sum(x32, dtype=np.float64)) == pytest.approx(1.0, abs=1e-5)
    assert np.linalg.norm(x32 - x64) / np.linalg.norm(x64) <= 1e-5
