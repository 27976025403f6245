"""Writes the golden fixtures of tests/golden from the reference's own code.

Run in a container that holds the raptr sources (builds oracle/_ref first):

    make -C oracle ref h5ref && python tests/golden/make_golden.py [full_angles | h5]

Fixtures (numpy .npz, data only):
  wq_small.npz    workqueue of the 20,480-channel / 32^3 workload computed
                  by the reference's getWorkqueue (getWorkqueue.hpp:193-243,
                  compiled from its sources): per-channel candidate counts,
                  total, SHA-256 of the (cnl, vxl) int32 arrays, the first and
                  last 4096 entries, the entries whose FP64 distance lies
                  within 1e-4 mm of the cylinder radius (near-boundary), and a
                  bounded call (n = 100,000 from cursor (3, 17)) with its
                  returned cursor.
  wq_tiny.npz     full workqueue of the 768-channel / 16^3 workload.
  trafo_ref.npz   reference pixel trafo outputs (MeasurementSetupTrafo2Cart
                  Coord.hpp, compiled) for all pixels/angles of the 4x4-head
                  setup of MeasurementSetupTrafo2CartCoordUT.cu:42-58 at 7
                  unit-cube points.
  wq_full_angles.npz  the 13x13-head, 180-angle (1 deg), 52^3 @ 2 mm
                  geometry of the full-size workload: the reference trafo of
                  both heads at all 180 angles (4 pixels x 3 unit-cube
                  points each), and the reference workqueue of one channel
                  per angle over the whole grid (pixels varied with the
                  angle), stored in full (voxel ids delta-coded per channel),
                  with the near-boundary pairs of those channels and their
                  in/out decision.
  density_ref.h5  a density file written by the reference's
                  H5DensityWriter (5 x 6 x 7 voxels, origin (-3, -1.5, 2),
                  voxel (0.5, 0.25, 1.5)); h5_ref.npz holds its input, and
                  the values the reference's H5Reader reads from the
                  measurement file raptr_amd.h5io writes for a fixed sparse
                  list on the 4x4-head setup.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as O  # noqa: E402
from raptr_amd import geometry as G  # noqa: E402


def osetup(s):
    return O.make_setup(s.pos0x, s.pos1x, s.na, s.n0z, s.n0y, s.n1z, s.n1y, s.da, s.segx,
                        s.segy, s.segz)


def ogrid(g):
    return O.make_grid(g.o, g.d, g.n)


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, np.int32).tobytes())
    return h.hexdigest()


def near_boundary(setup, grid, cnl_list, tol=1e-4):
    """(channel, voxel) pairs whose FP64 point-line distance is within tol of
    the cylinder radius (FP64 restatement of getWorkqueue.hpp:74-98)."""
    sn = np.sin(np.deg2rad(np.arange(setup.na) * setup.da))
    cs = np.cos(np.deg2rad(np.arange(setup.na) * setup.da))
    vc = grid.centres()
    r = 0.5 * (np.linalg.norm(grid.d) + np.linalg.norm([setup.segx, setup.segy, setup.segz]))
    out = []
    ida, id0z, id0y, id1z, id1y = setup.split(np.asarray(cnl_list))
    for k, c in enumerate(cnl_list):
        def centre(pos, idz, idy, n_z, n_y):
            x0 = pos
            y0 = (idy - 0.5 * (n_y - 1)) * setup.segy
            z0 = (idz - 0.5 * (n_z - 1)) * setup.segz
            a = ida[k]
            return np.array([cs[a] * x0 + sn[a] * z0, y0, -sn[a] * x0 + cs[a] * z0])
        a = centre(setup.pos0x, id0z[k], id0y[k], setup.n0z, setup.n0y)
        b = centre(setup.pos1x, id1z[k], id1y[k], setup.n1z, setup.n1y)
        ab = b - a
        d = np.linalg.norm(np.cross(ab, vc - a), axis=1) / np.linalg.norm(ab)
        for v in np.nonzero(np.abs(d - r) < tol)[0]:
            out.append((c, v))
    return np.array(out, np.int32).reshape(-1, 2)


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libraptr_ref.so missing: run `make -C oracle ref`")
    # ---- 20,480-channel workload
    setup, grid, _ = G.config_small()
    s, g = osetup(setup), ogrid(grid)
    lst = np.arange(setup.nchannels, dtype=np.int32)
    counts = np.zeros(len(lst), np.int64)
    cnl_all, vxl_all = [], []
    for c0 in range(0, len(lst), 1024):
        part = lst[c0:c0 + 1024]
        c, v, li, vi = O.ref_workqueue(s, g, part)
        assert li == len(part) and vi == 0
        cnl_all.append(c)
        vxl_all.append(v)
        counts[c0:c0 + len(part)] = np.bincount(c - c0, minlength=len(part))
    cnl = np.concatenate(cnl_all)
    vxl = np.concatenate(vxl_all)
    nb = near_boundary(setup, grid, lst)
    inq = np.zeros(len(nb), np.int8)
    if len(nb):
        key_all = cnl.astype(np.int64) * grid.nvox + vxl
        key_nb = nb[:, 0].astype(np.int64) * grid.nvox + nb[:, 1]
        inq = np.isin(key_nb, key_all).astype(np.int8)
    bc, bv, bli, bvi = O.ref_workqueue(s, g, lst, list_id=3, vxl_id=17, n=100000)
    np.savez_compressed(os.path.join(HERE, "wq_small.npz"), counts=counts.astype(np.int32),
                        total=np.int64(len(cnl)), sha256=np.array(sha(cnl, vxl)),
                        head_cnl=cnl[:4096], head_vxl=vxl[:4096], tail_cnl=cnl[-4096:],
                        tail_vxl=vxl[-4096:], near=nb, near_in=inq,
                        bounded_sha256=np.array(sha(bc, bv)), bounded_n=np.int64(len(bc)),
                        bounded_cursor=np.array([bli, bvi], np.int32))
    print("wq_small:", len(cnl), "entries,", len(nb), "near-boundary pairs,", inq.sum(), "inside")
    # ---- 768-channel workload (full list)
    setup, grid, _ = G.config_tiny()
    s, g = osetup(setup), ogrid(grid)
    lst = np.arange(setup.nchannels, dtype=np.int32)
    c, v, li, vi = O.ref_workqueue(s, g, lst)
    np.savez_compressed(os.path.join(HERE, "wq_tiny.npz"), cnl=c, vxl=v)
    print("wq_tiny:", len(c), "entries")
    # ---- trafo of the unit-test setup (MeasurementSetupTrafo2CartCoordUT.cu:42-58, in mm)
    ts = O.make_setup(-35.0, 35.0, 6, 4, 4, 4, 4, 60.0, 10.0, 10.0, 10.0)
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [.1, .1, .1], [.5, .5, .5],
                    [.9, .3, .7]], np.float32)
    sf = np.array([ts.pos0x, ts.pos1x, ts.da, ts.segx, ts.segy, ts.segz], np.float32)
    si = np.array([ts.na, ts.n0z, ts.n0y, ts.n1z, ts.n1y], np.int32)
    R = O.ref()
    out = np.zeros((2, 6, 4, 4, len(pts), 3), np.float32)
    for det in range(2):
        for a in range(6):
            for iz in range(4):
                for iy in range(4):
                    for k, p in enumerate(pts):
                        o = np.zeros(3, np.float32)
                        R.ref_trafo(O._p(sf, O.C.c_float), O._p(si, O.C.c_int), det, iz, iy, a,
                                    O._p(p, O.C.c_float), O._p(o, O.C.c_float))
                        out[det, a, iz, iy, k] = o
    np.savez_compressed(os.path.join(HERE, "trafo_ref.npz"), setup_f=sf, setup_i=si, pts=pts,
                        out=out)
    print("trafo_ref written")
    full_angles()
    h5_fixtures()


def h5_fixtures():
    """Reference-written density file, and the reference reader's view of a
    measurement file written by raptr_amd.h5io (needs oracle/_ref/
    libraptr_h5ref.so: `make -C oracle h5ref`)."""
    import tempfile
    from raptr_amd import h5io
    if not O.h5ref_available():
        raise SystemExit("oracle/_ref/libraptr_h5ref.so missing: run `make -C oracle h5ref`")
    nv = np.array([5, 6, 7], np.int32)
    org = np.array([-3.0, -1.5, 2.0], np.float32)
    vox = np.array([0.5, 0.25, 1.5], np.float32)
    dens = (np.arange(nv.prod(), dtype=np.float32) * np.float32(0.37) - 11).astype(np.float32)
    path = os.path.join(HERE, "density_ref.h5")
    assert O.h5ref().ref_h5_write_density(path.encode(), O._p(dens, O.C.c_float),
                                          O._p(org, O.C.c_float), O._p(vox, O.C.c_float),
                                          O._p(nv, O.C.c_int)) == 0
    setup = G.config_tiny()[0]
    rng = np.random.default_rng(9)
    ch = np.sort(rng.choice(setup.nchannels, 200, replace=False)).astype(np.int32)
    counts = rng.integers(1, 50, len(ch)).astype(np.float32)
    with tempfile.TemporaryDirectory() as d:
        mfn = os.path.join(d, "meas.h5")
        h5io.write_measurement(mfn, setup, ch, counts)
        read = O.ref_h5_read(mfn, False, setup.nchannels)
    np.savez_compressed(os.path.join(HERE, "h5_ref.npz"), density=dens, origin=org, voxel=vox,
                        nvox=nv, meas_channels=ch, meas_counts=counts, meas_ref_read=read)
    print("density_ref.h5 and h5_ref.npz written")


def full_angle_channels(setup):
    """One channel per angle of the 180-angle setup, pixels varied with the
    angle so that every head row and column occurs."""
    a = np.arange(setup.na)
    n = setup.n0z
    return np.asarray(setup.lin_id(a % n, (5 * a + 3) % n, (7 * a + 1) % n, (11 * a + 6) % n, a),
                      np.int32)


def full_angles():
    setup, grid, _ = G.config_full()
    s, g = osetup(setup), ogrid(grid)
    lst = full_angle_channels(setup)
    c, v, li, vi = O.ref_workqueue(s, g, lst)
    assert li == len(lst) and vi == 0
    counts = np.bincount(np.searchsorted(lst, c) if np.all(np.diff(lst) > 0) else
                         np.array([np.nonzero(lst == k)[0][0] for k in c]), minlength=len(lst))
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    dv = np.diff(v, prepend=0).astype(np.int32)
    dv[starts[counts > 0]] = v[starts[counts > 0]]
    nb = near_boundary(setup, grid, lst)
    key_all = c.astype(np.int64) * grid.nvox + v
    key_nb = nb[:, 0].astype(np.int64) * grid.nvox + nb[:, 1]
    inq = np.isin(key_nb, key_all).astype(np.int8)
    pix = np.array([[0, 0], [12, 12], [6, 6], [3, 9]], np.int32)
    pts = np.array([[.5, .5, .5], [0, 0, 0], [.9, .3, .7]], np.float32)
    sf = np.array([setup.pos0x, setup.pos1x, setup.da, setup.segx, setup.segy, setup.segz],
                  np.float32)
    si = np.array([setup.na, setup.n0z, setup.n0y, setup.n1z, setup.n1y], np.int32)
    R = O.ref()
    out = np.zeros((2, setup.na, len(pix), len(pts), 3), np.float32)
    for det in range(2):
        for a in range(setup.na):
            for j, (iz, iy) in enumerate(pix):
                for k, p in enumerate(pts):
                    o = np.zeros(3, np.float32)
                    R.ref_trafo(O._p(sf, O.C.c_float), O._p(si, O.C.c_int), det, int(iz),
                                int(iy), a, O._p(p, O.C.c_float), O._p(o, O.C.c_float))
                    out[det, a, j, k] = o
    np.savez_compressed(os.path.join(HERE, "wq_full_angles.npz"), channels=lst,
                        counts=counts.astype(np.int32), vxl_delta=dv,
                        sha256=np.array(sha(c, v)), near=nb, near_in=inq, setup_f=sf,
                        setup_i=si, trafo_pix=pix, trafo_pts=pts, trafo_out=out)
    print("wq_full_angles:", len(c), "entries,", len(nb), "near-boundary pairs,", inq.sum(),
          "inside")


if __name__ == "__main__":
    if sys.argv[1:] == ["full_angles"]:
        full_angles()
    elif sys.argv[1:] == ["h5"]:
        h5_fixtures()
    else:
        main()
