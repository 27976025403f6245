"""GPU parity at the full-size workload: 13x13-pixel heads, 180 angles at
1 deg (5,140,980 channels, 4,510,820 measured), 52^3 grid at 2 mm, 30 rays,
six-ellipsoid phantom, 2e7 Poisson events (the benchmark's workload).

The FP64 host recurrence (the oracle's MLEM, reco.cu:230-303 in double
precision) runs on the same matrix values as the device (the FP32 values of
the GPU build, which equal the oracle's build bit for bit on the smaller
workloads), with the GPU's FP64 sensitivity.  Bounds:
  * plain FP32 (one frame), 50 iterations: relative L2 <= 1e-4;
  * batched FP32, 64 frames on the matrix cores, 50 iterations: frames 0
    and 63 within 1e-4 of the oracle, every frame within 2e-5 of the plain
    single-frame run of the same counts;
  * 16-bit values (UNORM16, FP16, BF16 with per-row FP32 scale), 30
    iterations, two frames: every frame within 1e-5 of the FP64 recurrence
    on the dequantized matrix (the kernels' arithmetic), and within the
    stated bounds of the FP32-matrix oracle (the storage's own error).
And the 96^3 serving geometry as a single-frame plan (fused and separate
passes, 20 iterations, <= 1e-4).
"""
import os

import numpy as np
import pytest

from raptr_amd import geometry as G

pytestmark = pytest.mark.gpu

ITERS = 50
ITERS_16 = 30
# relative L2 / max-voxel bounds of the 16-bit storage at 30 iterations,
# per frame (measured on the 2e7- / 2e6-event frames: UNORM16 1.5e-3 / 8.0e-3
# rel L2, 1.1e-2 / 7.7e-2 max voxel; FP16 1.3e-3 / 3.0e-3, 5.4e-3 / 7.1e-3;
# BF16 1.1e-2 / 1.6e-2, 4.9e-2 / 4.6e-2); against the FP64 recurrence on
# the dequantized matrix every type is within 8e-7
BOUNDS_16 = {"u16": (1.5e-2, 1.5e-1), "f16": (6e-3, 2e-2), "bf16": (3e-2, 1e-1)}


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _maxerr(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def full(gpu, oracle):
    from raptr_amd import synthetic, sysmat
    from raptr_amd.device import DeviceArray
    setup, grid, nrays = G.config_full()
    print("full-size workload: data", flush=True)
    image, _ = synthetic.random_ellipsoids(grid, n=6, seed=2)
    meas, y, ybar = synthetic.measurement(setup, grid, nrays, image, 2e7, seed=1)
    d_list = DeviceArray.from_numpy(meas)
    A = sysmat.build_csr(setup, grid, meas, nrays, d_list=d_list, capacity="auto")
    sens = sysmat.sensitivity(setup, grid, nrays)
    order = sysmat.locality_order(setup, d_list)
    print(f"full-size workload: m={len(meas)} nnz={A.nnz}", flush=True)
    rp, col, val = A.to_host()
    oracle.lib().ro_set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    ym = ybar[meas]
    totals = np.concatenate([[float(y.sum())], np.geomspace(2e6, 2e7, 63)])
    Y = np.empty((len(meas), 64), np.float32)
    Y[:, 0] = y
    for f in range(1, 64):
        Y[:, f] = synthetic.poisson_counts(ym, totals[f], seed=100 + f)
    return dict(setup=setup, grid=grid, nrays=nrays, meas=meas, A=A, sens=sens, order=order,
                rp=rp, col=col, val=val, s64=sens.numpy(), Y=Y, cache={})


def _oracle(oracle, F, frame, niter, val64=None, key=None):
    """FP64 recurrence of one frame (cached per key)."""
    k = (key, frame, niter)
    if k not in F["cache"]:
        n = F["grid"].nvox
        v = val64 if val64 is not None else F["val"].astype(np.float64)
        F["cache"][k] = oracle.mlem_f64(F["rp"], F["col"], v, F["Y"][:, frame].astype(np.float64),
                                        F["s64"], np.full(n, 1.0 / n), niter)
    return F["cache"][k]


def test_full_plain_mlem_vs_fp64(full, oracle):
    """Plain single-frame FP32 plan with the separate passes (locality row
    order, LDS-staged forward, per-XCD backward, 16-bit indices) over 50
    iterations."""
    from raptr_amd.recon import Reconstructor
    F = full
    n = F["grid"].nvox
    R = Reconstructor(F["A"], F["sens"], "f32", 1, row_order=F["order"])
    R.set_fused(False)  # the separate passes; the fused iteration is tested below
    x = R.mlem(F["Y"][:, 0].copy(), np.full(n, 1.0 / n, np.float32), ITERS)
    R.close()
    x64 = _oracle(oracle, F, 0, ITERS)
    rel, mx = _rel(x, x64), _maxerr(x, x64)
    print(f"plain f32: rel L2 {rel:.2e}, max voxel {mx:.2e}")
    assert rel <= 1e-4
    assert abs(float(np.sum(x, dtype=np.float64)) - 1.0) <= 1e-5
    F["cache"]["plain_x"] = x


def test_full_plain_fused_vs_fp64(full, oracle):
    """The fused single-frame iteration (one pass over the matrix per
    iteration) over 50 iterations: within 1e-4 of the FP64 oracle, within
    2e-6 of the separate passes, and the same bits on a second run."""
    from raptr_amd.recon import Reconstructor
    F = full
    n = F["grid"].nvox
    from raptr_amd import sysmat
    from raptr_amd.device import DeviceArray
    # the single-frame default row order (y-major), as bench.py's plain plan
    order_y = sysmat.locality_order(F["setup"], DeviceArray.from_numpy(F["meas"]), kind="y")
    R = Reconstructor(F["A"], F["sens"], "f32", 1, row_order=order_y)
    assert R.fused()["available"] == 1 and R.fused()["on"] == 2  # the default
    # the plan's choices at this size: groups of 256 rows, 8 waves, the
    # column-tile reduce (tiles of 1,024 columns, parts of 512 segments)
    assert R.unions()["rows"] == 256
    assert R.fused()["waves"] == 8 and R.fused()["per_xcd"] == 2
    assert R.fused_tiles()["columns"] == 1024
    y, x0 = F["Y"][:, 0].copy(), np.full(n, 1.0 / n, np.float32)
    x = R.mlem(y, x0, ITERS)
    np.testing.assert_array_equal(x, R.mlem(y, x0, ITERS))
    R.close()
    x64 = _oracle(oracle, F, 0, ITERS)
    rel, mx = _rel(x, x64), _maxerr(x, x64)
    print(f"plain fused f32: rel L2 {rel:.2e}, max voxel {mx:.2e}")
    assert rel <= 1e-4
    assert abs(float(np.sum(x, dtype=np.float64)) - 1.0) <= 1e-5
    if "plain_x" in F["cache"]:
        assert _rel(x, F["cache"]["plain_x"].astype(np.float64)) <= 2e-6


def test_full_batched64_vs_fp64_and_plain(full, oracle):
    """The benchmark's batched mode: 64 frames (totals 2e6-2e7) in one plan
    with the grouped layouts on the matrix cores, 50 iterations; frames 0
    and 63 against the oracle, every frame against the plain plan."""
    from raptr_amd.recon import Reconstructor
    F = full
    n = F["grid"].nvox
    R = Reconstructor(F["A"], F["sens"], "f32", 64, block_rows=32768, row_order=F["order"],
                      grid=F["grid"], group=16)
    assert R.grouped()["on"] == 1
    X = R.mlem(F["Y"], np.full((n, 64), 1.0 / n, np.float32), ITERS)
    R.close()
    for f in (0, 63):
        x64 = _oracle(oracle, F, f, ITERS)
        rel = _rel(X[:, f], x64)
        print(f"batched frame {f}: rel L2 {rel:.2e}, max voxel {_maxerr(X[:, f], x64):.2e}")
        assert rel <= 1e-4
    P = Reconstructor(F["A"], F["sens"], "f32", 1, row_order=F["order"])
    worst = 0.0
    for f in range(64):
        x1 = P.mlem(F["Y"][:, f].copy(), np.full(n, 1.0 / n, np.float32), ITERS)
        worst = max(worst, _rel(X[:, f], x1.astype(np.float64)))
    P.close()
    print(f"batched vs plain, worst frame: {worst:.2e}")
    assert worst <= 2e-5


def _dequantized(val, rp, vt):
    """FP64 values of the plan's 16-bit storage: per-row scale (row maximum;
    /65535 for UNORM16) times the stored 16-bit value of value/scale, with
    the rounding of the plan build (round to nearest even)."""
    m = len(rp) - 1
    cnt = np.diff(rp)
    mx = np.zeros(m, np.float32)
    nz = cnt > 0
    mx[nz] = np.maximum.reduceat(np.abs(val), rp[:-1][nz])
    sc = np.where(mx > 0, mx, np.float32(1.0)).astype(np.float32)
    scr = np.repeat(sc, cnt)
    q = (val / scr).astype(np.float32)
    if vt == "f16":
        d = q.astype(np.float16).astype(np.float64)
        eff = sc.astype(np.float64)
    elif vt == "bf16":
        u = q.view(np.uint32).astype(np.uint64)
        u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
        d = u.astype(np.uint32).view(np.float32).astype(np.float64)
        eff = sc.astype(np.float64)
    else:
        d = np.minimum(np.rint(q * np.float32(65535.0)), 65535.0).astype(np.float64)
        eff = (sc * np.float32(1.0 / 65535.0)).astype(np.float64)
    return d * np.repeat(eff, cnt)


@pytest.mark.parametrize("vt", ["u16", "f16", "bf16"])
def test_full_mixed_precision_per_frame(full, oracle, vt):
    """16-bit matrix values on the full-size workload, two frames (the 2e7-
    and the 2e6-event data sets), 30 iterations: per-frame relative L2 and
    max-voxel error against the FP64 oracle on the FP32 matrix (the cost of
    the storage) and against the FP64 recurrence on the dequantized matrix
    (the kernels' FP32 arithmetic, <= 1e-5)."""
    from raptr_amd.recon import Reconstructor
    F = full
    n = F["grid"].nvox
    R = Reconstructor(F["A"], F["sens"], vt, 2, row_order=F["order"])
    X = R.mlem(F["Y"][:, :2].copy(), np.full((n, 2), 1.0 / n, np.float32), ITERS_16)
    R.close()
    vq = _dequantized(F["val"], F["rp"], vt)
    for f in range(2):
        xq = _oracle(oracle, F, f, ITERS_16, val64=vq, key=vt)
        x64 = _oracle(oracle, F, f, ITERS_16)
        rel_q, rel, mx = _rel(X[:, f], xq), _rel(X[:, f], x64), _maxerr(X[:, f], x64)
        print(f"{vt} frame {f}: vs dequantized FP64 {rel_q:.2e}; vs FP64 rel L2 {rel:.2e}, "
              f"max voxel {mx:.2e}")
        assert rel_q <= 1e-5
        assert rel <= BOUNDS_16[vt][0] and mx <= BOUNDS_16[vt][1]


def test_full_96_plain_fused_vs_fp64(gpu, oracle):
    """The largest serving geometry as a single-frame plan: 16x16-pixel
    heads, 60 angles at 3 deg (3,932,160 channels), 96^3 grid at 2 mm, 20
    rays, 5e6 events; rows of about 640 entries.  The fused iteration (its
    default shape here: 32-row union groups, 3 register-held steps, 8
    waves per group, the column-tile reduce) and the
    separate passes, 20 iterations, against the FP64 oracle on the same
    matrix (<= 1e-4) and each other (<= 2e-6); the fused run repeats bit for
    bit."""
    from raptr_amd import synthetic, sysmat
    from raptr_amd.device import DeviceArray
    from raptr_amd.recon import Reconstructor
    setup, grid = G.two_head_setup(16, 60, 3.0), G.centred_grid(96, 2.0)
    image, _ = synthetic.random_ellipsoids(grid, n=6, seed=2)
    meas, y, _ = synthetic.measurement(setup, grid, 20, image, 5e6, seed=1)
    d_list = DeviceArray.from_numpy(meas)
    A = sysmat.build_csr(setup, grid, meas, 20, d_list=d_list, capacity="auto")
    sens = sysmat.sensitivity(setup, grid, 20)
    order = sysmat.locality_order(setup, d_list, kind="y")  # the single-frame default
    n, it = grid.nvox, 20
    print(f"96^3 workload: m={len(meas)} nnz={A.nnz}", flush=True)
    R = Reconstructor(A, sens, "f32", 1, row_order=order)
    fz = R.fused()
    assert fz["on"] == 2 and fz["steps"] == 3 and fz["waves"] == 8 and fz["per_xcd"] == 2
    assert R.unions()["rows"] == 32
    y32, x0 = y.astype(np.float32), np.full(n, 1.0 / n, np.float32)
    xf = R.mlem(y32, x0, it)
    np.testing.assert_array_equal(xf, R.mlem(y32, x0, it))
    R.set_fused(False)
    xs = R.mlem(y32, x0, it)
    R.close()
    rp, col, val = A.to_host()
    oracle.lib().ro_set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    x64 = oracle.mlem_f64(rp, col, val.astype(np.float64), y.astype(np.float64), sens.numpy(),
                          np.full(n, 1.0 / n), it)
    rel_f, rel_s = _rel(xf, x64), _rel(xs, x64)
    print(f"96^3: fused rel L2 {rel_f:.2e}, separate {rel_s:.2e}")
    assert rel_f <= 1e-4 and rel_s <= 1e-4
    assert _rel(xf, xs.astype(np.float64)) <= 2e-6
