"""Fixed-shape pass loop on the full-size workload, for profilers.

    python -m raptr_amd.passes [--frames F] [--value-type f32] [--block-rows R]
        [--order y|slope|angle|angle_slope] [--xcd-blocks B] [--active A] [--reps N]
        [--separate]

Builds the benchmark's full-size workload (13x13-pixel heads, 180 angles,
52^3 grid at 2 mm, 30 rays, six-ellipsoid phantom, 2e7 Poisson events), the
matrix of its measured channels and the FP64 sensitivity, then one plan
with the given shape (no autotuning), and runs `reps` forward and backward
passes; a single-frame plan runs `reps` fused iterations (its default)
unless --separate is given.  A 2 GiB streaming copy (k_stream_copy) runs first: the known-size
dispatch that calibrates the HBM traffic counters.  Progress goes to
stderr.  Nothing is timed here; rocprofv3 measures the kernels.
"""
import argparse
import sys

import numpy as np

from . import _capi as capi
from . import geometry as G
from . import synthetic, sysmat
from .device import DeviceArray, Stream
from .recon import Reconstructor, default_order_kind


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m raptr_amd.passes")
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--value-type", default="f32", choices=("f32", "f16", "bf16", "u16"))
    ap.add_argument("--block-rows", type=int, default=0, help="0: the plan's default")
    ap.add_argument("--order", default=None, choices=("y", "slope", "angle", "angle_slope"),
                    help="locality row order (default: the plan's, recon.default_order_kind)")
    ap.add_argument("--xcd-blocks", type=int, default=-1,
                    help="64 frames: row blocks per XCD and launch (-1: the plan's default)")
    ap.add_argument("--active", type=int, default=64, choices=(16, 32, 64),
                    help="64 frames: frames each pass computes (an MLEM iteration per rep)")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--separate", action="store_true",
                    help="one frame: the separate forward and backward passes instead of the "
                         "fused iteration")
    a = ap.parse_args(argv)
    st_obj = Stream()
    st = st_obj.handle
    nb = 2 << 30
    ca, cb = DeviceArray(nb // 4, np.float32), DeviceArray(nb // 4, np.float32)
    capi.call("raptr_memset", ca.ptr, 0, nb, st)
    capi.call("raptr_stream_copy", cb.ptr, ca.ptr, nb // 4, st)
    st_obj.sync()
    del ca, cb
    setup, grid, nrays = G.config_full()
    image, _ = synthetic.random_ellipsoids(grid, n=6, seed=2)
    meas, y, _ = synthetic.measurement(setup, grid, nrays, image, 2e7, seed=1)
    m, n, F = len(meas), grid.nvox, a.frames
    d_list = DeviceArray.from_numpy(meas)
    A = sysmat.build_csr(setup, grid, meas, nrays, d_list=d_list, stream=st, capacity="auto")
    sens = sysmat.sensitivity(setup, grid, nrays, stream=st)
    kind = a.order or default_order_kind(F)
    order = sysmat.locality_order(setup, d_list, stream=st, kind=kind)
    R = Reconstructor(A, sens, a.value_type, F, block_rows=a.block_rows, stream=st,
                      row_order=order, grid=grid if F == 64 else None, group=16)
    if F == 64 and a.xcd_blocks >= 0:
        R.set_xcd_blocks(a.xcd_blocks)
    log(f"plan: {R.info()} (order {kind})")
    Y = np.repeat(y[:, None], F, axis=1).astype(np.float32)
    dy = DeviceArray.from_numpy(Y.reshape(-1))
    dx = DeviceArray.from_numpy(np.full(n * F, 1.0 / n, np.float32))
    de = DeviceArray(m * F, np.float32)
    fused = F == 1 and not a.separate and R.fused()["on"]
    if F == 1 and a.separate and R.fused()["available"]:
        R.set_fused(False)
    for r in range(a.reps):
        if fused:
            R.run(dy, dx, 1, 1, use_graph=False, stream=st)
        elif a.active < 64:
            R.run(dy, dx, 1, F, use_graph=False, stream=st, active=a.active)
        else:
            capi.call("raptr_recon_forward", R.handle, F, dx.ptr, dy.ptr, de.ptr, st)
            capi.call("raptr_recon_backward", R.handle, F, dx.ptr, st)
        st_obj.sync()
        log(f"passes {r + 1}/{a.reps}")
    R.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
