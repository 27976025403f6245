"""Drivers of the reference's executables over the device path, reading and
writing its HDF5 files.

* ``reco``: list-mode MLEM from a measurement file (tests/reco.cu:53-330):
  measured channels and counts from "messung", sensitivity from a density
  file when one is given (reco.cu:171-204) or computed on the device over all
  channels (FP64), start image from a density file or uniform 1/N
  (reco.cu:135-158), one density file per iteration named
  "<iteration>_<output>" (reco.cu:304-311) or only the last one.
* ``backprojection``: summed back-projection A^T y normalised to sum 1
  (tests/backprojection.cu:144-185).
* ``sm`` / ``system_matrix``: the matrix of the measured channels only,
  with the reference's per-chunk element counts
  (tests/pureSMCalculation.cu:53-176).

The geometry is an argument (the reference compiles it in from
real_measurementsetup_defines.h / voxelgrid_defines.h); ``--config`` names
the workloads of raptr_amd.geometry on the command line.
"""
import argparse
import os
import sys

import numpy as np

from . import geometry as G
from . import h5io, sysmat
from .device import DeviceArray, Stream
from .recon import Reconstructor

CONFIGS = {"tiny": G.config_tiny, "small": G.config_small, "full": G.config_full}


def system_matrix(measurement, setup, grid, nrays, seed=1, stream=None, keep_zeros=False,
                  raygen=sysmat.capi.RAYGEN_RAND):
    """The matrix of the measured channels (rows in list order).  keep_zeros:
    every culled candidate stored, as the reference's fused kernel does (nnz
    = workqueue length); raygen RAYGEN_RAND_ELEMENT: rays per element."""
    channels, _ = h5io.read_measurement(measurement, setup)
    return channels, sysmat.build_csr(setup, grid, channels, nrays, seed=seed, stream=stream,
                                      capacity="auto", keep_zeros=keep_zeros, raygen=raygen)


def chunk_counts(row_ptr, eff_m, lim_m):
    """Entries per chunk of lim_m measured channels (wrappers.hpp:429-464
    chunk math over the measurement list), from a row pointer."""
    n_chunks = (eff_m + lim_m - 1) // lim_m
    return [int(row_ptr[min((k + 1) * lim_m, eff_m)] - row_ptr[k * lim_m])
            for k in range(n_chunks)]


def backprojection(measurement, output, setup, grid, nrays, seed=1):
    channels, counts = h5io.read_measurement(measurement, setup)
    x = sysmat.backproject(setup, grid, nrays, channels, counts, seed=seed)
    h5io.write_density(output, x.astype(np.float32), grid)
    return x


def reco(measurement, output, setup, grid, nrays, niter, sensitivity=None, guess=None,
         every_iteration=True, value_type="f32", seed=1, log=None):
    """Returns the final image (float32, grid order)."""
    st = Stream()
    channels, counts = h5io.read_measurement(measurement, setup)
    A = sysmat.build_csr(setup, grid, channels, nrays, seed=seed, stream=st.handle,
                         capacity="auto")
    if sensitivity and os.path.exists(sensitivity):
        s, sgrid = h5io.read_density(sensitivity)
        if s.size != grid.nvox:
            raise ValueError(f"sensitivity {sensitivity} has {s.size} voxels, grid {grid.nvox}")
        sens = DeviceArray.from_numpy(s.reshape(-1).astype(np.float64), stream=st.handle)
    else:
        sens = sysmat.sensitivity(setup, grid, nrays, seed=seed, stream=st.handle)
    if guess and os.path.exists(guess):
        x0, _ = h5io.read_density(guess)
        x0 = x0.reshape(-1).astype(np.float32)
        if x0.size != grid.nvox:
            raise ValueError(f"guess {guess} has {x0.size} voxels, grid {grid.nvox}")
    else:
        x0 = np.full(grid.nvox, 1.0 / grid.nvox, np.float32)
    # plan rows in locality order (the single-frame forward then stages x in
    # LDS); the images do not depend on it beyond FP32 summation order
    R = Reconstructor(A, sens, value_type, 1, stream=st.handle, row_order=setup)
    dy = DeviceArray.from_numpy(counts.astype(np.float32), stream=st.handle)
    dx = DeviceArray.from_numpy(x0, stream=st.handle)
    head, tail = os.path.split(output)
    steps = [1] * niter if every_iteration else [niter]
    done = 0
    for k in steps:
        R.run(dy, dx, k, 1, use_graph=k > 1, stream=st.handle)
        done += k
        x = dx.numpy(stream=st.handle)
        if every_iteration or done == niter:
            name = os.path.join(head, f"{done - 1}_{tail}") if every_iteration else output
            h5io.write_density(name, x, grid)
            if log:
                log(f"iteration {done}: sum {float(x.astype(np.float64).sum()):.6f} -> {name}")
    R.close()
    return x


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m raptr_amd.drivers")
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("reco", help="MLEM reconstruction (reco.cu)")
    r.add_argument("measurement")
    r.add_argument("output")
    r.add_argument("nrays", type=int)
    r.add_argument("sensitivity", help="density file, or a missing path to compute it")
    r.add_argument("niter", type=int)
    r.add_argument("guess", help="density file, or a missing path for a uniform start")
    r.add_argument("--last-only", action="store_true", help="write only the final image")
    b = sub.add_parser("backprojection", help="summed back-projection (backprojection.cu)")
    b.add_argument("measurement")
    b.add_argument("output")
    b.add_argument("nrays", type=int)
    m = sub.add_parser("sm", help="system matrix only (pureSMCalculation.cu)")
    m.add_argument("measurement")
    m.add_argument("nrays", type=int)
    m.add_argument("--drop-zeros", action="store_true",
                   help="store exact non-zeros only (the reconstruction plans' form); "
                        "default: every culled candidate, as the reference")
    m.add_argument("--element-rays", action="store_true",
                   help="rays per (channel, voxel) element (the reference's per-thread "
                        "generators) instead of per channel")
    m.add_argument("--lim-m", type=int, default=0,
                   help="channels per reported chunk (default: the reference's LIMM, "
                        "512 MiB of float values / voxels)")
    m.add_argument("--output", help="write the matrix as .npz (row_ptr, col, val, channels)")
    for p in (r, b, m):
        p.add_argument("--config", choices=sorted(CONFIGS), default="full")
    a = ap.parse_args(argv)
    setup, grid, _ = CONFIGS[a.config]()
    if a.cmd == "reco":
        reco(a.measurement, a.output, setup, grid, a.nrays, a.niter, a.sensitivity, a.guess,
             every_iteration=not a.last_only, log=lambda s: print(s, flush=True))
    elif a.cmd == "sm":
        import time
        t = time.perf_counter()
        raygen = sysmat.capi.RAYGEN_RAND_ELEMENT if a.element_rays else sysmat.capi.RAYGEN_RAND
        channels, A = system_matrix(a.measurement, setup, grid, a.nrays,
                                    keep_zeros=not a.drop_zeros, raygen=raygen)
        secs = time.perf_counter() - t
        rp = A.row_ptr.numpy(A.m + 1)
        lim = a.lim_m or max(1, (512 * 1024 * 1024 // 4) // grid.nvox)
        counts = chunk_counts(rp, len(channels), lim)
        print(f"Calculate system matrix with #LOR = {len(channels)} and #voxels = {grid.nvox}. "
              f"Partition into chunks of maximum #LOR = {lim}: {len(counts)} chunks.")
        for k, c in enumerate(counts):
            print(f"Finished chunk {k} of {len(counts)}, found {c} elements.")
        print(f"Found: {A.nnz} elements.")
        print(f"Time for SM calculation: {secs:.3f} s")
        if a.output:
            rp, col, val = A.to_host()
            np.savez(a.output, row_ptr=rp, col=col, val=val, channels=channels)
    else:
        x = backprojection(a.measurement, a.output, setup, grid, a.nrays)
        print(f"Norm: {float(x.sum()):.6f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
