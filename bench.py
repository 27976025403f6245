"""Benchmark of the raptr reconstruction hot path on one MI355X.

Workload (default): plain MLEM at full size -- 13x13-pixel heads (4x4x20 mm,
460 mm apart), 180 angles at 1 deg (5,140,980 channels), 52^3 grid at 2 mm,
30 Monte Carlo rays per channel, phantom of 6 random ellipsoids (semi-axes
5-25 mm, activity log-uniform 0.5-8) on a 0.1 background, 2e7 Poisson events,
FP32 CSR, 50 MLEM iterations per frame.

A step is one frame: 50 MLEM iterations (forward+ratio, backward+update,
FP64 normalisation) over the device-resident matrix of the measured
channels, replayed from a HIP graph.  The matrix, its transpose and the FP64
sensitivity image are built once before timing (their times are reported in
`build`).  `value` is frames/s.

Prints one JSON line on stdout; progress on stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iters", type=int, default=50, help="MLEM iterations per frame")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=20)
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU "
              "row-sharded design is described in DESIGN.md and not run here.", file=sys.stderr)
        sys.exit(2)

    import ctypes as C

    from raptr_amd import _capi as capi
    from raptr_amd import geometry as G
    from raptr_amd import sparse, synthetic, sysmat
    from raptr_amd.device import DeviceArray, Event, Stream, Workspace, device_name

    dev = device_name()
    log("device:", dev)
    setup, grid, nrays = G.config_full()
    niter = args.iters
    total_events = 2e7

    t0 = time.perf_counter()
    image, _ = synthetic.random_ellipsoids(grid, n=6, seed=2)
    meas, y, ybar = synthetic.measurement(setup, grid, nrays, image, total_events, seed=1)
    t_data = time.perf_counter() - t0
    m, n = len(meas), grid.nvox
    log(f"data: {m} measured channels of {setup.nchannels} ({t_data:.1f}s)")

    stream = Stream()
    st = stream.handle
    ev0, ev1 = Event(), Event()

    def timed(fn):
        ev0.record(st)
        fn()
        ev1.record(st)
        return ev0.elapsed_ms(ev1)

    d_list = DeviceArray.from_numpy(meas)
    res = {}
    ms_count = timed(lambda: res.setdefault("A0", sysmat.build_csr(
        setup, grid, meas, nrays, d_list=d_list, stream=st, capacity=None)))
    A = res["A0"]
    log(f"matrix: nnz={A.nnz} ({A.nnz / m:.1f}/row), count+fill {ms_count:.0f} ms")
    ms_fill = timed(lambda: res.__setitem__("A", sysmat.build_csr(
        setup, grid, meas, nrays, d_list=d_list, stream=st, capacity=A.nnz)))
    A = res["A"]
    ms_t = timed(lambda: res.__setitem__("At", sparse.transpose(A, stream=st)))
    At = res["At"]
    ms_s = timed(lambda: res.__setitem__("sens", sysmat.sensitivity(setup, grid, nrays,
                                                                     stream=st)))
    sens = res["sens"]
    log(f"build: fill {ms_fill:.0f} ms, transpose {ms_t:.0f} ms, sensitivity {ms_s:.0f} ms")

    d_y = DeviceArray.from_numpy(y)
    x0 = np.full(n, 1.0 / n, np.float32)
    d_x = DeviceArray.from_numpy(x0)
    L = capi.lib()
    ws = Workspace(L.raptr_mlem_workspace_size(m, n))
    from raptr_amd import mlem as mlem_mod
    cfg, tune_ms = mlem_mod.autotune(A, At, d_y, d_x, reps=3, stream=st)
    log(f"autotune: forward cfg {cfg[0]} ({tune_ms[0]:.3f} ms), backward cfg {cfg[1]} "
        f"({tune_ms[1]:.3f} ms)")

    def frame():
        d_x.copy_from(x0, stream=st)
        capi.call("raptr_mlem", m, n, A.row_ptr.ptr, A.col.ptr, A.val.ptr, At.row_ptr.ptr,
                  At.col.ptr, At.val.ptr, d_y.ptr, sens.ptr, d_x.ptr, niter, 1, cfg[0], cfg[1],
                  ws.ptr, ws.nbytes, st)

    for _ in range(args.warmup):
        frame()
    stream.sync()
    ev0.record(st)
    t_wall = time.perf_counter()
    for _ in range(args.steps):
        frame()
    ev1.record(st)
    stream.sync()
    wall_s = time.perf_counter() - t_wall
    gpu_ms = ev0.elapsed_ms(ev1)
    ms_per_step = gpu_ms / args.steps
    fps = 1000.0 / ms_per_step
    x = d_x.numpy()
    log(f"MLEM: {ms_per_step:.2f} ms/frame ({niter} it), {fps:.2f} frames/s, wall {wall_s:.2f}s")

    # ---- dominant kernels timed alone with HIP events on the same stream
    lf, lb = cfg
    d_e = DeviceArray(m, np.float32)
    d_s = DeviceArray.from_numpy((sens.numpy() / sens.numpy().sum()).astype(np.float32))
    d_xx = DeviceArray(n, np.float32)
    d_part = DeviceArray(L.raptr_mlem_partial_slots(n), np.float64)
    R = args.kernel_reps

    def fwd():
        capi.call("raptr_mlem_forward_ratio", m, A.row_ptr.ptr, A.col.ptr, A.val.ptr, d_x.ptr,
                  d_y.ptr, d_e.ptr, lf, st)

    def bwd():
        capi.call("raptr_mlem_backward_update", n, At.row_ptr.ptr, At.col.ptr, At.val.ptr,
                  d_e.ptr, d_x.ptr, d_s.ptr, d_xx.ptr, d_part.ptr, lb, st)

    fwd(), bwd()
    ms_f = timed(lambda: [fwd() for _ in range(R)]) / R
    ms_b = timed(lambda: [bwd() for _ in range(R)]) / R
    bytes_f = 8 * A.nnz + 8 * (m + 1) + 4 * m + 4 * m + 4 * n
    bytes_b = 8 * A.nnz + 8 * (n + 1) + 4 * m + 4 * n * 3 + 8 * (n // 4)
    gbs_f = bytes_f / (ms_f * 1e-3) / 1e9
    gbs_b = bytes_b / (ms_b * 1e-3) / 1e9
    dom = ("mlem_forward_ratio", ms_f, bytes_f, gbs_f) if ms_f >= ms_b else \
        ("mlem_backward_update", ms_b, bytes_b, gbs_b)
    log(f"kernels: fwd {ms_f:.3f} ms {gbs_f:.0f} GB/s; bwd {ms_b:.3f} ms {gbs_b:.0f} GB/s")
    bytes_it = 2 * 8 * A.nnz + 8 * (m + 1) + 4 * 3 * m + 4 * 6 * n

    # ---- CPU baseline: oracle FP64 MLEM iteration on the same matrix
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle as O
        rp, col, val = A.to_host()
        v64 = val.astype(np.float64)
        s64 = sens.numpy()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        O.lib().ro_set_threads(threads)
        xh = np.full(n, 1.0 / n)
        t = time.perf_counter()
        O.mlem_f64(rp, col, v64, y.astype(np.float64), s64, xh, 1)
        it_s = time.perf_counter() - t
        cpu = {"value": round(1.0 / (it_s * niter), 6), "unit": "frames/s", "cores": threads,
               "kind": "port",
               "sample": f"1 MLEM iteration (FP64 C oracle, OpenMP forward, serial "
                         f"back-projection) over the same {A.nnz}-nnz matrix, {it_s:.2f} s; "
                         f"frames/s = 1/({niter} x iteration time)"}
        log(f"cpu baseline: {it_s:.2f} s/iteration -> {cpu['value']} frames/s")

    out = {
        "metric": "frames/s batched vs plain MLEM at <=1e-4 rel err; p50/p99 latency; HBM GB/s",
        "value": round(fps, 4),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, Poisson counts from the FP64 projection)",
        "config": {"workload": "plain MLEM full size: 13x13 heads x 180 angles (5,140,980 "
                               "channels), 52^3 grid @ 2 mm, 30 rays, 2e7 events, "
                               f"{niter} iterations/frame",
                   "measured_channels": m, "nnz": A.nnz, "voxels": n,
                   "iterations_per_frame": niter},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": round(dom[3], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dom[3] / HBM_PEAK_GBS, 4),
                     "traffic": None, "algorithmic_bytes": dom[2],
                     "avg_launch_ms": round(dom[1], 4)},
        "kernels": {"forward_cfg": cfg[0], "backward_cfg": cfg[1],
                    "forward_ratio_ms": round(ms_f, 4), "forward_GBps": round(gbs_f, 1),
                    "backward_update_ms": round(ms_b, 4), "backward_GBps": round(gbs_b, 1),
                    "iteration_bytes": bytes_it,
                    "iteration_GBps": round(bytes_it / (ms_per_step / niter * 1e-3) / 1e9, 1)},
        "build": {"count_fill_ms": round(ms_count, 1), "fill_ms": round(ms_fill, 1),
                  "transpose_ms": round(ms_t, 1), "sensitivity_ms": round(ms_s, 1),
                  "data_s": round(t_data, 2)},
        "cpu_baseline": cpu,
        "device": dev,
        "x_sum": float(x.astype(np.float64).sum()),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
