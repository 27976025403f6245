"""Benchmark of the raptr reconstruction hot path on one MI355X.

Workload (configuration of the metric): plain MLEM at full size --
13x13-pixel heads (4x4x20 mm, 460 mm apart), 180 angles at 1 deg
(5,140,980 channels), 52^3 grid at 2 mm, 30 Monte Carlo rays per channel,
phantom of 6 random ellipsoids (semi-axes 5-25 mm, activity log-uniform
0.5-8) on a 0.1 background, 2e7 Poisson events, 50 MLEM iterations per
frame.  The system matrix of the measured channels, its row-blocked
transpose and the FP64 sensitivity are built once and stay in HBM.

Modes on that workload (a step = one batch of frames, each frame run for
50 iterations, replayed from a HIP graph):
  plain    one frame at a time, FP32 values            (the plain GPU path)
  batched  F frames per launch, FP32 values; three batches side by side
           (the plan and two clones sharing its matrix, on their own
           streams), so a step is 3 x F frames (--side-by-side 1: one batch)
  mixed    as batched, 16-bit values (UNORM16 fixed point, FP16, BF16) with
           per-row FP32 scale, FP32 accumulation, FP64 normalisation
Frame 0 of every mode is the 2e7-event data set; its image is compared with
the FP64 host recurrence (oracle) run on the same matrix -- that run is also
the timed CPU baseline.  `value` = frames/s of the fastest mode whose error
is <= 1e-4 relative L2; every mode's frames/s and error are in `modes`.

Also reported: the 64-frame batched workload on the small geometry
(20,480 channels, 32^3 at 2.5 mm, 10 rays, decaying ellipsoids, counts
log-uniform 1e4-2e6 per frame, 30 iterations) in `batched_small`; the
largest serving geometry (16x16 heads, 60 angles, 96^3 at 2 mm, 20 rays,
5e6 events) as a single-frame plan in `plain_96`.

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
ERR_BAR = 1e-4
ERR_ITERS = 30  # iterations of the mixed-precision error figures
CPU_ITERS = 5  # iterations per timed run of the CPU baseline

# HBM bytes per launch measured with rocprofv3 (FETCH_SIZE scaled x2.00,
# WRITE_SIZE x1.00: the factors a 2 GiB streaming copy of this library
# reads back) by `python -m raptr_amd.passes` on this workload, for the
# kernel shapes the autotuner picks; other shapes report no traffic figure.
# Row order: y-major locality order for one frame.
MEASURED_TRAFFIC = {
    # forward (kernel, lanes, unroll, row order[, rows per union group]);
    # one launch per pass
    ("k_fwd1l", 16, 4, "y", 64): 2 * 5156424 * 1024 + 27546 * 1024,
    # backward with per-XCD row ranges (unroll, row blocks, row order,
    # 16-bit rows, lanes per column); one launch per pass (its reduce,
    # k_bwd1_xreduce, moves 0.006 GB more)
    ("k_bwd1x", 2, 69, "y", "r16", 32): 2 * 5467831 * 1024 + 10202 * 1024,
    ("k_bwd1x", 2, 368, "y", "r16", 32): 2 * 5855002 * 1024 + 8618 * 1024,
    # the fused iteration's products kernel (kernel, mode, waves, lanes per
    # row, row order, plan rows per union group) and its column reduce
    # (kernel, lanes per column or segment slot, row order, rows per union
    # group); one launch each per iteration
    ("k_fb1l", 2, 4, 64, "y", 64): 2 * 5035683 * 1024 + 465498 * 1024,
    ("k_fb1_reduce8", 16, "y", 64): 2 * 1028086 * 1024 + 6883 * 1024,
    ("k_fb1l", 2, 8, 64, "y", 256): 2 * 4908446 * 1024 + 221626 * 1024,
    ("k_fb1_treduce", 64, "y", 256, 1024, 512): 2 * 267869 * 1024 + 7869 * 1024,
}



def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2, help="timed batches per mode")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iters", type=int, default=50, help="MLEM iterations per frame")
    ap.add_argument("--frames", type=int, default=64, help="frames per batch")
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the FP64 host recurrence (no error figures either)")
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--side-by-side", type=int, default=3,
                    help="batched modes: batches of the same frames run side by side, the "
                         "plan and plan clones on their own streams (a step is this many x F "
                         "frames); 1: one batch")
    ap.add_argument("--kernel-reps", type=int, default=10)
    ap.add_argument("--modes",
                    default="plain,batched,mixed_u16,mixed_f16,mixed_bf16,plain_u16,plain_f16",
                    help="comma-separated subset of the modes to run")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=120)
    ap.add_argument("--serving-rate", type=float, default=20.0, help="arrivals per second")
    ap.add_argument("--serving-chunk", type=int, default=2,
                    help="most iterations a lane runs before admitting new requests")
    ap.add_argument("--serving-slots", type=int, default=64,
                    help="frame slots per geometry lane (64: grouped layouts on the matrix cores)")
    ap.add_argument("--row-order", default="locality", choices=("locality", "list", "both"),
                    help="plan row order: locality (angle, midpoint, slope), the measurement "
                         "list's, or try both")
    ap.add_argument("--locality-tile", type=int, default=8,
                    help="midpoint tile of the locality row order")
    ap.add_argument("--group", type=int, default=16, choices=(16, 32),
                    help="rows / voxels per group of the 64-frame grouped layouts")
    ap.add_argument("--block-slices-mb", default="2,8,32",
                    help="ratio-slice sizes (MB) of the batched plans' row blocks to try")
    return ap.parse_args()


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (got --gpus {args.gpus}); the row-sharded multi-GPU "
              "design is described in DESIGN.md and is not run here.", file=sys.stderr)
        sys.exit(2)

    from raptr_amd import _capi as capi
    from raptr_amd import geometry as G
    from raptr_amd import synthetic, sysmat
    from raptr_amd.device import DeviceArray, Event, Stream, device_name
    from raptr_amd.recon import Reconstructor

    dev = device_name()
    log("device:", dev)
    stream = Stream()
    st = stream.handle
    # the other batches of a batched mode (plan clones), one stream each
    st_more = [Stream() for _ in range(max(0, args.side_by_side - 1))]
    joined = [Event() for _ in st_more]
    ev0, ev1 = Event(), Event()

    def timed(fn):
        ev0.record(st)
        fn()
        ev1.record(st)
        return ev0.elapsed_ms(ev1)

    copy = copy_peak(timed, st)
    log(f"device copy: {copy['GBps']:.0f} GB/s (read + write) over {copy['bytes'] >> 20} MiB")
    if args.modes == "copy":
        print(json.dumps({"partial": True, "copy": copy, "device": dev}), flush=True)
        return
    if args.modes == "serving":
        serve = bench_serving(log, args.serving_requests, args.serving_rate, args.serving_slots,
                              args.serving_chunk)
        print(json.dumps({"partial": True, "serving": serve, "device": dev}), flush=True)
        return

    # ------------------------------------------------------------ data
    setup, grid, nrays = G.config_full()
    niter, F = args.iters, args.frames
    t0 = time.perf_counter()
    image, _ = synthetic.random_ellipsoids(grid, n=6, seed=2)
    meas, y, ybar = synthetic.measurement(setup, grid, nrays, image, 2e7, seed=1)
    t_data = time.perf_counter() - t0
    m, n = len(meas), grid.nvox
    log(f"data: {m} measured channels of {setup.nchannels} ({t_data:.1f} s)")

    # -------------------------------------------------------- matrix
    d_list = DeviceArray.from_numpy(meas)
    res = {}
    ms_build = timed(lambda: res.__setitem__("A", sysmat.build_csr(
        setup, grid, meas, nrays, d_list=d_list, stream=st, capacity="auto")))
    A = res["A"]
    ms_sens = timed(lambda: res.__setitem__("sens", sysmat.sensitivity(setup, grid, nrays,
                                                                       stream=st)))
    sens = res["sens"]
    log(f"matrix: nnz={A.nnz} ({A.nnz / m:.1f}/row) built in {ms_build:.0f} ms; "
        f"sensitivity (all channels, FP64) {ms_sens:.0f} ms")

    # frames: frame 0 = the data set, others Poisson draws of the same mean
    # with totals log-uniform in [2e6, 2e7] (rows with no count contribute 0)
    ym = ybar[meas]
    totals = np.concatenate([[float(y.sum())], np.geomspace(2e6, 2e7, F - 1)])
    Y = np.empty((m, F), np.float32)
    Y[:, 0] = y
    for f in range(1, F):
        Y[:, f] = synthetic.poisson_counts(ym, totals[f], seed=100 + f)
    x0 = np.full(n, 1.0 / n, np.float32)

    modes = {}
    plans = {}
    wanted = args.modes.split(",")
    orders = ["locality", "list"] if args.row_order == "both" else [args.row_order]
    # locality orders: y-major for the single-frame plans, slope tiles over
    # angle groups for the batched ones (recon.default_order_kind)
    from raptr_amd.recon import default_order_kind
    d_orders = ({k: sysmat.locality_order(setup, d_list, stream=st, tile=args.locality_tile,
                                          kind=k)
                 for k in {default_order_kind(1), default_order_kind(F)}}
                if "locality" in orders else {})
    slices_mb = [float(v) for v in args.block_slices_mb.split(",")]
    best_rows = best_order = None
    plain_images = None  # every frame by the plain plan (per-frame errors)
    for name, vt, nf in (("plain", "f32", 1), ("batched", "f32", F), ("mixed_u16", "u16", F),
                         ("mixed_f16", "f16", F), ("mixed_bf16", "bf16", F),
                         ("plain_u16", "u16", 1), ("plain_f16", "f16", 1)):
        if name not in wanted:
            continue
        yy = Y[:, 0].copy() if nf == 1 else Y
        dy = DeviceArray.from_numpy(yy.reshape(-1))
        xx0 = x0 if nf == 1 else np.repeat(x0[:, None], nf, axis=1)
        dx = DeviceArray.from_numpy(xx0.reshape(-1))
        # row blocks of the backward: for F frames try several ratio-slice
        # sizes (the first batched mode picks, the others reuse its choice)
        # one frame: a few row-block sizes are tried
        if nf == 1 and name != "plain":
            rows_try = [("blocked", 0)]
        elif nf == 1:
            # (block_rows 0: the plan's default, 65,536 rows for one frame)
            rows_try = [("blocked", 0), ("blocked", 12288), ("blocked", 8192)]
        elif best_rows is not None:
            rows_try = [("blocked", best_rows)]
        else:
            rows_try = [("blocked", max(4, int(mb * 2**20 / (4 * nf)) // 4 * 4))
                        for mb in slices_mb]
        rows_try = [(lay, rows, o) for lay, rows in rows_try
                    for o in (orders if nf == 1 or best_order is None else [best_order])]
        R, t_plan, cfg, ms_it = None, 0.0, None, None
        for layout, rows, order in rows_try:
            t = time.perf_counter()
            Rc = Reconstructor(A, sens, vt, nf, block_rows=rows, stream=st, layout=layout,
                               row_order=(d_orders[default_order_kind(nf)]
                                          if order == "locality" else None),
                               grid=grid, group=args.group)
            tp = time.perf_counter() - t
            cc, mc = Rc.autotune(dy, dx, nf, reps=2, stream=st)
            if len(rows_try) > 1:
                log(f"{name}: {layout} layout, block_rows {rows} "
                    f"({Rc.info()['row_blocks']} blocks), {order} row order "
                    f"-> {mc:.2f} ms/iteration")
            if ms_it is None or mc < ms_it:
                if R is not None:
                    R.close()
                R, t_plan, cfg, ms_it = Rc, tp, cc, mc
                chosen_rows, chosen_order = rows, order
                cfg = dict(cfg, layout=layout,
                           row_order=default_order_kind(nf) if order == "locality" else order)
            else:
                Rc.close()
        if nf > 1:
            best_rows, best_order = chosen_rows, chosen_order
        plans[name] = (R, dy, dx, xx0, nf)
        # batches: a second batch of the same frames on a clone of the plan
        # (shared matrix, own work vectors) on a second stream, side by side
        sides = []
        if nf > 1:
            for k in range(max(0, args.side_by_side - 1)):
                sides.append((R.clone(stream=st_more[k].handle),
                              DeviceArray.from_numpy(yy.reshape(-1)),
                              DeviceArray.from_numpy(xx0.reshape(-1)), st_more[k]))
        nb = 1 + len(sides)

        def batch():
            dx.copy_from(xx0, stream=st)
            R.run(dy, dx, niter, nf, use_graph=True, stream=st)
            for C_, dy_, dx_, s_ in sides:
                dx_.copy_from(xx0, stream=s_.handle)
                C_.run(dy_, dx_, niter, nf, use_graph=True, stream=s_.handle)
            for k, (_, _, _, s_) in enumerate(sides):
                joined[k].record(s_.handle)
                stream.wait(joined[k])

        # the 16-bit batched modes miss the error bar (never the value):
        # a shorter timing
        steps, warm = ((min(args.steps, 5), min(args.warmup, 1)) if name.startswith("mixed")
                       else (args.steps, args.warmup))
        for _ in range(warm):
            batch()
        stream.sync()
        ms = timed(lambda: [batch() for _ in range(steps)]) / steps
        xr = dx.numpy().reshape(n, nf)
        same = all(np.array_equal(dx_.numpy(stream=s_.handle).reshape(n, nf), xr)
                   for _, _, dx_, s_ in sides)
        for C_, _, _, _ in sides:
            C_.close()
        sides = []
        fps = nb * nf * 1000.0 / ms
        # the mixed-precision error figure is quoted at 30 iterations
        dx.copy_from(xx0, stream=st)
        R.run(dy, dx, ERR_ITERS, nf, use_graph=True, stream=st)
        x30 = dx.numpy(stream=st).reshape(n, nf)[:, 0].copy()
        info = R.info()
        modes[name] = {"value_type": vt, "frames_per_batch": nf, "batches_side_by_side": nb,
                       "ms_per_step": round(ms, 3), "steps_timed": steps,
                       "frames_per_s": round(fps, 3), "ms_per_iteration": round(ms / niter, 4),
                       "config": dict(cfg, row_blocks=info["row_blocks"]),
                       "plan_build_s": round(t_plan, 2),
                       "plan_device_GB": round(info["device_bytes"] / 1e9, 2),
                       "image0": xr[:, 0].copy(), "image0_30": x30}
        de_m = DeviceArray(m * nf, np.float32)

        def pf():
            capi.call("raptr_recon_forward", R.handle, nf, dx.ptr, dy.ptr, de_m.ptr, st)

        def pb():
            capi.call("raptr_recon_backward", R.handle, nf, dx.ptr, st)

        pf(), pb()
        modes[name]["passes_ms"] = {"forward": round(timed(lambda: [pf() for _ in range(3)]) / 3, 3),
                                    "backward": round(timed(lambda: [pb() for _ in range(3)]) / 3, 3)}
        del de_m
        if nb > 1:
            modes[name]["side_by_side_images_equal"] = same
        log(f"{name}: {ms:.1f} ms per step of {nb} x {nf} frames -> {fps:.2f} frames/s "
            f"(cfg {cfg}, {info['row_blocks']} row blocks, plan {t_plan:.1f} s, "
            f"passes {modes[name]['passes_ms']})")
        if name == "plain" and nf < F and F > 1:
            # every frame of the batch by the plain plan: the per-frame
            # reference of the batched modes (FP32, 1.2e-6 from FP64)
            plain_images = np.empty((n, F), np.float32)
            dyf = DeviceArray(m, np.float32)
            for f in range(F):
                dyf.copy_from(Y[:, f].copy(), stream=st)
                dx.copy_from(x0, stream=st)
                R.run(dyf, dx, niter, 1, use_graph=True, stream=st)
                plain_images[:, f] = dx.numpy(stream=st)
            del dyf
        if nf > 1:
            modes[name]["roofline"] = batched_roofline(
                R, info, nf, m, n, A.nnz, vt, args.kernel_reps, timed, capi, dx, dy, st)
            log(f"{name}: roofline {modes[name]['roofline']}")
        if nf > 1 and plain_images is not None:
            per = [rel_l2(xr[:, f], plain_images[:, f].astype(np.float64)) for f in range(nf)]
            modes[name]["rel_l2_vs_plain_per_frame"] = [float(f"{v:.3e}") for v in per]
            modes[name]["rel_l2_vs_plain_max"] = max(per)
            modes[name]["image63"] = xr[:, nf - 1].copy()
        if name != "plain":
            R.close()
            del plans[name], R, dy, dx

    if "plain" not in modes:
        # a partial run (profiling a subset of modes): per-mode figures only
        for name in modes:
            modes[name].pop("image0")
            modes[name].pop("image0_30")
            modes[name].pop("image63", None)
        print(json.dumps({"partial": True, "modes": modes, "device": dev}), flush=True)
        return

    # ------------------------------------------- roofline of the plain path
    R, dy, dx, _, _ = plans["plain"]
    de = DeviceArray(m, np.float32)
    reps = args.kernel_reps

    def fwd():
        capi.call("raptr_recon_forward", R.handle, 1, dx.ptr, dy.ptr, de.ptr, st)

    def bwd():
        capi.call("raptr_recon_backward", R.handle, 1, dx.ptr, st)

    fwd(), bwd()
    ms_f = timed(lambda: [fwd() for _ in range(reps)]) / reps
    ms_b = timed(lambda: [bwd() for _ in range(reps)]) / reps
    info = R.info()
    nb = info["row_blocks"]
    cmp_ = max(info["compact_rows"], 0)
    cols16, rows16 = bool(cmp_ & 2), bool(cmp_ & 1)
    fwd_kernel = ("k_fwd1l" if info["fwd_kind"] == 2 else
                  "k_fwd1c" if cols16 else "k_fwd1")
    urows = R.unions()["rows"]
    fwd_name = f"{fwd_kernel} (forward + ratio, one launch" + (
        f", x staged in LDS per {urows} rows, 16-bit ranks)" if fwd_kernel == "k_fwd1l" else
        ", 16-bit columns)" if cols16 else ")")
    # bytes as stored: 16-bit indices move 2 bytes less per entry than the
    # 4-byte index of the algorithmic count (plus 8 bytes of split points per
    # row for the columns)
    stored_f = stored_b = None
    # algorithmic bytes: matrix (value + column index), row pointers, vectors
    bytes_f = 8 * info["padded_nnz"] + 8 * (m + 1) + 4 * m + 4 * m + 4 * n
    if fwd_kernel == "k_fwd1l":
        # 2-byte ranks; per union entry its column and x (4 + 4 bytes);
        # union offsets per group of rows; no x term (read through the unions)
        nu = R.forward_union_entries()
        stored_f = (bytes_f - 2 * info["padded_nnz"] - 4 * n + 8 * nu
                    + 8 * (-(-m // urows) + 1))
    elif cols16:
        stored_f = bytes_f - 2 * info["padded_nnz"] + 8 * m
    if info["bwd_staged"] == 3:
        # one pass: matrix (value + row), two segment bounds per column and
        # block, the ratio vector (counted once; each XCD reads its eighth),
        # the 8 per-XCD partial sums written and read, x/s read, x' written
        bytes_b = (8 * info["transpose_nnz"] + 8 * nb * (n + 1) + 4 * m + 2 * 8 * 4 * n
                   + 12 * n)
        if rows16:
            stored_b = bytes_b - 2 * info["transpose_nnz"]
        bwd_name = (f"k_bwd1x + k_bwd1_xreduce (backward + update, one pass over {nb} row "
                    "blocks, per-XCD row ranges" + (", 16-bit rows)" if rows16 else ")"))
    else:
        bytes_b = (8 * info["transpose_nnz"] + 8 * nb * (n + 1) + 4 * m
                   + 4 * n * (2 * nb + 2) + 8 * (n // 4))
        bwd_name = f"k_bwd (backward + update, pass of {nb} row-block launches)"
    gbs_f = bytes_f / (ms_f * 1e-3) / 1e9
    gbs_b = bytes_b / (ms_b * 1e-3) / 1e9
    log(f"forward pass {ms_f:.3f} ms ({gbs_f:.0f} GB/s), backward pass {ms_b:.3f} ms "
        f"({bwd_name}, {gbs_b:.0f} GB/s)")
    # launches per pass: the blocked backward is one launch per group of
    # row blocks; the roofline is quoted per launch (bytes and time averaged)
    cfg_p = modes["plain"]["config"]
    nl_f = 1
    nl_b = 1 if info["bwd_staged"] else -(-nb // cfg_p["bwd_blocks_per_launch"])
    key_f = (fwd_kernel, cfg_p["fwd_lanes"], cfg_p["fwd_unroll"], cfg_p["row_order"])
    if fwd_kernel == "k_fwd1l":
        key_f += (urows,)
    key_b = ("k_bwd", cfg_p["bwd_unroll"], nb, cfg_p["bwd_blocks_per_launch"],
             cfg_p["bwd_staged"], cfg_p["row_order"])
    if info["bwd_staged"] == 3:
        lpc = cfg_p["bwd_lanes_per_column"]
        key_b = ("k_bwd1x", cfg_p["bwd_unroll"], nb, cfg_p["row_order"],
                 "r16" if rows16 else "r32", lpc if lpc in (16, 32) else 64)
    dom = {"kernel": fwd_name, "ms": ms_f, "bytes": bytes_f, "gbs": gbs_f, "launches": nl_f,
           "key": key_f, "stored": stored_f}
    if ms_b > ms_f:
        dom = {"kernel": bwd_name, "ms": ms_b, "bytes": bytes_b, "gbs": gbs_b, "launches": nl_b,
               "key": key_b, "stored": stored_b}
    traffic = MEASURED_TRAFFIC.get(dom["key"]) if dom["key"] else None
    # the fused iteration (the plain plan's default where built): the
    # products kernel k_fb1l reads the matrix once for both products; its
    # algorithmic bytes are those of the two SpMVs it replaces
    fused = None
    if info.get("fused"):
        def fb1():
            capi.call("raptr_recon_fused_pass", R.handle, dx.ptr, dy.ptr, 1, st)

        def fb2():
            capi.call("raptr_recon_fused_pass", R.handle, dx.ptr, dy.ptr, 2, st)

        fb1(), fb2()
        ms_fb = timed(lambda: [fb1() for _ in range(reps)]) / reps
        ms_red = timed(lambda: [fb2() for _ in range(reps)]) / reps
        nu = R.forward_union_entries()
        urows = R.unions()["rows"]
        tz = R.fused_tiles()
        spmv = 8 * A.nnz + 4 * (m + 1) + 4 * m + 4 * n  # one direction
        bytes_fb = 2 * spmv
        # stored: 6 bytes per padded entry (FP32 value, 16-bit union rank),
        # row pointers, y, e, per union entry its column and x, the partial
        # sums written
        stored_fb = (6 * info["padded_nnz"] + 8 * (m + 1) + 8 * m + 8 * nu + 4 * nu
                     + 8 * (-(-m // urows) + 1))
        it_bytes = 2 * 8 * A.nnz + 8 * (m + 1) + 4 * 3 * m + 4 * 6 * n  # one iteration
        ms_it = modes["plain"]["ms_per_iteration"]
        key = ("k_fb1l", info["fused"], info["fused_waves"], info["fused_lanes_per_row"],
               cfg_p["row_order"], urows)
        fused = {"kernel": f"k_fb1l<f32, {info['fused_waves']} waves, "
                           f"{info['fused_lanes_per_row']} lanes per row, "
                           + ("one shared column-sum region" if info["fused"] == 2 else
                              "a column-sum region per wave")
                           + f"> (forward + ratio + backward partial sums over each {urows}-row "
                             "group's column union, one pass over the matrix)",
                 "ms": ms_fb, "bytes": bytes_fb, "gbs": bytes_fb / (ms_fb * 1e-3) / 1e9,
                 "stored": stored_fb, "key": key,
                 "reduce": {"kernel": (
                                f"k_fb1_treduce<{info['fused_lanes_per_column']} lanes per "
                                f"slot> + k_fb1_tcombine (per part of a tile of "
                                f"{tz['columns']} columns, {tz['segments_per_workgroup']} "
                                f"segments: the groups' partials in written order into LDS "
                                f"column sums; then the parts in order and the update)"
                                if info["fused_reduce_per_xcd"] == 2 else
                                "k_fb1_reduce8 + k_bwd1_xreduce (group partials per column "
                                "in per-XCD group ranges, then the update)"
                                if info["fused_reduce_per_xcd"] else
                                "k_fb1_reduce (group partials per column, update)"),
                            "ms": round(ms_red, 4),
                            # mode 2: each partial and its 16-bit column
                            # inside its tile once, the segment bounds, the
                            # part sums and x, s, x'; else each partial and
                            # its position once
                            "algorithmic_bytes": int(
                                6 * nu + 8 * tz["segments"] + 4 * tz["max_parts"] * n + 12 * n
                                if info["fused_reduce_per_xcd"] == 2 else
                                4 * 2 * nu + 8 * (n + 1) + 4 * 4 * n),
                            "traffic": MEASURED_TRAFFIC.get(
                                ("k_fb1_treduce", info["fused_lanes_per_column"],
                                 cfg_p["row_order"], urows, tz["columns"],
                                 tz["segments_per_workgroup"])
                                if info["fused_reduce_per_xcd"] == 2 else
                                ("k_fb1_reduce8", info["fused_lanes_per_column"],
                                 cfg_p["row_order"], urows))},
                 "iteration": {"ms": ms_it, "algorithmic_bytes": it_bytes,
                               "achieved": round(it_bytes / (ms_it * 1e-3) / 1e9, 1),
                               "frac": round(it_bytes / (ms_it * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                               "bytes_basis": "one MLEM iteration by the survey's formula: "
                                              "the matrix counted once per product"}}
        log(f"fused products {ms_fb:.3f} ms ({fused['gbs']:.0f} GB/s of the two products' "
            f"bytes), reduce {ms_red:.3f} ms, iteration {ms_it:.3f} ms")

    # ------------------------ FP64 host recurrence: errors + CPU baseline
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle as O
        rp, col, val = A.to_host()
        v64 = val.astype(np.float64)
        del val
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        O.lib().ro_set_threads(threads)
        t = time.perf_counter()
        x64_30 = O.mlem_f64(rp, col, v64, y.astype(np.float64), sens.numpy(),
                            x0.astype(np.float64), ERR_ITERS)
        x64 = O.mlem_f64(rp, col, v64, y.astype(np.float64), sens.numpy(), x64_30,
                         niter - ERR_ITERS)
        cpu_s = time.perf_counter() - t
        # the batch's last frame (the per-frame check of the batched modes)
        x64_63 = None
        if plain_images is not None:
            x64_63 = O.mlem_f64(rp, col, v64, Y[:, F - 1].astype(np.float64), sens.numpy(),
                                x0.astype(np.float64), niter)
        # one host thread on a bounded sample: 3 iterations, scaled to niter
        O.lib().ro_set_threads(1)
        t = time.perf_counter()
        O.mlem_f64(rp, col, v64, y.astype(np.float64), sens.numpy(), x0.astype(np.float64), 3)
        cpu1_s = (time.perf_counter() - t) * niter / 3.0
        O.lib().ro_set_threads(threads)
        del v64
        # the reference's FP32 recurrence restated on the host cores: one
        # warm-up, then the median of 5 runs of CPU_ITERS iterations, scaled
        # to niter
        val32 = A.val.numpy(A.nnz)
        sens64 = sens.numpy()
        O.mlem_f32(rp, col, val32, y, sens64, x0, 1)
        t32 = []
        for _ in range(5):
            t = time.perf_counter()
            O.mlem_f32(rp, col, val32, y, sens64, x0, CPU_ITERS)
            t32.append(time.perf_counter() - t)
        cpu32_s = float(np.median(t32)) * niter / CPU_ITERS
        del rp, col, val32
        for name in modes:
            if x64_63 is not None and "image63" in modes[name]:
                modes[name]["rel_l2_vs_fp64_last_frame"] = rel_l2(modes[name].pop("image63"),
                                                                  x64_63)
            modes[name]["rel_l2_vs_fp64"] = rel_l2(modes[name]["image0"], x64)
            modes[name][f"rel_l2_vs_fp64_{ERR_ITERS}it"] = rel_l2(modes[name]["image0_30"], x64_30)
            modes[name][f"max_voxel_err_vs_fp64_{ERR_ITERS}it"] = float(
                np.max(np.abs(modes[name]["image0_30"] - x64_30)) / np.max(np.abs(x64_30)))
            modes[name]["max_voxel_err_vs_fp64"] = float(
                np.max(np.abs(modes[name]["image0"] - x64)) / np.max(np.abs(x64)))
            log(f"{name}: rel L2 vs FP64 oracle {modes[name]['rel_l2_vs_fp64']:.2e}")
        cpu = {"value": round(1.0 / cpu32_s, 6), "unit": "frames/s", "cores": threads,
               "kind": "port",
               "sample": f"the reference's FP32 MLEM recurrence restated in C (OpenMP) on the "
                         f"same {A.nnz}-entry matrix: median of 5 runs of {CPU_ITERS} "
                         f"iterations after one warm-up, scaled to {niter} iterations "
                         f"({cpu32_s * 1e3 / niter:.1f} ms per iteration)",
               "ms_per_iteration": round(cpu32_s * 1e3 / niter, 2)}
        cpu["fp64"] = {"value": round(1.0 / cpu_s, 6), "unit": "frames/s", "cores": threads,
                       "sample": f"one full frame ({niter} iterations) by the FP64 oracle: "
                                 f"{cpu_s:.1f} s"}
        cpu["single_thread"] = {"value": round(1.0 / cpu1_s, 6), "unit": "frames/s", "cores": 1,
                                "sample": f"FP64, 3 iterations on one thread, scaled to {niter}"}
        log(f"cpu baseline: FP32 {cpu32_s:.1f} s/frame, FP64 {cpu_s:.1f} s/frame on {threads} "
            f"threads, FP64 {cpu1_s:.1f} s/frame on one thread")

    for name in modes:
        modes[name].pop("image0")
        modes[name].pop("image0_30")
        modes[name].pop("image63", None)
    # a mode qualifies when frame 0 (and the batch's last frame) meet the bar
    # against the FP64 oracle and every frame is within it of the plain path
    ok = [k for k in modes if modes[k].get("rel_l2_vs_fp64", 0.0) <= ERR_BAR
          and modes[k].get("rel_l2_vs_fp64_last_frame", 0.0) <= ERR_BAR
          and modes[k].get("rel_l2_vs_plain_max", 0.0) <= ERR_BAR]
    best = max(ok, key=lambda k: modes[k]["frames_per_s"]) if ok else "plain"
    plain_fps = modes["plain"]["frames_per_s"]
    for k in modes:
        modes[k]["speedup_vs_plain"] = round(modes[k]["frames_per_s"] / plain_fps, 2)

    # ------------------------------------------------ batched small workload
    small = None
    if not args.no_small:
        small = bench_small(log, timed, stream, st,
                            0 if args.no_cpu_baseline else
                            int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))

    p96 = None
    if not args.no_small:
        p96 = bench_plain_96(log, timed, stream, st, niter,
                             0 if args.no_cpu_baseline else
                             int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))

    phases = None
    if not args.no_small:
        phases = bench_build_phases(log, timed, stream, st,
                                    0 if args.no_cpu_baseline else
                                    int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))

    # ------------------------------------------------------------ serving
    serve = None
    nnz_full = A.nnz
    if not args.no_serving:
        del plans, R, dy, dx, de
        res.pop("A")  # free the full-size matrix first
        A = None
        serve = bench_serving(log, args.serving_requests, args.serving_rate, args.serving_slots,
                              args.serving_chunk)

    separate = None
    if fused is not None:  # the plain path's dominant kernel is the fused one
        separate = {"forward": {"kernel": fwd_name, "ms": round(ms_f, 4),
                                "GBps": round(gbs_f, 1)},
                    "backward": {"kernel": bwd_name, "ms": round(ms_b, 4),
                                 "GBps": round(gbs_b, 1), "frac": round(gbs_b / HBM_PEAK_GBS, 4)},
                    "ms_per_iteration_measured_by": "autotune (the separate passes are "
                                                    "slower; not run in the timed region)"}
        traffic = MEASURED_TRAFFIC.get(fused["key"])
        dom = {"kernel": fused["kernel"], "ms": fused["ms"], "bytes": fused["bytes"],
               "gbs": fused["gbs"], "launches": 1, "key": fused["key"],
               "stored": fused["stored"]}
    plain_roofline = {"bound": "hbm", "kernel": dom["kernel"], "achieved": round(dom["gbs"], 1),
                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(dom["gbs"] / HBM_PEAK_GBS, 4),
                 "traffic": int(traffic) if traffic else None,
                 "traffic_source": ("rocprofv3 FETCH_SIZE x2.00 + WRITE_SIZE x1.00 per launch, "
                                    "factors from a 2 GiB streaming copy") if traffic else None,
                 "algorithmic_bytes": int(dom["bytes"] / dom["launches"]),
                 "bytes_basis": ("both products' bytes (8 per matrix entry each, the "
                                 "survey's SpMV formula twice): the kernel reads the matrix once "
                                 "for two products, so frac can exceed 1; traffic/peak is the "
                                 "HBM use" if fused is not None else
                                 "4-byte value + 4-byte index per stored entry"),
                 "stored_bytes": (int(dom["stored"] / dom["launches"])
                                  if dom["stored"] else None),
                 "stored_GBps": (round(dom["stored"] / (dom["ms"] * 1e-3) / 1e9, 1)
                                 if dom["stored"] else None),
                 "avg_launch_ms": round(dom["ms"] / dom["launches"], 4),
                 "launches_per_pass": dom["launches"], "pass_ms": round(dom["ms"], 4),
                 "copy_GBps_measured": copy["GBps"]}
    if traffic:
        plain_roofline["traffic_frac"] = round(traffic / (dom["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                               4)
    if fused is not None:
        plain_roofline["reduce"] = fused["reduce"]
        plain_roofline["iteration"] = fused["iteration"]
        plain_roofline["separate_passes"] = separate
    # the roofline line describes the headline mode's dominant kernel (the
    # batched backward when a batched mode is the value); the plain path's is
    # kept beside it
    headline_roofline = dict(modes[best].get("roofline", plain_roofline))
    headline_roofline["copy_GBps_measured"] = copy["GBps"]
    headline_roofline["mode"] = best
    out = {
        "metric": "frames/s batched vs plain MLEM at <=1e-4 rel err; p50/p99 latency; HBM GB/s",
        "value": modes[best]["frames_per_s"],
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": modes[best]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": modes[best]["value_type"],
        "data": "synthetic (seeded ellipsoid phantom, Poisson counts from the FP64 projection)",
        "config": {"workload": "full-size MLEM: 13x13 heads x 180 angles (5,140,980 channels), "
                               "52^3 grid @ 2 mm, 30 rays, 2e7 events, "
                               f"{niter} iterations/frame, mode '{best}'",
                   "measured_channels": m, "nnz": nnz_full, "voxels": n,
                   "frames_per_batch": modes[best]["frames_per_batch"],
                   "batches_side_by_side": modes[best]["batches_side_by_side"]},
        "modes": modes,
        "plain_frames_per_s": plain_fps,
        "speedup_vs_plain": modes[best]["speedup_vs_plain"],
        "roofline": headline_roofline,
        "roofline_plain": plain_roofline,
        "passes": {"forward_kernel": fwd_name,
                   "forward_ms": round(ms_f, 4), "forward_GBps": round(gbs_f, 1),
                   "backward_ms": round(ms_b, 4), "backward_GBps": round(gbs_b, 1),
                   "backward_kernel": bwd_name},
        "build": {"matrix_ms": round(ms_build, 1), "sensitivity_ms": round(ms_sens, 1),
                  "data_s": round(t_data, 2)},
        "batched_small": small,
        "plain_96": p96,
        "build_phases_small": phases,
        "serving": serve,
        "device_copy": copy,
        "cpu_baseline": cpu,
        "device": dev,
    }
    print(json.dumps(out), flush=True)


# HBM bytes per launch of the 64-frame grouped backward (rocprofv3
# FETCH_SIZE x2.00 + WRITE_SIZE x1.00, the factors of the 2 GiB streaming
# copy), full-size workload, slope-tiled locality row order, keyed by
# (kernel, group,
# row blocks, row blocks per XCD and launch); per launch of k_bwdm64x
MEASURED_TRAFFIC_BATCHED = {
    # (kernel, group, row blocks, row blocks per XCD and launch): bytes per
    # launch of the 64-frame backward (5 launches per pass here)
    ("k_bwdm64x", 16, 138, 4): 2 * 5455964 * 1024 + 291307 * 1024,
    # (kernel, group, unroll): the forward, one launch per pass
    ("k_fwdm64", 16, 8): 2 * 9235221 * 1024 + 1127705 * 1024,
}


def batched_roofline(R, info, nf, m, n, nnz, vt, reps, timed, capi, dx, dy, st):
    """Roofline of an F-frame plan's passes: the backward (the slower pass
    and the dominant kernel of the batched iteration) and the forward, timed
    with HIP events on the work stream over `reps` launches each.
    Algorithmic bytes per pass (matrix value + 4-byte index per stored
    entry, 4-byte row pointers, F-frame vectors):
      forward   (b_v + 4)*nnz + 4*(m+1) + F*4*n (x) + F*4*m (y) + F*4*m (e)
      backward  (b_v + 4)*nnz + 4*(m+1) + F*4*m (e) + F*4*n (x) + 4*n (s)
                + F*4*n (x')"""
    from raptr_amd.device import DeviceArray
    bv = 4 if vt == "f32" else 2
    e = DeviceArray(m * nf, np.float32)
    fwd = lambda: capi.call("raptr_recon_forward", R.handle, nf, dx.ptr, dy.ptr, e.ptr, st)  # noqa
    bwd = lambda: capi.call("raptr_recon_backward", R.handle, nf, dx.ptr, st)  # noqa: E731
    fwd(), bwd()
    ms_f = timed(lambda: [fwd() for _ in range(reps)]) / reps
    ms_b = timed(lambda: [bwd() for _ in range(reps)]) / reps
    del e
    bytes_f = (bv + 4) * nnz + 4 * (m + 1) + nf * 4 * n + nf * 8 * m
    bytes_b = (bv + 4) * nnz + 4 * (m + 1) + nf * 4 * m + nf * 8 * n + 4 * n
    g = info.get("group", 0)
    kind = info.get("grouped_kind", 0)
    grouped = bool(info.get("grouped64")) and nf == 64
    key = None
    nl = 1
    if grouped and kind == 3:
        xb = capi.lib().raptr_recon_get_xcd_blocks(R.handle)
        per_xcd = -(-info["row_blocks"] // 8)
        nl = -(-per_xcd // xb) if xb > 0 else 1
        kb = (f"k_bwdm64x<{vt}, {g}> ({nl} launches per pass, {xb or per_xcd} row blocks per "
              "XCD each, per-XCD row ranges) + k_bwd_xreduce")
        key = ("k_bwdm64x", g, info["row_blocks"], xb)
    else:
        kb = "general F-frame backward"
    stored_b = None
    if grouped:
        # union entries (key + G values), list pointers per (row block, group)
        stored_b = (info["grouped_bwd_entries"] * (4 + g * bv) + 8 * info["row_blocks"] *
                    (-(-n // g) + 1) + nf * 4 * m + nf * 8 * n + 4 * n)
    traffic = MEASURED_TRAFFIC_BATCHED.get(key) if key else None
    traffic_f = (MEASURED_TRAFFIC_BATCHED.get(("k_fwdm64", g, info.get("grouped_unroll")))
                 if grouped else None)
    gb_b = bytes_b / (ms_b * 1e-3) / 1e9
    gb_f = bytes_f / (ms_f * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": kb, "achieved": round(gb_b, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(gb_b / HBM_PEAK_GBS, 4),
            "traffic": int(traffic) if traffic else None,
            "traffic_source": ("rocprofv3 FETCH_SIZE x2.00 + WRITE_SIZE x1.00 per launch, "
                               "factors from a 2 GiB streaming copy") if traffic else None,
            "algorithmic_bytes": int(bytes_b / nl), "bytes_basis":
                "per launch of the pass's nl launches: 4-byte value (2 for 16-bit) + 4-byte "
                "index per matrix entry, F-frame vectors",
            "stored_bytes": int(stored_b / nl) if stored_b else None,
            "stored_GBps": round(stored_b / (ms_b * 1e-3) / 1e9, 1) if stored_b else None,
            "avg_launch_ms": round(ms_b / nl, 4), "launches_per_pass": nl,
            "pass_ms": round(ms_b, 4), "algorithmic_bytes_per_pass": int(bytes_b),
            "frames": nf,
            "forward": {"kernel": f"k_fwdm64<{vt}, {g}>" if grouped else "general F-frame forward",
                        "ms": round(ms_f, 4), "algorithmic_bytes": int(bytes_f),
                        "achieved": round(gb_f, 1), "frac": round(gb_f / HBM_PEAK_GBS, 4),
                        "traffic": int(traffic_f) if traffic_f else None}}


def copy_peak(timed, st, nbytes=2 << 30, reps=5):
    """Device-to-device copy of nbytes by the library's streaming kernel
    (k_stream_copy): the box's attainable read+write bandwidth, and the
    known-size dispatch that calibrates the FETCH_SIZE / WRITE_SIZE
    counters."""
    from raptr_amd import _capi as capi
    from raptr_amd.device import DeviceArray
    a = DeviceArray(nbytes // 4, np.float32)
    b = DeviceArray(nbytes // 4, np.float32)
    capi.call("raptr_memset", a.ptr, 0, nbytes, st)
    capi.call("raptr_stream_copy", b.ptr, a.ptr, nbytes // 4, st)
    ms = timed(lambda: [capi.call("raptr_stream_copy", b.ptr, a.ptr, nbytes // 4, st)
                        for _ in range(reps)]) / reps
    del a, b
    return {"bytes": nbytes, "ms": round(ms, 4), "GBps": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}


def bench_serving(log, nreq, rate, slots=64, chunk=2):
    """Poisson arrivals of reconstruction requests over three geometries
    (8x8 heads / 32^3 / 180 angles at 1 deg, 13x13 / 52^3 / 120 at 1.5 deg,
    16x16 / 96^3 / 60 at 3 deg; 2 mm voxels, 20 rays), 1e4-5e6 events and
    10-50 iterations per request, served by continuous batching (`slots` frame
    slots per geometry; 64: grouped-only plans on the matrix cores) from a
    pre-built device matrix cache.  Latency = arrival to image on the host."""
    from raptr_amd import geometry as G
    from raptr_amd import serving, synthetic, sysmat
    geoms = [serving.Geometry(G.two_head_setup(p, na, 180.0 / na), G.centred_grid(v, 2.0), 20)
             for p, v, na in ((8, 32, 180), (13, 52, 120), (16, 96, 60))]
    cache = serving.MatrixCache(200 << 30, frames=slots)
    t = time.perf_counter()
    builds = {}
    for g in geoms:
        lane = cache.admit(g)
        builds[f"{g.setup.n0z}x{g.setup.n0y}/{g.grid.n[0]}^3"] = {
            "channels": g.nchannels, "nnz": lane.nnz, "build_s": round(lane.build_s, 2),
            "device_GB": round(lane.bytes / 1e9, 2)}
        log(f"serving lane {g.setup.n0z}x{g.setup.n0y} heads, {g.grid.n[0]}^3: "
            f"nnz {lane.nnz}, built in {lane.build_s:.1f} s, {lane.bytes / 1e9:.1f} GB")
    ybar = {}
    for g in geoms:
        img, _ = synthetic.random_ellipsoids(g.grid, n=6, seed=5)
        ybar[g] = sysmat.forward_project(g.setup, g.grid, img.astype(np.float32), g.nrays)
    reqs = serving.synthetic_requests(geoms, rate, nreq, seed=11, ybar=ybar)
    t_prep = time.perf_counter() - t
    stats = serving.Server(cache, max_chunk=chunk).serve(reqs, realtime=True)
    out = stats.summary()
    out.update({"workload": f"{nreq} requests, Poisson arrivals at {rate}/s, geometries "
                            "8x8/32^3/180 angles, 13x13/52^3/120, 16x16/96^3/60 (2 mm voxels, "
                            f"20 rays), 1e4-5e6 events, 10-50 iterations, {slots} slots per "
                            f"geometry, chunks of at most {chunk} iterations",
                "lanes": builds, "prep_s": round(t_prep, 1)})
    log(f"serving: p50 {out['p50_ms']:.0f} ms, p99 {out['p99_ms']:.0f} ms, "
        f"{out['requests_per_s']:.1f} req/s over {out['wall_s']:.1f} s")
    for g in list(cache.lanes):
        cache.lanes.pop(g).close()
    return out


def bench_build_phases(log, timed, stream, st, cpu_threads=0):
    """The matrix build on the small geometry (20,480 channels x 32^3 voxels,
    10 rays): the workqueue (cylinder cull + emit of (channel, voxel)
    pairs), MC ray sampling of the workqueue's entries, and the fused
    cull + sample + compaction into CSR that plans are built from.  GPU:
    all channels, device buffers, HIP events.  CPU: the C oracle on the
    first 2,048 channels (one tenth), scaled by 10: the workqueue serially
    (the reference's getWorkqueue is a serial cursor walk), the sampling
    and the fused build on cpu_threads threads."""
    import ctypes as C
    from raptr_amd import geometry as G
    from raptr_amd import _capi as capi
    from raptr_amd import sysmat
    from raptr_amd.device import DeviceArray, Workspace
    setup, grid, nrays = G.config_small()
    m, n = setup.nchannels, grid.nvox
    allc = np.arange(m, dtype=np.int32)
    s, g = setup.struct(), grid.struct()
    L = capi.lib()
    d_list = DeviceArray.from_numpy(allc)
    cnt = DeviceArray(m, np.int64)
    ms_cnt = timed(lambda: capi.call("raptr_workqueue_counts", C.byref(s), C.byref(g),
                                     d_list.ptr, m, cnt.ptr, st))
    nwq = int(cnt.numpy().sum())
    ws = Workspace(L.raptr_workqueue_workspace_size(C.byref(g), m))
    d_cnl, d_vxl = DeviceArray(nwq, np.int32), DeviceArray(nwq, np.int32)
    li, vi, found = C.c_int(0), C.c_int(0), C.c_int64()

    def wq():
        li.value, vi.value = 0, 0
        capi.call("raptr_workqueue", C.byref(s), C.byref(g), d_list.ptr, m, C.byref(li),
                  C.byref(vi), -1, d_cnl.ptr, d_vxl.ptr, nwq, C.byref(found), ws.ptr,
                  ws.nbytes, st)
    wq()
    ms_wq = timed(wq)
    assert found.value == nwq
    d_val = DeviceArray(nwq, np.float32)
    sample = lambda: capi.call("raptr_sm_from_wq", C.byref(s), C.byref(g), d_cnl.ptr,
                               d_vxl.ptr, d_val.ptr, nwq, nrays, 1, capi.RAYGEN_RAND, st)
    sample()
    ms_sm = timed(sample)
    res = {}
    build = lambda: res.__setitem__("A", sysmat.build_csr(setup, grid, allc, nrays, stream=st,
                                                          d_list=d_list))
    build()
    ms_csr = timed(build)
    nnz = res["A"].nnz
    out = {"workload": "20,480-channel heads x 5 angles, 32^3 @ 2.5 mm, 10 rays (all channels)",
           "candidate_tests": m * n, "workqueue_entries": nwq, "nnz": nnz,
           "gpu_ms": {"workqueue_counts": round(ms_cnt, 3), "workqueue": round(ms_wq, 3),
                      "sm_from_wq": round(ms_sm, 3), "csr_build_fused": round(ms_csr, 3)},
           "gpu_rates": {"cull_tests_per_s": round(m * n / (ms_wq * 1e-3) / 1e9, 2),
                         "workqueue_GBps_written": round(8 * nwq / (ms_wq * 1e-3) / 1e9, 1),
                         "ray_evaluations_per_s": round(nwq * nrays / (ms_sm * 1e-3) / 1e9, 2),
                         "unit": "G/s (tests, evaluations), GB/s (bytes)"}}
    log(f"build phases (small): workqueue {ms_wq:.2f} ms ({nwq} entries), sampling "
        f"{ms_sm:.2f} ms, fused CSR build {ms_csr:.2f} ms (nnz {nnz})")
    del d_cnl, d_vxl, d_val, res
    if cpu_threads:
        from oracle import pyoracle as O
        os_, og = O.make_setup(*[getattr(setup, f) for f in (
            "pos0x", "pos1x", "na", "n0z", "n0y", "n1z", "n1y", "da", "segx", "segy", "segz")]), \
            O.make_grid(grid.o, grid.d, grid.n)
        part = allc[:m // 10]
        O.lib().ro_set_threads(cpu_threads)
        t = time.perf_counter()
        ccnt = O.workqueue_counts(os_, og, part)
        t_cnt = time.perf_counter() - t
        t = time.perf_counter()
        ccnl, cvxl, _, _ = O.workqueue(os_, og, part, capacity=int(ccnt.sum()))
        t_wq = time.perf_counter() - t
        t = time.perf_counter()
        O.sm_from_wq(os_, og, ccnl, cvxl, nrays)
        t_sm = time.perf_counter() - t
        t = time.perf_counter()
        O.build_csr(os_, og, part, nrays)
        t_csr = time.perf_counter() - t
        scale = m / len(part)
        out["cpu_ms"] = {"workqueue_counts": round(1e3 * t_cnt * scale, 1),
                         "workqueue": round(1e3 * t_wq * scale, 1),
                         "sm_from_wq": round(1e3 * t_sm * scale, 1),
                         "csr_build_fused": round(1e3 * t_csr * scale, 1)}
        out["cpu"] = {"kind": "port", "threads": {"workqueue": 1, "workqueue_counts": cpu_threads,
                                                  "sm_from_wq": cpu_threads,
                                                  "csr_build_fused": cpu_threads},
                      "sample": f"first {len(part)} channels, scaled by {scale:.0f}"}
        log(f"build phases (small), CPU oracle: workqueue {out['cpu_ms']['workqueue']:.0f} ms "
            f"(1 thread), sampling {out['cpu_ms']['sm_from_wq']:.0f} ms, fused build "
            f"{out['cpu_ms']['csr_build_fused']:.0f} ms ({cpu_threads} threads)")
    return out


def bench_plain_96(log, timed, stream, st, niter, cpu_threads=0):
    """The largest serving geometry as a single-frame plan: 16x16-pixel
    heads, 60 angles at 3 deg, 96^3 grid at 2 mm, 20 rays, 5e6 events, niter
    iterations per frame; the autotuned plan (the fused iteration where it
    wins) against the separate passes, and the image against the FP64
    oracle on the same matrix."""
    from raptr_amd import geometry as G
    from raptr_amd import synthetic, sysmat
    from raptr_amd.device import DeviceArray
    from raptr_amd.recon import Reconstructor
    setup, grid, nrays = G.two_head_setup(16, 60, 3.0), G.centred_grid(96, 2.0), 20
    image, _ = synthetic.random_ellipsoids(grid, n=6, seed=2)
    meas, y, _ = synthetic.measurement(setup, grid, nrays, image, 5e6, seed=1)
    m, n = len(meas), grid.nvox
    d_list = DeviceArray.from_numpy(meas)
    A = sysmat.build_csr(setup, grid, meas, nrays, d_list=d_list, stream=st, capacity="auto")
    sens = sysmat.sensitivity(setup, grid, nrays, stream=st)
    from raptr_amd.recon import default_order_kind
    order = sysmat.locality_order(setup, d_list, stream=st, kind=default_order_kind(1))
    R = Reconstructor(A, sens, "f32", 1, stream=st, row_order=order)
    dy = DeviceArray.from_numpy(y.astype(np.float32))
    x0 = np.full(n, 1.0 / n, np.float32)
    dx = DeviceArray.from_numpy(x0)
    cfg, _ = R.autotune(dy, dx, 1, reps=2, stream=st)

    def frame():
        dx.copy_from(x0, stream=st)
        R.run(dy, dx, niter, 1, use_graph=True, stream=st)

    frame()
    stream.sync()
    ms = timed(frame)
    x = dx.numpy(stream=st)
    sep = None
    if R.fused()["on"]:
        R.set_fused(False)
        frame()
        sep = timed(frame)
    out = {"workload": f"16x16 heads x 60 angles (3,932,160 channels), 96^3 grid @ 2 mm, "
                       f"20 rays, 5e6 events, {niter} iterations, FP32, one frame",
           "measured_channels": m, "nnz": int(A.nnz), "voxels": n,
           "frames_per_s": round(1000.0 / ms, 3), "ms_per_iteration": round(ms / niter, 4),
           "config": cfg,
           "separate_passes_ms_per_iteration": round(sep / niter, 4) if sep else None}
    R.close()
    log(f"plain 96^3: {ms / niter:.3f} ms/iteration (separate passes "
        f"{(sep / niter) if sep else float('nan'):.3f}), cfg {cfg}")
    if cpu_threads:
        from oracle import pyoracle as O
        rp, col, val = A.to_host()
        O.lib().ro_set_threads(cpu_threads)
        x64 = O.mlem_f64(rp, col, val.astype(np.float64), y.astype(np.float64), sens.numpy(),
                         x0.astype(np.float64), niter)
        out["rel_l2_vs_fp64"] = rel_l2(x, x64)
        log(f"plain 96^3: rel L2 vs FP64 {out['rel_l2_vs_fp64']:.2e}")
    return out


def bench_small(log, timed, stream, st, cpu_threads=0):
    """64 dynamic frames on the small geometry (shared matrix of the union of
    measured channels), 30 iterations, FP32, one batch, vs one frame at a time."""
    from raptr_amd import geometry as G
    from raptr_amd import synthetic, sysmat
    from raptr_amd.device import DeviceArray
    from raptr_amd.recon import Reconstructor
    setup, grid, nrays = G.config_small()
    F, niter = 64, 30
    rng = np.random.default_rng(7)
    base, params = synthetic.random_ellipsoids(grid, n=6, seed=3, semi=(3.0, 12.0))
    half = rng.uniform(10, 120, len(params))
    c = grid.centres()
    totals = np.exp(rng.uniform(np.log(1e4), np.log(2e6), F))
    ys = []
    for f in range(F):
        img = np.full(grid.nvox, 0.1)
        for (ctr, ax, a), h in zip(params, half):
            img[np.sum(((c - ctr) / ax) ** 2, axis=1) <= 1.0] += a * 0.5 ** (f / h)
        yb = sysmat.forward_project(setup, grid, img.astype(np.float32), nrays)
        ys.append(synthetic.poisson_counts(yb, totals[f], seed=1000 + f))
    Yall = np.stack(ys, axis=1)
    union = np.nonzero(Yall.sum(axis=1) > 0)[0].astype(np.int32)
    Y = Yall[union].astype(np.float32)
    A = sysmat.build_csr(setup, grid, union, nrays)
    sens = sysmat.sensitivity(setup, grid, nrays)
    n = grid.nvox
    out = {"workload": "64 decaying-ellipsoid frames, 20,480-channel heads x 5 angles, 32^3 @ "
                       "2.5 mm, 10 rays, 1e4-2e6 events/frame, 30 iterations, FP32",
           "union_channels": int(len(union)), "nnz": int(A.nnz)}
    for name, nf in (("plain", 1), ("batched", F)):
        R = Reconstructor(A, sens, "f32", nf, stream=st)
        if nf == 1:
            dys = [DeviceArray.from_numpy(Y[:, f].copy()) for f in range(F)]
            dx = DeviceArray(n, np.float32)
            x0 = np.full(n, 1.0 / n, np.float32)

            def run():
                for f in range(F):
                    dx.copy_from(x0, stream=st)
                    R.run(dys[f], dx, niter, 1, use_graph=True, stream=st)
        else:
            dy = DeviceArray.from_numpy(Y.reshape(-1))
            X0 = np.full((n, F), 1.0 / n, np.float32)
            dx = DeviceArray.from_numpy(X0.reshape(-1))

            def run():
                dx.copy_from(X0, stream=st)
                R.run(dy, dx, niter, F, use_graph=True, stream=st)
        run()
        stream.sync()
        ms = timed(run)
        out[f"{name}_frames_per_s"] = round(F * 1000.0 / ms, 1)
        out[f"{name}_ms_for_64_frames"] = round(ms, 2)
        log(f"small {name}: {ms:.1f} ms for 64 frames")
        if nf > 1:
            Xb = dx.numpy().reshape(n, F)
    out["speedup"] = round(out["batched_frames_per_s"] / out["plain_frames_per_s"], 2)
    if cpu_threads:
        # FP64 oracle on the same matrix: frames 0 and 63 (error of the batched
        # images) and the host time of those two frames, scaled to 64
        from oracle import pyoracle as O
        rp, col, val = A.to_host()
        v64 = val.astype(np.float64)
        O.lib().ro_set_threads(cpu_threads)
        t = time.perf_counter()
        errs = []
        for f in (0, F - 1):
            x64 = O.mlem_f64(rp, col, v64, Y[:, f].astype(np.float64), sens.numpy(),
                             np.full(n, 1.0 / n), niter)
            errs.append(float(np.linalg.norm(Xb[:, f] - x64) / np.linalg.norm(x64)))
        cpu_s = (time.perf_counter() - t) * F / 2
        out["batched_rel_l2_vs_fp64"] = max(errs)
        out["cpu_baseline"] = {"value": round(F / cpu_s, 2), "unit": "frames/s",
                               "cores": cpu_threads, "kind": "port",
                               "sample": "frames 0 and 63 by the FP64 C oracle, scaled to 64"}
        log(f"small: batched rel L2 vs FP64 {max(errs):.2e}; cpu {F / cpu_s:.1f} frames/s")
    return out


if __name__ == "__main__":
    main()
