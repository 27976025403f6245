"""Device-resident MLEM reconstruction plan (Python side of raptr_recon_*).

A plan takes the CSR of the measured channels (from ``sysmat.build_csr``)
and the FP64 sensitivity, builds the forward (row-padded CSR) and backward
(row-blocked transpose) layouts once, and runs plain or batched MLEM
(reco.cu:230-303) on F frames at a time.
"""
import ctypes as C

import numpy as np

from . import _capi as capi
from .device import DeviceArray

VALUE_TYPES = {"f32": 0, "f16": 1, "bf16": 2, "u16": 3}
LAYOUTS = {"blocked": 0}


def default_order_kind(max_frames):
    """Locality order of a plan computed from a setup: y-major for one frame
    (the plain backward's ratio gathers touch fewer lines: 2.72 -> 2.56 ms
    at full size); for batches slope tiles over angles in groups of 4 (16
    consecutive rows = 4 slopes x 4 angles of one midpoint: the 64-frame
    forward's 16-row unions hold 19 % fewer entries than with the
    angle-major slope tiles, 9.1 -> 7.9 ms at full size)."""
    return "y" if max_frames == 1 else "angle_slope"


def _row_order(row_order, A, stream, kind):
    if row_order is None:
        return None
    if isinstance(row_order, DeviceArray):
        return row_order
    if isinstance(row_order, tuple):
        from . import sysmat
        return sysmat.locality_order(row_order[0], row_order[1], stream=stream, kind=kind)
    if hasattr(row_order, "nchannels"):
        from . import sysmat
        return sysmat.locality_order(row_order, A.channels, stream=stream, kind=kind)
    return DeviceArray.from_numpy(np.ascontiguousarray(row_order, np.int32), stream=stream)


class Reconstructor:
    """layout: "blocked" (row-padded CSR + row-blocked transpose, any frame
    count) or "auto" (= blocked).  block_rows: row-block height."""

    def __init__(self, A, sens, value_type="f32", max_frames=1, block_rows=0, stream=None,
                 layout="auto", row_order=None, grid=None, group=0, grouped_only=False,
                 order_kind=None):
        """row_order: None (the CSR's row order), a device or host int32
        permutation (plan row p = CSR row row_order[p]), or a
        ``MeasurementSetup`` with the CSR's channel ids given as
        ``A.channels`` / a (setup, channels) pair, from which the locality
        order (sysmat.locality_order) is computed.  Counts, images and the
        forward's ratios keep the CSR's row order at the interface.
        grid (VoxelGrid) / group (16, 32; 0 = 16): the 64-frame grouped
        layouts' group size, with backward groups as bricks of the grid.
        grouped_only (max_frames 64): keep only the grouped layouts (64-frame
        passes only, ~40 % less device memory).  order_kind: the locality
        order computed from a setup ("angle", "y", "slope"; None: y for one
        frame, slope for batches)."""
        self.m, self.n, self.nnz = A.m, A.n, A.nnz
        self.value_type = value_type
        self.max_frames = max_frames
        if layout == "auto":
            layout = "blocked"
        self.layout = layout
        order = _row_order(row_order, A, stream, order_kind or default_order_kind(max_frames))
        p = C.c_void_p()
        self._grid = grid.struct() if grid is not None else None
        opts = capi.raptr_recon_options(LAYOUTS[layout], order.ptr if order is not None else None,
                                        C.pointer(self._grid) if self._grid is not None else None,
                                        int(group), 1 if grouped_only else 0)
        self.grouped_only = bool(grouped_only)
        capi.call("raptr_recon_create_opts", C.byref(p), A.m, A.n, A.nnz, A.row_ptr.ptr,
                  A.col.ptr, A.val.ptr, sens.ptr, VALUE_TYPES[value_type], int(max_frames),
                  int(block_rows), C.byref(opts), stream)
        self.handle = p.value
        self.row_ordered = bool(capi.lib().raptr_recon_row_ordered(self.handle))

    def info(self):
        b, p = C.c_int64(), C.c_int64()
        nb, fl, fu, bu = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        capi.call("raptr_recon_info", self.handle, C.byref(b), C.byref(p), C.byref(nb),
                  C.byref(fl), C.byref(fu), C.byref(bu))
        st, lpc, wgs = C.c_int(), C.c_int(), C.c_int()
        capi.call("raptr_recon_get_backward", self.handle, C.byref(st), C.byref(lpc),
                  C.byref(wgs))
        out = {"layout": self.layout, "device_bytes": b.value, "padded_nnz": p.value,
               "row_blocks": nb.value, "fwd_lanes": fl.value, "fwd_unroll": fu.value,
               "bwd_unroll": bu.value, "bwd_staged": st.value,
               "bwd_lanes_per_column": lpc.value, "bwd_workgroups": wgs.value,
               "bwd_blocks_per_launch": capi.lib().raptr_recon_get_blocks_per_launch(self.handle),
               "transpose_nnz": capi.lib().raptr_recon_transpose_nnz(self.handle),
               "row_ordered": bool(capi.lib().raptr_recon_row_ordered(self.handle)),
               "compact_rows": capi.lib().raptr_recon_get_compact_rows(self.handle),
               "fwd_kind": self.forward_kind()[0]}
        fz = self.fused()
        out.update({"fused": fz["on"], "fused_available": fz["available"],
                    "fused_waves": fz["waves"], "fused_lanes_per_row": fz["lanes_per_row"],
                    "fused_lanes_per_column": fz["lanes_per_column"],
                    "fused_reduce_per_xcd": fz["per_xcd"], "fused_steps": fz["steps"],
                    "union_rows": self.unions()["rows"]})
        gi = self.grouped()
        if gi["group"]:
            out.update({"grouped64": gi["on"], "group": gi["group"],
                        "grouped_unroll": gi["unroll"], "grouped_kind": gi["kind"],
                        "grouped_fwd_entries": gi["fwd_entries"],
                        "grouped_bwd_entries": gi["bwd_entries"]})
        return out

    def set_broadcast(self, on):
        capi.call("raptr_recon_set_broadcast", self.handle, 1 if on else 0)

    # grouped kernel shapes per group size: (unroll, kind); kind 2 matrix
    # cores, 3 matrix cores with per-XCD row ranges in the backward
    GROUP_SHAPES = {16: ((8, 3), (16, 3), (32, 3), (8, 2), (16, 2)),
                    32: ((16, 3), (8, 3))}

    def grouped(self):
        """dict of the grouped 64-frame layouts: group size (0: none), on,
        unroll, kind, forward / backward union entries."""
        v = [C.c_int() for _ in range(4)] + [C.c_int64(), C.c_int64()]
        capi.call("raptr_recon_get_grouped", self.handle, *[C.byref(t) for t in v])
        return dict(zip(("group", "on", "unroll", "kind", "fwd_entries", "bwd_entries"),
                        [t.value for t in v]))

    def set_grouped(self, on, unroll=None, kind=None):
        """unroll / kind None: the group size's first shape."""
        g = self.grouped()["group"]
        shapes = self.GROUP_SHAPES.get(g, ((8, 3),))
        if unroll is None or kind is None:
            unroll, kind = [(u, k) for u, k in shapes
                            if (unroll is None or u == unroll) and (kind is None or k == kind)][0]
        capi.call("raptr_recon_set_grouped", self.handle, 1 if on else 0, int(unroll), int(kind))

    def set_compact_rows(self, on):
        """16-bit transpose rows (per-XCD single-frame backward) and forward
        columns, where the plan has them."""
        capi.call("raptr_recon_set_compact_rows", self.handle, 1 if on else 0)

    def set_forward_kind(self, kind):
        """Single-frame forward kernel: 0 one entry per lane, 2 x staged in
        LDS over the column union of each group of 64 plan rows (both
        bit-identical)."""
        capi.call("raptr_recon_set_forward_kind", self.handle, int(kind))

    def forward_kind(self):
        """(kind, available kinds)."""
        k, a = C.c_int(), C.c_int()
        capi.call("raptr_recon_get_forward_kind", self.handle, C.byref(k), C.byref(a), None)
        return k.value, [j for j in range(3) if a.value >> j & 1]

    def set_fused(self, on, waves=None, lanes_per_row=None, lanes_per_column=None):
        """Single-frame iterations in one pass over the matrix (forward,
        ratio and backward per group of 64 plan rows, then a fixed-order
        column reduce).  on: False / 0 off; True / 1 one column-sum region
        per wave; 2 one region per workgroup, the waves adding in turn.
        Waves per group 2 or 4 (or 8 with a shared region), lanes per row
        16, 32 or 64, reduce lanes per column 16, 32 or 64 (None: keep the
        plan's)."""
        mode = (1 if on else 0) if isinstance(on, bool) else int(on)
        cur = self.fused()
        waves = cur["waves"] if waves is None else waves
        lanes_per_row = cur["lanes_per_row"] if lanes_per_row is None else lanes_per_row
        lanes_per_column = cur["lanes_per_column"] if lanes_per_column is None else lanes_per_column
        capi.call("raptr_recon_set_fused", self.handle, mode, int(waves),
                  int(lanes_per_row), int(lanes_per_column))

    def set_fused_steps(self, steps):
        """Register-held steps of 256 entries per row of the fused products
        kernel with 64 lanes per row (2, 3 or 4)."""
        capi.call("raptr_recon_set_fused_steps", self.handle, int(steps))

    def set_fused_reduce(self, mode):
        """Fused iteration's column reduce: 0 (or False) all of a column's
        group partials in one pass; 1 (or True, the default) per-XCD group
        ranges; 2 column tiles over the partials in written order."""
        capi.call("raptr_recon_set_fused_reduce", self.handle, int(mode))

    def set_union_rows(self, rows=0, stream=None):
        """Plan rows per column-union group (0: the plan's own choice, else a
        power of two in 16..512); rebuilds the unions and the fused pass's
        index (before any clone)."""
        capi.call("raptr_recon_set_union_rows", self.handle, int(rows), stream)

    def unions(self):
        """{"rows", "largest", "entries"} of the column unions (zeros: none)."""
        r, l, e = C.c_int(), C.c_int(), C.c_int64()
        capi.call("raptr_recon_get_unions", self.handle, C.byref(r), C.byref(l), C.byref(e))
        return {"rows": r.value, "largest": l.value, "entries": e.value}

    def set_fused_tiles(self, columns, segments_per_workgroup=None, stream=None):
        """Shape of the column-tile reduce: columns per tile (power of two,
        128-2048) and segments per workgroup (None: keep the plan's);
        rebuilds its segment table (before any clone)."""
        spw = self.fused_tiles()["segments_per_workgroup"] if segments_per_workgroup is None \
            else segments_per_workgroup
        capi.call("raptr_recon_set_fused_tiles", self.handle, int(columns), int(spw), stream)

    def fused_tiles(self):
        """{"columns", "segments_per_workgroup", "segments", "workgroups",
        "max_parts"} of the column-tile reduce's table (zeros: none)."""
        c, p, n, w, k = C.c_int(), C.c_int(), C.c_int64(), C.c_int(), C.c_int()
        capi.call("raptr_recon_get_fused_tiles", self.handle, C.byref(c), C.byref(p),
                  C.byref(n), C.byref(w), C.byref(k))
        return {"columns": c.value, "segments_per_workgroup": p.value, "segments": n.value,
                "workgroups": w.value, "max_parts": k.value}

    def fused(self):
        """{"available", "on", "waves", "lanes_per_row", "lanes_per_column",
        "per_xcd", "steps"} of the fused pass."""
        a, o, w, r, l = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        capi.call("raptr_recon_get_fused", self.handle, C.byref(a), C.byref(o), C.byref(w),
                  C.byref(r), C.byref(l))
        return {"available": a.value, "on": o.value, "waves": w.value,
                "lanes_per_row": r.value, "lanes_per_column": l.value,
                "per_xcd": capi.lib().raptr_recon_get_fused_reduce(self.handle),
                "steps": capi.lib().raptr_recon_get_fused_steps(self.handle)}

    def forward_union_entries(self):
        """Total size of the LDS-staged forward's column unions (0: none)."""
        u = C.c_int64()
        capi.call("raptr_recon_get_forward_kind", self.handle, None, None, C.byref(u))
        return u.value

    def set_xcd_blocks(self, blocks):
        """Kind 3 backward: row blocks each XCD walks per launch (0: all)."""
        capi.call("raptr_recon_set_xcd_blocks", self.handle, int(blocks))

    def set_blocks_per_launch(self, blocks):
        capi.call("raptr_recon_set_blocks_per_launch", self.handle, int(blocks))

    def set_config(self, fwd_lanes, fwd_unroll, bwd_unroll):
        capi.call("raptr_recon_set_config", self.handle, int(fwd_lanes), int(fwd_unroll),
                  int(bwd_unroll))

    def set_backward(self, staged, lanes_per_column=16, workgroups=512):
        """Single-frame backward: 3 (per-XCD row ranges; True means 3) or
        block by block through L2 (0 / False)."""
        mode = (3 if staged else 0) if isinstance(staged, bool) else int(staged)
        capi.call("raptr_recon_set_backward", self.handle, mode, int(lanes_per_column),
                  int(workgroups))

    def run(self, y, x, niter, nframes=1, use_graph=True, stream=None, active=None):
        """In-place MLEM on device arrays y ([m][F]) and x ([n][F]).
        active (64-frame grouped plans: 16 or 32): iterate only the first
        `active` frames of each 64-frame row (raptr_recon_mlem_active)."""
        if active is None or active == nframes:
            capi.call("raptr_recon_mlem", self.handle, int(nframes), y.ptr, x.ptr, int(niter),
                      1 if use_graph else 0, stream)
        else:
            capi.call("raptr_recon_mlem_active", self.handle, int(nframes), int(active), y.ptr,
                      x.ptr, int(niter), 1 if use_graph else 0, stream)

    def mlem(self, y, x0, niter, use_graph=True, stream=None):
        """Host convenience: y [m] or [m, F]; x0 [n] or [n, F]; returns x."""
        y = np.asarray(y, np.float32)
        x0 = np.asarray(x0, np.float32)
        F = 1 if y.ndim == 1 else y.shape[1]
        # every copy on the work stream: a non-blocking stream is not ordered
        # with the null stream
        dy = DeviceArray.from_numpy(np.ascontiguousarray(y).reshape(-1), stream=stream)
        dx = DeviceArray.from_numpy(np.ascontiguousarray(x0).reshape(-1), stream=stream)
        self.run(dy, dx, niter, F, use_graph=use_graph and stream is not None, stream=stream)
        out = dx.numpy(stream=stream)
        return out if F == 1 else out.reshape(self.n, F)

    def autotune(self, y, x, nframes=1, reps=3, stream=None):
        """Time the kernel shapes on this plan with HIP events and keep the
        fastest: forward shapes first (lanes per row, unroll) with the
        current backward, then backward shapes (unroll; for one frame also
        the per-XCD pass's lanes per column).
        Returns (config dict, ms per iteration)."""
        from .device import Event
        e0, e1 = Event(), Event()
        x_save = DeviceArray(x.count, np.float32)
        capi.call("raptr_memcpy_d2d", x_save.ptr, x.ptr, x.nbytes, stream)

        def timed():
            self.run(y, x, 1, nframes, use_graph=False, stream=stream)
            e0.record(stream)
            self.run(y, x, reps, nframes, use_graph=False, stream=stream)
            e1.record(stream)
            return e0.elapsed_ms(e1) / reps

        inf = self.info()
        bu = inf["bwd_unroll"]
        grp = nframes == 64 and inf.get("group", 0) > 0
        if self.grouped_only:
            (g_u, g_p, bpl, xb), ms_best = self._tune_grouped(inf, timed)
            capi.call("raptr_memcpy_d2d", x.ptr, x_save.ptr, x.nbytes, stream)
            return ({"grouped64": 1, "grouped_unroll": g_u, "grouped_kind": g_p,
                     "bwd_blocks_per_launch": bpl, "xcd_blocks": xb, "grouped_only": 1},
                    ms_best)
        if grp:  # tune the general kernels first
            self.set_grouped(False)
        fz0 = self.fused() if nframes == 1 else {"available": 0}
        if fz0["available"]:  # the separate passes first
            self.set_fused(False)
        if nframes == 1:
            fshapes = [(fl, fu) for fl in (16, 32, 64) for fu in (1, 2, 4)]
        else:
            fshapes = [(inf["fwd_lanes"], fu) for fu in (1, 2, 4)]
        # single frame: every forward kind the plan has
        kinds = self.forward_kind()[1] if nframes == 1 else [None]
        best_f = None
        for fk in kinds:
            if fk is not None:
                self.set_forward_kind(fk)
            for fl, fu in fshapes:
                self.set_config(fl, fu, bu)
                ms = timed()
                if best_f is None or ms < best_f[1]:
                    best_f = ((fl, fu, fk), ms)
        fl, fu, fp = best_f[0]
        if fp is not None:
            self.set_forward_kind(fp)
        bshapes = [(0, 16, 512, u, bpl) for u in (1, 2, 4) for bpl in (1, 4, 16)
                   if bpl == 1 or bpl < inf["row_blocks"]]
        staged_ok = nframes == 1 and self.max_frames == 1
        if staged_ok:
            try:  # per-XCD row ranges need the plan's partial column sums
                self.set_backward(3, 32)
            except capi.RaptrError:
                staged_ok = False
        if staged_ok:
            bshapes += [(3, lpc, 512, u, 1) for lpc in (16, 32, 64) for u in (1, 2, 4)]
        best = None
        for st, lpc, wgs, u, bpl in bshapes:
            if staged_ok:
                self.set_backward(st, lpc, wgs)
            self.set_config(fl, fu, u)
            self.set_blocks_per_launch(bpl)
            ms = timed()
            if best is None or ms < best[1]:
                best = ((st, lpc, wgs, u, bpl), ms)
        st, lpc, wgs, u, bpl = best[0]
        if staged_ok:
            self.set_backward(st, lpc, wgs)
        self.set_config(fl, fu, u)
        self.set_blocks_per_launch(bpl)
        fz_cfg = (0, 0, 0, 0, 0, 0, 0, 0, 0)
        if fz0["available"]:
            fz_cfg = self._tune_fused(fz0, timed, stream, y, x)
        bc = 0
        if nframes == 64:
            # the 64-frame kernels with in-row broadcast of values and indices
            base = timed()
            self.set_broadcast(True)
            if timed() < base:
                bc = 1
            else:
                self.set_broadcast(False)
        ms_best = timed()
        g_on, g_u, g_p, xb = 0, 0, 0, 0
        if grp:
            best_g = self._tune_grouped(inf, timed)
            if best_g[1] < ms_best:
                (g_u, g_p, bpl, xb), ms_best = best_g
                g_on = 1
            else:
                self.set_grouped(False)
                self.set_xcd_blocks(0)
            self.set_blocks_per_launch(bpl)
        capi.call("raptr_memcpy_d2d", x.ptr, x_save.ptr, x.nbytes, stream)
        cfg = {"fwd_lanes": fl, "fwd_unroll": fu, "fwd_kind": self.forward_kind()[0],
               "bwd_unroll": u, "bwd_staged": st,
               "bwd_lanes_per_column": lpc, "bwd_workgroups": wgs, "bwd_blocks_per_launch": bpl,
               "broadcast64": bc, "grouped64": g_on, "grouped_unroll": g_u,
               "grouped_kind": g_p, "xcd_blocks": xb, "fused": fz_cfg[0],
               "fused_waves": fz_cfg[1], "fused_lanes_per_row": fz_cfg[2],
               "fused_lanes_per_column": fz_cfg[3], "fused_reduce_per_xcd": fz_cfg[4],
               "fused_steps": fz_cfg[5], "union_rows": fz_cfg[6] or self.unions()["rows"],
               "fused_tiles": fz_cfg[7], "fused_segments_per_workgroup": fz_cfg[8]}
        return cfg, ms_best

    def _tune_fused(self, fz0, timed, stream, y, x):
        """The fused iteration's shape: plan rows per union group (the
        plan's choice, half and twice it), then per group size the
        products kernel (4 or 8 waves, the plan's register-held steps and
        one more), then the column reduce (column tiles of 512-2,048
        columns with 32 or 64 lanes per slot, the per-XCD and the by-column
        passes).  Kept if faster than the separate passes, which are timed
        last over the kept unions.  The reduces are timed alone (their
        share of an iteration is within its timing noise).  Returns (mode,
        waves, lanes per row,
        lanes per column, reduce mode, steps, union rows, tile columns,
        segments per workgroup), zeros if the separate passes win.  Each
        shape keeps the better of two timings, and a later shape replaces
        an earlier one only if 1 % faster (the plan's defaults come first),
        so that timing noise does not pick among equals."""
        def t2():
            return min(timed(), timed())

        from .device import Event
        r0, r1 = Event(), Event()

        def t_reduce(reps=20):
            capi.call("raptr_recon_fused_pass", self.handle, x.ptr, y.ptr, 1, stream)
            best = None
            for _ in range(2):
                r0.record(stream)
                for _ in range(reps):
                    capi.call("raptr_recon_fused_pass", self.handle, x.ptr, y.ptr, 2, stream)
                r1.record(stream)
                ms = r0.elapsed_ms(r1) / reps
                best = ms if best is None else min(best, ms)
            return best

        rows0 = self.unions()["rows"]
        red0 = (fz0["lanes_per_column"], fz0["per_xcd"])  # the plan's own reduce
        best = None
        for rows in [rows0] + [r for r in (rows0 // 2, rows0 * 2) if 16 <= r <= 512]:
            if rows != self.unions()["rows"]:
                try:
                    self.set_union_rows(rows, stream=stream)
                except capi.RaptrError:
                    continue
            fz = self.fused()
            if not fz["available"]:
                continue
            s0 = fz["steps"]
            shapes = [(2, w, 64, st_) for w in (fz["waves"], 12 - fz["waves"])
                      for st_ in sorted({s0, min(s0 + 1, 4)})]
            if rows == rows0:
                shapes += [(2, 2, 64, s0), (2, 4, 32, s0), (1, 2, 64, s0)]
            for fmode, fw, flr, fst in shapes:
                try:
                    self.set_fused(fmode, fw, flr, red0[0])
                    self.set_fused_reduce(red0[1])
                    self.set_fused_steps(fst)
                except capi.RaptrError:  # too little LDS for that many waves
                    continue
                ms = t2()
                if best is None or ms < 0.99 * best[1]:
                    best = ((fmode, fw, flr, fst, rows), ms)
        if best is None:
            self.set_fused(False)
            return (0,) * 9
        fmode, fw, flr, fst, rows = best[0]
        if rows != self.unions()["rows"]:
            self.set_union_rows(rows, stream=stream)
        self.set_fused_steps(fst)
        reds = [(16, 1, None), (64, 1, None), (64, 0, None)]
        if self.fused_tiles()["columns"]:
            reds = [(lpc, 2, t) for t in ((1024, 512), (512, 1024), (2048, 512))
                    for lpc in (64, 32)] + reds
        first = (red0[0], red0[1], (1024, 512) if red0[1] == 2 else None)
        reds = [first] + [r for r in reds if r != first]
        best_r = None
        for lpc, xr, tiles in reds:
            if tiles is not None:
                try:
                    self.set_fused_tiles(*tiles, stream=stream)
                except capi.RaptrError:  # a table shared with clones stays
                    continue
            self.set_fused(fmode, fw, flr, lpc)
            self.set_fused_reduce(xr)
            ms = t_reduce()
            if best_r is None or ms < 0.99 * best_r[1]:
                best_r = ((lpc, xr, tiles), ms)
        lpc, xr, tiles = best_r[0]
        if tiles is not None:
            self.set_fused_tiles(*tiles, stream=stream)
        self.set_fused(fmode, fw, flr, lpc)
        self.set_fused_reduce(xr)
        ms_fused = t2()
        self.set_fused(False)
        if ms_fused >= t2():
            return (0,) * 9
        self.set_fused(fmode, fw, flr, lpc)
        self.set_fused_reduce(xr)
        tz = self.fused_tiles()
        return (fmode, fw, flr, lpc, xr, fst, rows, tz["columns"], tz["segments_per_workgroup"])

    # rows each XCD walks per launch of kind 3's backward: windows of about
    # 131,072 plan rows (32 MB of 64-frame ratios) measured best at full size
    # with 16,384- and 32,768-row blocks and 4 x 1 x 4 voxel bricks; tried
    # with half and twice that window and in one launch
    XCD_WINDOW_ROWS = 131072

    def _xcd_blocks_options(self, inf):
        rows = max(1, -(-self.m // max(1, inf["row_blocks"])))
        t = max(1, self.XCD_WINDOW_ROWS // rows)
        return sorted({0, t, max(1, t // 2), 2 * t})

    def _tune_grouped(self, inf, timed):
        """Grouped 64-frame shapes: union entries per step and kind, with
        row blocks per launch (kinds 0-2) or per XCD and launch (kind 3);
        leaves the fastest set.  Returns ((unroll, kind, bpl, xcd_blocks),
        ms)."""
        best = None
        for gu, gp in self.GROUP_SHAPES[inf["group"]]:
            opts = ([(1, xb) for xb in self._xcd_blocks_options(inf)] if gp >= 3 else
                    [(gb, 0) for gb in (1, 4, 16)])
            for gb, xb in opts:
                if gb > 1 and gb >= inf["row_blocks"]:
                    continue
                self.set_grouped(True, gu, gp)
                self.set_blocks_per_launch(gb)
                self.set_xcd_blocks(xb)
                ms = timed()
                if best is None or ms < best[1]:
                    best = ((gu, gp, gb, xb), ms)
        (gu, gp, gb, xb), _ = best
        self.set_grouped(True, gu, gp)
        self.set_blocks_per_launch(gb)
        self.set_xcd_blocks(xb)
        return best

    def clone(self, stream=None):
        """A second plan over the same matrix (raptr_recon_clone): it shares
        this plan's layouts and owns only its work vectors, with this plan's
        current kernel shapes.  Run it on its own stream beside this plan
        (two batches at once).  It keeps this plan alive; close it first."""
        h = C.c_void_p()
        capi.call("raptr_recon_clone", self.handle, C.byref(h), stream)
        c = object.__new__(type(self))
        c.__dict__.update({k: v for k, v in self.__dict__.items() if k != "handle"})
        c.handle = h.value
        c._parent = self
        return c

    def clone_count(self):
        return capi.lib().raptr_recon_clone_count(self.handle)

    def close(self):
        """Destroy the plan (raises while clones of it live)."""
        if getattr(self, "handle", None):
            capi.call("raptr_recon_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
