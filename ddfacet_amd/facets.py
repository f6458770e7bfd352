"""Device-resident facet set: one visibility chunk and all facets of a major cycle on the GPU.

This is the device counterpart of ClassFacetMachine's per-facet job loop
(DDFacet/Imager/ClassFacetMachine.py:1499-1592 gridding, 1859-1943 degridding): every buffer
is uploaded once; ``grid_plain`` / ``degrid_plain`` issue one launch per facet on one stream
(the "plain" path, one facet per launch), the batched path lives in ``cycle.py``.
Degridding sums the model of every facet into the shared visibilities; launches on one stream
are ordered, so no row locks are needed (the reference serialises rows with named semaphores,
degridder.h:236-263).
"""
import ctypes as C

import numpy as np
import torch

from . import abi, marshal
from . import bda as bdamod
from ._lib import check, lib
from .device import require_gpu, stream_handle, to_device


class DeviceFacetSet:
    def __init__(self, P, device=None, grid_dtype=torch.complex64, with_model=True):
        self.P = P
        self.dev = device or require_gpu()
        d = self.dev
        self.nf = len(P.facets)
        t = {}
        t["uvw"] = to_device(np.ascontiguousarray(P.uvw, np.float64), d)
        vis = P.vis
        t["vis"] = vis if isinstance(vis, torch.Tensor) else to_device(np.ascontiguousarray(vis, np.complex64), d)
        t["flags"] = to_device(np.ascontiguousarray(P.flags), d)
        t["weights"] = to_device(np.ascontiguousarray(P.weights, np.float32), d)
        t["freqs"] = to_device(np.ascontiguousarray(P.freqs, np.float64), d)
        t["chan_map"] = to_device(np.int32(P.chan_map), d)
        t["chan_map_degrid"] = to_device(np.int32(P.chan_map_degrid), d)
        t["bda_grid"] = to_device(P.bda_grid, d)
        t["bda_degrid"] = t["bda_grid"] if P.bda_degrid is P.bda_grid else to_device(P.bda_degrid, d)
        t["bda_grid_start"] = to_device(bdamod.block_starts(P.bda_grid), d)
        t["bda_degrid_start"] = (t["bda_grid_start"] if P.bda_degrid is P.bda_grid
                                 else to_device(bdamod.block_starts(P.bda_degrid), d))
        self.t = t
        if P.decorr is not None:
            t["uvw_dt"] = to_device(np.ascontiguousarray(P.decorr[0], np.float64), d)
        if P.sparsification is not None:
            t["sparse"] = to_device(np.ascontiguousarray(P.sparsification, np.bool_), d)
        self.nrow, self.nchan, self.ncorr = [int(x) for x in t["vis"].shape]
        self.fmean, self.wlmean = marshal.freq_means(P.freqs)
        self.equi = int(P.extra.get("equidistant", 1))
        # Jones tables are shared by all facets; only the direction index differs
        j = P.jones
        self.jones = None
        if j is not None:
            jd = {"mode": j.get("mode", 1), "interp": j.get("interp", 0)}
            for k in ("kms", "kms_time", "kms_chan", "beam", "beam_time", "beam_chan", "coefs"):
                if k in j and j[k] is not None and np.asarray(j[k]).size:
                    jd[k] = to_device(np.ascontiguousarray(j[k]), d)
            jd["a0"] = to_device(np.int32(P.a0), d)
            jd["a1"] = to_device(np.int32(P.a1), d)
            for k in ("apply_amp", "apply_phase", "scale", "calib_error"):
                jd[k] = j.get(k, 1 if k.startswith("apply") else 0)
            self.jones = jd
        # per-facet CF tables, grids and accumulators
        self.cf = []
        for f in P.facets:
            pk = marshal.pack_cfs(f.wplanes, f.wplanes_conj, f.winfos)
            # get() passes the lists swapped (WplanesConj first): same data, swapped offsets
            off = pk.offset
            self.cf.append(dict(data=to_device(pk.data, d), offset=to_device(off, d),
                                offset_degrid=to_device(np.concatenate([off[pk.nw:], off[:pk.nw]]), d),
                                support=to_device(pk.support, d), nw=pk.nw, overs=pk.overs,
                                ref_wave=pk.ref_wave, wmax=pk.wmax, max_support=int(pk.support.max())))
        shape = (P.nband, P.npol, P.npix, P.npix)
        self.grids = torch.zeros((self.nf,) + shape, dtype=grid_dtype, device=d)
        self.models = torch.zeros((self.nf,) + shape, dtype=torch.complex64, device=d) if with_model else None
        self.sumwt = torch.zeros((self.nf, P.nband, P.npol), dtype=torch.float64, device=d)
        self.sum_jones = torch.zeros((self.nf, 2, P.nband), dtype=torch.float64, device=d)
        self.sum_jones_chan = torch.zeros((self.nf, 2, self.nchan), dtype=torch.float64, device=d)
        self.nblocks_grid = bdamod.nblocks_of(P.bda_grid)
        self.nblocks_degrid = bdamod.nblocks_of(P.bda_degrid)

    # ---- struct assembly ---------------------------------------------------------------------
    def _desc(self, dopsf=0):
        P = self.P
        g = abi.ddf_grid_desc()
        g.grid_f64 = 1 if self.grids.dtype == torch.complex128 else 0
        g.nband, g.npol, g.ny, g.nx = P.nband, P.npol, P.npix, P.npix
        g.incr[0], g.incr[1] = float(P.incr[0]), float(P.incr[1])
        g.dopsf = int(dopsf)
        return g

    def _vis(self, degrid=False, vis=None):
        P, t = self.P, self.t
        v = abi.ddf_vis()
        v.vis = (vis if vis is not None else t["vis"]).data_ptr()
        v.uvw = t["uvw"].data_ptr()
        v.flags = t["flags"].data_ptr()
        v.weights = t["weights"].data_ptr()
        v.freqs = t["freqs"].data_ptr()
        v.chan_map = t["chan_map_degrid" if degrid else "chan_map"].data_ptr()
        v.nrow, v.nchan, v.ncorr = self.nrow, self.nchan, self.ncorr
        v.chan_equidistant = self.equi
        key = "bda_degrid" if degrid else "bda_grid"
        v.bda = t[key].data_ptr()
        v.bda_len = int(t[key].numel())
        v.nblocks = self.nblocks_degrid if degrid else self.nblocks_grid
        v.bda_start = t[key + "_start"].data_ptr()
        v.sparsification = 0 if (degrid or "sparse" not in t) else t["sparse"].data_ptr()
        v.ncorr_codes = len(P.corr_codes)
        for i, c in enumerate(P.corr_codes):
            v.corr_codes[i] = int(c)
        v.nstokes = len(P.stokes_codes)
        for i, c in enumerate(P.stokes_codes):
            v.stokes_codes[i] = int(c)
        v.freq_mean, v.wavelength_mean = self.fmean, self.wlmean
        return v

    def _facet(self, i, degrid=False):
        P = self.P
        f = abi.ddf_facet()
        f.grid = (self.models[i] if degrid else self.grids[i]).data_ptr()
        f.sumwt = self.sumwt[i].data_ptr()
        cf = self.cf[i]
        f.cf.data = cf["data"].data_ptr()
        f.cf.offset = cf["offset_degrid" if degrid else "offset"].data_ptr()
        f.cf.support = cf["support"].data_ptr()
        f.cf.nw, f.cf.overs = cf["nw"], cf["overs"]
        f.cf.ref_wave, f.cf.wmax = cf["ref_wave"], cf["wmax"]
        f.cf.max_support = cf["max_support"]
        fa = P.facets[i]
        f.cu, f.cv, f.l0, f.m0 = fa.cu, fa.cv, fa.l0, fa.m0
        if P.decorr is not None:
            _, dt, dnu, ts, fs = P.decorr
            f.decorr.enabled = 1
            f.decorr.uvw_dt = self.t["uvw_dt"].data_ptr()
            f.decorr.dt, f.decorr.dnu = float(dt), float(dnu)
            f.decorr.tsmear, f.decorr.fsmear = int(ts), int(fs)
            f.decorr.l0, f.decorr.m0 = fa.l0, fa.m0
        j = self.jones
        if j is not None:
            f.jones.enabled = 1
            f.jones.mode = j["mode"]
            f.jones.interp_mode = j["interp"]
            if "kms" in j:
                f.jones.kms = j["kms"].data_ptr()
                for k in range(4):
                    f.jones.kms_dims[k] = int(j["kms"].shape[k])
                f.jones.kms_time = j["kms_time"].data_ptr()
                f.jones.kms_chan = j["kms_chan"].data_ptr()
                f.jones.kms_dir = fa.kms_dir
            if "beam" in j:
                f.jones.beam = j["beam"].data_ptr()
                for k in range(4):
                    f.jones.beam_dims[k] = int(j["beam"].shape[k])
                f.jones.beam_time = j["beam_time"].data_ptr()
                f.jones.beam_chan = j["beam_chan"].data_ptr()
                f.jones.beam_dir = fa.beam_dir
            f.jones.a0 = j["a0"].data_ptr()
            f.jones.a1 = j["a1"].data_ptr()
            if "coefs" in j:
                f.jones.coefs = j["coefs"].data_ptr()
                f.jones.ncoefs = int(j["coefs"].numel())
            f.jones.apply_amp = int(j["apply_amp"])
            f.jones.apply_phase = int(j["apply_phase"])
            f.jones.scale_jones = int(j["scale"])
            f.jones.calib_error = float(j["calib_error"])
            f.jones.sum_jones = self.sum_jones[i].data_ptr()
            f.jones.sum_jones_chan = self.sum_jones_chan[i].data_ptr()
        return f

    # ---- plain path: one launch per facet ---------------------------------------------------------
    def zero_grids(self):
        self.grids.zero_()
        self.sumwt.zero_()
        self.sum_jones.zero_()
        self.sum_jones_chan.zero_()

    def grid_plain(self, dopsf=0, stream=None, facets=None):
        g = self._desc(dopsf)
        v = self._vis(False)
        sh = C.c_void_p(stream_handle(stream))
        L = lib()
        for i in (range(self.nf) if facets is None else facets):
            f = self._facet(i, False)
            check(L.ddf_grid(C.byref(g), C.byref(v), C.byref(f), sh))

    def degrid_plain(self, vis=None, stream=None, facets=None):
        g = self._desc(0)
        v = self._vis(True, vis)
        sh = C.c_void_p(stream_handle(stream))
        L = lib()
        for i in (range(self.nf) if facets is None else facets):
            f = self._facet(i, True)
            check(L.ddf_degrid(C.byref(g), C.byref(v), C.byref(f), sh))

    # ---- batched path: one plan, a few persistent launches per cycle ----------------------------
    def plan_batched(self, tile=0, chunk=0, blocks_per_item=0, group=0, precision="mixed", workgroups=0,
                     stream=None):
        """Build (or rebuild) the batched plan of this chunk (see ddf_plan_create)."""
        self.destroy_plan()
        o = abi.ddf_plan_opts(tile, chunk, blocks_per_item, group, abi.PRECISION[precision], workgroups)
        g = self._desc(0)
        vg = self._vis(False)
        vd = self._vis(True)
        fa = (abi.ddf_facet * self.nf)(*[self._facet(i, False) for i in range(self.nf)])
        models = (C.c_void_p * self.nf)(*[self.models[i].data_ptr() for i in range(self.nf)])
        out = C.c_void_p()
        check(lib().ddf_plan_create(C.byref(g), C.byref(vg), C.byref(vd), fa, models, self.nf, C.byref(o),
                                    C.byref(out), C.c_void_p(stream_handle(stream))))
        self._plan = out
        return self.plan_info()

    def plan_info(self):
        a = (C.c_int64 * 11)()
        check(lib().ddf_plan_info(self._plan, a, 11))
        keys = ["grid_records", "degrid_records", "grid_items", "degrid_items", "degrid_vis", "device_bytes",
                "tile", "facet_group", "grid_tile_workgroups", "degrid_tile_workgroups", "adjoint_tile_workgroups"]
        return dict(zip(keys, list(a)))

    def placements(self, degrid=False, stream=None):
        """The plan's integer placement maps (ddf_plan_placements) decoded to int32
        [nfacets, nblocks, 8] rows [valid, iw, conj (W <= 0), locx, locy, io, jo, band]."""
        nb = self.nblocks_degrid if degrid else self.nblocks_grid
        raw = torch.empty((nb, self.nf), dtype=torch.int64, device=self.dev)
        check(lib().ddf_plan_placements(self._plan, int(bool(degrid)), C.c_void_p(raw.data_ptr()), raw.numel(),
                                        C.c_void_p(stream_handle(stream))))
        torch.cuda.synchronize(self.dev)
        r = raw.cpu().numpy().view(np.uint64).T
        nw = self.cf[0]["nw"]
        valid = (r >> np.uint64(63)).astype(np.int32)
        plane = ((r >> np.uint64(32)) & np.uint64(0xFF)).astype(np.int32)
        # plane in put order: the degrid map serves W > 0 from the put-order conjugate list
        conj = (plane >= nw) if not degrid else (plane < nw)
        out = np.stack([valid, plane % nw, conj.astype(np.int32),
                        (r & np.uint64(0xFFFF)).astype(np.int32),
                        ((r >> np.uint64(16)) & np.uint64(0xFFFF)).astype(np.int32),
                        ((r >> np.uint64(40)) & np.uint64(0xFF)).astype(np.int32),
                        ((r >> np.uint64(48)) & np.uint64(0xFF)).astype(np.int32),
                        ((r >> np.uint64(56)) & np.uint64(0x7F)).astype(np.int32)], axis=-1)
        out[valid == 0] = 0
        return out

    def plan_timing(self, enable=True):
        """Bracket every batched pass with HIP events on its launch stream (ddf_plan_timing)."""
        check(lib().ddf_plan_timing(self._plan, int(bool(enable))))
        self._timing = bool(enable)

    def capture_cycle(self, stream=None):
        """One major cycle of the batched path (zero the grids, grid, degrid) captured into a HIP
        graph on a side stream; ``graph.replay()`` runs it on the current stream's device.  The
        passes contain no memset nodes (work queues are reset by a kernel, grids by a fill
        kernel) and make no device queries, so the capture is complete.  Pass timing is off
        while capturing."""
        if getattr(self, "_plan", None) is None:
            self.plan_batched(stream=stream)
        timing = getattr(self, "_timing", False)
        if timing:
            self.plan_timing(False)
        cur = torch.cuda.current_stream() if stream is None else stream
        side = torch.cuda.Stream(device=self.dev)
        side.wait_stream(cur)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side):
            self.zero_grids()
            self.grid_batched(stream=side)
            self.degrid_batched(stream=side)
        cur.wait_stream(side)
        if timing:
            self.plan_timing(True)
        return g

    def pass_ms(self):
        """Duration (ms) of the latest run of each batched pass, None for passes not run."""
        a = (C.c_float * len(abi.PASS_NAMES))()
        check(lib().ddf_plan_pass_ms(self._plan, a, len(abi.PASS_NAMES)))
        return {k: (float(x) if x >= 0 else None) for k, x in zip(abi.PASS_NAMES, a)}

    def destroy_plan(self):
        if getattr(self, "_plan", None) is not None:
            lib().ddf_plan_destroy(self._plan)
            self._plan = None

    def grid_batched(self, dopsf=0, stream=None, vis=None):
        if getattr(self, "_plan", None) is None:
            self.plan_batched(stream=stream)
        check(lib().ddf_plan_grid(self._plan, C.c_void_p(0 if vis is None else vis.data_ptr()), int(dopsf),
                                  C.c_void_p(stream_handle(stream))))

    def degrid_batched(self, stream=None, vis=None):
        if getattr(self, "_plan", None) is None:
            self.plan_batched(stream=stream)
        check(lib().ddf_plan_degrid(self._plan, C.c_void_p(0 if vis is None else vis.data_ptr()),
                                    C.c_void_p(stream_handle(stream))))

    def __del__(self):
        try:
            self.destroy_plan()
        except Exception:
            pass
