"""Throughput of one grid + degrid major cycle over all facets (vis/s), 81-facet MeerKAT-like.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode batched|plain] [--no-cpu]

A step = grid the chunk's visibilities onto all 81 facet grids, then degrid all 81 model grids
and subtract them from the visibilities (vis -= sum over facets).  Inputs are resident in HBM
before timing starts.  value = Nvis / t_step with Nvis = rows x channels (4-correlation
visibilities presented, flagged ones included).

Roofline (HBM-bound scatter/gather).  `roofline` is for the dominant pass (the batched pass with
the largest mean duration, timed live with HIP events on its launch stream through
ddf_plan_timing): achieved = the pass's share of the per-facet algorithm's bytes (per
visibility-facet pair 40 + 64 + 48 / C: grid pass A carries 40 + 24 / C per pair, degrid pass B
64 + 24 / C, the tile passes the per-facet grid and CF bytes) / its mean duration, frac =
achieved / 8 TB/s (`passes_frac` for every pass).  `unique_frac` prices the same pass on the
bytes it must move once (a visibility read once for all facets):
  grid pass A    per vis 40 (vis 32, flags 4, weight 4), per row 24 (uvw), per BDA word 4,
                 per record 8 (placement) + 8 per Stokes (block sum written), Jones tables once
  grid pass B    per record 12 (placement 8, index 4) + 8 per Stokes (block sum), per facet the CF
                 table once and the grid read + written once (2 x grid bytes)
  degrid pass A  per record 12 + 8 per Stokes (written), per facet the CF table and the model grid
  degrid pass B  per vis 64 (read + write) + 12 (vis list), per row 24, per record 8 per Stokes,
                 Jones tables once.
`roofline.cycle` is the cycle-level figure of the per-facet algorithm: per visibility-facet pair
40 + 64 + 48 / C bytes, plus per facet 3 grids, the BDA maps and the CF table, over t_step; a
design that reads a visibility once for all facets can exceed 1 there.
The CPU baseline is the reference gridder (oracle/_ref, built from the reference sources) on the
host cores (at most 16 worker processes, the CPU share of one GPU, over all facet jobs) on a
stride sample of the same BDA blocks, scaled to the full cycle; it runs before the GPU is
initialised.  Reported beside t_step, not in it: the plan build (`plan`: ms and device bytes,
once per chunk), the image side (`image_side_ms`: facet FFTs, stitch, model cut) and the same
cycle in the FP64 precision (`fp64_ms_per_step`, with `rel_err_vs_fp64`).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cf_bytes(P):
    f = P.facets[0]
    return sum(p.size * 8 for p in f.wplanes) + sum(p.size * 8 for p in f.wplanes_conj)


def _jones_bytes(P):
    if P.jones is None:
        return 0
    return int(np.asarray(P.jones["kms"]).size) * 8


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "batched_cycle_81facet_hbm_traffic.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the 81-facet cycle, from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE scaled by a copy-kernel calibration); None when absent."""
    try:
        text = open(TRAFFIC_FILE).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        d = json.loads(block)
        return int(d["bytes_per_launch"][kernel]["traffic_bytes"])
    except (OSError, IndexError, KeyError, ValueError):
        return None


def cycle_bytes(P, nf):
    """Bytes of one major cycle of the per-facet algorithm (each facet a pass over the
    visibilities): per visibility-facet pair 40 (grid in) + 64 (degrid read + write) + 48 / C
    (uvw in both passes); per facet 3 grids (zero/write after gridding, model read), the BDA maps
    and the CF table."""
    R, C = P.uvw.shape[0], P.freqs.shape[0]
    nvis = R * C
    gbytes = P.nband * P.npol * P.npix * P.npix * 8
    per_facet = 3 * gbytes + 4 * (P.bda_grid.size + P.bda_degrid.size) + _cf_bytes(P)
    return nf * (nvis * (104 + 48.0 / C) + per_facet)


def pass_bytes(P, nf, info, ns):
    """Algorithmic bytes of one launch of each batched pass (see the module docstring)."""
    R, C = P.uvw.shape[0], P.freqs.shape[0]
    nvis = R * C
    gbytes = P.nband * P.npol * P.npix * P.npix * 8
    ng, nd = info["grid_records"], info["degrid_records"]
    cf = _cf_bytes(P)
    jb = _jones_bytes(P)
    return {
        "grid_pass_a": nvis * 40 + R * 24 + 4 * P.bda_grid.size + ng * (8 + 8 * ns) + jb,
        "grid_pass_b": ng * (12 + 8 * ns) + nf * (cf + 2 * gbytes),
        "degrid_pass_a": nd * (12 + 8 * ns) + nf * (cf + gbytes),
        "degrid_pass_b": info["degrid_vis"] * 76 + R * 24 + nd * 8 * ns + jb,
    }


def survey_pass_bytes(P, nf):
    """The per-facet algorithm's bytes (per visibility-facet pair 40 + 64 + 48 / C, per facet 3
    grids and the CF table) split over the batched passes: grid pass A carries the gridding side
    of every pair (40 + 24 / C), degrid pass B the degridding side (64 + 24 / C), the tile passes
    the per-facet terms (grid B: grid zero + write and the CF table, degrid A: model read and the
    CF table)."""
    R, C = P.uvw.shape[0], P.freqs.shape[0]
    pairs = R * C * nf
    gbytes = P.nband * P.npol * P.npix * P.npix * 8
    cf = _cf_bytes(P)
    return {
        "grid_pass_a": pairs * (40 + 24.0 / C),
        "grid_pass_b": nf * (2 * gbytes + cf),
        "degrid_pass_a": nf * (gbytes + cf),
        "degrid_pass_b": pairs * (64 + 24.0 / C),
    }


def plain_bytes(P, kind):
    """Algorithmic bytes of one per-facet plain launch (grid_plain / degrid_plain)."""
    R, C = P.uvw.shape[0], P.freqs.shape[0]
    nvis = R * C
    gbytes = P.nband * P.npol * P.npix * P.npix * 8
    if kind == "grid":
        return nvis * (40 + 24.0 / C) + 2 * gbytes + 4 * P.bda_grid.size + _cf_bytes(P)
    return nvis * (64 + 24.0 / C) + gbytes + 4 * P.bda_degrid.size + _cf_bytes(P)


_CPU = {}


def _cpu_facet_job(fi):
    """Grid then degrid the sample onto facet `fi` with the reference module (forked worker)."""
    mod, S = _CPU["mod"], _CPU["S"]
    g = S.new_grid()
    sw = np.zeros((S.nband, S.npol))
    mod.pyGridderWPol32(*S.grid_args(fi, g, sw))
    v = S.vis.copy()
    mod.pyDeGridderWPol(*S.degrid_args(fi, g, v))
    return fi


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(P, target_s=12.0, max_workers=16):
    """The reference gridder (oracle/_ref, built from DDFacet/Gridder) on the host cores, before
    the GPU is touched: one facet job at a time per worker process (the reference's process pool
    runs one facet job per core with one OpenMP thread, AsyncProcessPool.py:831-861), every facet
    of the cycle gridded then degridded over the same stride sample of the BDA blocks; 3373 row
    semaphores as the reference deploys.  Workers: the affinity cores, at most `max_workers` (the
    CPU share of one GPU on the box).  value = vis-facet pairs per second over all workers / the
    number of facets."""
    import multiprocessing as mp
    from ddfacet_amd import bda as bdamod
    from ddfacet_amd.problems import Problem
    from oracle import refmod
    if not refmod.available():
        return None
    mod = refmod.module(nsem=3373)
    affinity = len(os.sched_getaffinity(0))
    cores = max(1, min(max_workers, affinity))
    nb = bdamod.nblocks_of(P.bda_grid)
    starts = bdamod.block_starts(P.bda_grid)
    sizes = P.bda_grid[2:2 + nb]
    rng = np.random.default_rng(5)

    def sample(stride):
        sel = np.arange(0, nb, stride)
        rows = np.unique(np.concatenate([P.bda_grid[starts[b] + 2:starts[b] + sizes[b]] for b in sel]))
        remap = -np.ones(P.uvw.shape[0], np.int64)
        remap[rows] = np.arange(rows.size)
        recs, szs = [], []
        for b in sel:
            r = P.bda_grid[starts[b]:starts[b] + sizes[b]].astype(np.int64).copy()
            r[2:] = remap[r[2:]]
            recs.append(r)
            szs.append(sizes[b])
        m = np.concatenate([np.array([sel.size, 0], np.int64), np.array(szs, np.int64)] + recs).astype(np.int32)
        nvis_s = int(sum((r[1] - r[0]) * (r.size - 2) for r in recs))
        # values do not change the work: unit-variance noise in place of the sky
        shape = (rows.size, P.freqs.size, 4)
        vs = (rng.standard_normal(shape, dtype=np.float32) + 1j * rng.standard_normal(shape, dtype=np.float32))
        j = None
        if P.jones is not None:
            j = dict(P.jones)
            j["kms_time"] = P.jones["kms_time"][rows]
        S = Problem(uvw=np.ascontiguousarray(P.uvw[rows]), vis=np.ascontiguousarray(vs.astype(np.complex64)),
                    flags=np.ascontiguousarray(P.flags[rows]), weights=np.ascontiguousarray(P.weights[rows]),
                    freqs=P.freqs, a0=np.int32(P.a0[rows]), a1=np.int32(P.a1[rows]), chan_map=P.chan_map,
                    chan_map_degrid=P.chan_map_degrid, bda_grid=m, bda_degrid=m, facets=P.facets,
                    npix=P.npix, nband=P.nband, npol=P.npol, incr=P.incr, corr_codes=P.corr_codes,
                    stokes_codes=P.stokes_codes, jones=j)
        return S, nvis_s

    nf = len(P.facets)
    # calibrate the stride on one facet in this process, then every facet over the workers
    _CPU.update(mod=mod)
    S, nv = sample(400)
    _CPU["S"] = S
    t0 = time.perf_counter()
    _cpu_facet_job(nf // 2)
    t1 = time.perf_counter() - t0
    stride = int(max(1, min(4000, round(400 * nf * t1 / (cores * target_s)))))
    S, nv = sample(stride)
    _CPU["S"] = S
    facets = list(range(nf))
    ctx = mp.get_context("fork")
    with ctx.Pool(cores) as pool:
        t0 = time.perf_counter()
        pool.map(_cpu_facet_job, facets, chunksize=1)
        t = time.perf_counter() - t0
    _CPU.clear()
    vf_per_s = nv * nf / t
    return dict(value=vf_per_s / nf, unit="vis/s", cores=cores, kind="reference", cpu_model=cpu_model(),
                affinity_cores=affinity,
                sample="reference _pyGridderSmearPols3x (oracle/_ref, built from DDFacet/Gridder sources), "
                       "%d worker processes (of %d affinity cores; one GPU's CPU share is 16) running all %d "
                       "facet jobs, each gridding then degridding every %d-th BDA block (%d vis), %.1f s wall; "
                       "value = vis-facet pairs/s / %d facets" % (cores, affinity, nf, stride, nv, t, nf))


def error_vs_fp64(fs, vis0, tile, group, blocks_per_item, stream, steps=3):
    """Relative L2 distance of the timed (mixed-precision) cycle's facet grids and predicted
    visibilities to the same cycle run by an FP64-precision plan of the same chunk, and that
    plan's cycle time (mean of `steps` cycles after one warm-up)."""
    import torch

    def cycle():
        fs.zero_grids()
        fs.t["vis"].copy_(vis0)
        fs.grid_batched(stream=stream)
        # degridding subtracts the prediction: from zeros, so that the FP32 rounding of the
        # observed visibilities does not enter the comparison
        fs.t["vis"].zero_()
        fs.degrid_batched(stream=stream)
        torch.cuda.synchronize()
        out = (fs.grids.clone(), (-fs.t["vis"]).clone())
        fs.t["vis"].copy_(vis0)
        return out

    def rel(x, y):
        return float(torch.linalg.vector_norm((x - y).to(torch.complex128)) /
                     torch.linalg.vector_norm(y.to(torch.complex128)))

    g_mixed, p_mixed = cycle()
    fs.plan_batched(tile=tile, group=group, blocks_per_item=blocks_per_item, precision="fp64", stream=stream)
    g64, p64 = cycle()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fs.zero_grids()
        fs.grid_batched(stream=stream)
        fs.degrid_batched(stream=stream)
    torch.cuda.synchronize()
    ms64 = (time.perf_counter() - t0) * 1e3 / steps
    fs.t["vis"].copy_(vis0)
    fs.destroy_plan()
    return {"grids": rel(g_mixed, g64), "predicted_vis": rel(p_mixed, p64), "fp64_ms_per_step": ms64,
            "note": "the timed cycle's arithmetic (mixed precision) against the FP64-precision plan"}


def device_bda_map(P, steps):
    """Time the chunk's BDA map built on the device (ddf_bda_build_*, the mode-1 builder with the
    workload's GridDecorr 0.02 and facet field of view), inputs resident in HBM.  Returns the
    timing and the map (host copy)."""
    import torch
    from ddfacet_amd import bda as bdamod
    cell = P.extra["cell_arcsec"] * np.pi / 180 / 3600
    dfreq = float(P.freqs[1] - P.freqs[0])
    t_uvw = torch.from_numpy(np.ascontiguousarray(P.uvw)).cuda()
    t_a0 = torch.from_numpy(np.int32(P.a0)).cuda()
    t_a1 = torch.from_numpy(np.int32(P.a1)).cuda()
    t_cm = torch.from_numpy(np.int32(P.chan_map)).cuda()
    args = (t_uvw, t_a0, t_a1, P.freqs, dfreq, 0.02, cell, P.extra["facet_npix"], t_cm)
    m = bdamod.bda_map_device(*args)
    torch.cuda.synchronize()
    ts = []
    for _ in range(max(1, steps)):
        t0 = time.perf_counter()
        bdamod.bda_map_device(*args)
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    return {"ms": float(np.median(ts)), "blocks": bdamod.nblocks_of(m.cpu().numpy()[:2]), "words": int(m.numel()),
            "map": m.cpu().numpy()}


def image_side_ms(P, fs, steps):
    """Mean ms of the image side of the cycle, reported apart from t_step (ClassFacetMachine.py:
    1331-1416 stitch, ClassDDEGridMachine.py:1186-1190 GridToIm, ClassImToGrid.py:138-264 +
    get()'s FFT): GridToIm of every facet grid (batched rocFFT + shifts, x OverS^2), the stitch
    onto the 4095^2 main image, and the model cut + forward FFT of every facet."""
    import torch
    from ddfacet_amd import imaging
    from ddfacet_amd.imaging import FacetGeometry
    nfx = P.extra["nfx"]
    fn = P.extra["facet_npix"]
    geom = FacetGeometry(nfx * fn, nfx, fn, P.npix, P.extra["cell_arcsec"] * np.pi / 180 / 3600)
    dev = fs.grids.device
    sphe = torch.from_numpy(np.stack([np.float32(f.sphe) for f in P.facets])).to(dev)
    inv = torch.from_numpy(np.stack([np.float32(1.0 / np.float64(f.sphe)) for f in P.facets])).to(dev)
    sw = torch.from_numpy(geom.spatial_weights()).to(dev)
    norm = torch.zeros((geom.m, geom.m), dtype=torch.float32, device=dev)
    imaging.facet_norm(geom, sw, norm)
    image = torch.zeros((P.nband, P.npol, geom.m, geom.m), dtype=torch.float32, device=dev)
    overs = int(P.facets[0].winfos[3])
    sj = fs.sum_jones if fs.jones is not None else None
    parts = {"grid_to_image": lambda: imaging.grid_to_image(fs.grids, overs),
             "stitch": lambda: imaging.stitch(geom, sphe, inv, sw, fs.grids, fs.sumwt, sj, norm, image),
             "model_to_grid": lambda: imaging.model_to_grid(geom, sphe, sw, image, norm, fs.models, overs)}
    keep = fs.models.clone()
    out = {}
    for k, fn_ in parts.items():
        fn_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn_()
        torch.cuda.synchronize()
        out[k] = (time.perf_counter() - t0) * 1e3 / steps
    out["total"] = sum(out.values())
    fs.models.copy_(keep)
    fs.zero_grids()
    return out


def time_graph(fs, steps):
    """Mean ms of one batched cycle replayed from a HIP graph (capture excluded)."""
    import torch
    g = fs.capture_cycle()
    g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def run_uneven(a):
    """121 facets x 4 bands, baseline-dependent averaging: plain per-facet launches against the
    autotuned batched cycle over the same major cycle (grid all facets, degrid all facets)."""
    import torch
    from ddfacet_amd import autotune, configs
    from ddfacet_amd.facets import DeviceFacetSet
    P = configs.cfg4(log=log)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    fs = DeviceFacetSet(P, device=dev)
    gen = torch.Generator(device=dev).manual_seed(7)
    fs.models.copy_(torch.randn(fs.models.shape, generator=gen, device=dev, dtype=torch.complex64) * 1e-3)
    vis0 = fs.t["vis"].clone()
    s = torch.cuda.current_stream()
    R, C = P.uvw.shape[0], P.freqs.shape[0]

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        fs.t["vis"].copy_(vis0)
        return (time.perf_counter() - t0) * 1e3 / steps

    def plain():
        fs.zero_grids()
        fs.grid_plain(stream=s)
        fs.degrid_plain(stream=s)

    def batched():
        fs.zero_grids()
        fs.grid_batched(stream=s)
        fs.degrid_batched(stream=s)

    ms_plain = timed(plain, a.steps, a.warmup)
    log("plain: %.2f ms per cycle" % ms_plain)
    space = {"tile": (16, 32, 64), "chunk": (1024, 4096), "blocks_per_item": (32, 128)}
    best, results = autotune.autotune(fs, space=space, reps=2, log=log)
    ms_batched = timed(batched, a.steps, a.warmup)
    ms_graph = time_graph(fs, a.steps)
    fs.t["vis"].copy_(vis0)
    info = fs.plan_info()
    out = {
        "metric": "vis/sec grid+degrid, 121-facet uneven BDA, 4 bands",
        "value": R * C / (ms_batched * 1e-3), "unit": "vis/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_batched, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (fp32 grids, fp64 phase base)", "data": "synthetic",
        "config": {"workload": "121 facets (11 x 11, 567 -> 975 px at 2\"), 4 bands, 27 antennas 20 km disk, "
                               "100 x 10 s, 16 ch, baseline-dependent averaging 1-16 by length quintile",
                   "rows": R, "channels": C, "nvis": R * C, "facets": len(P.facets),
                   "rows_per_factor_1_2_4_8_16": P.extra["rows_per_factor"],
                   "grid_records": info["grid_records"], "grid_items": info["grid_items"]},
        "plain_ms_per_step": ms_plain, "plain_value": R * C / (ms_plain * 1e-3),
        "speedup_vs_plain": ms_plain / ms_batched, "graph_ms_per_step": ms_graph,
        "autotune": {"best": best, "candidates": [[c, ms] for c, ms in results]},
    }
    print(json.dumps(out))


def run_single(a):
    """The plain single-facet workload (561 600 visibilities, one 512^2 grid): the reference
    gridder (oracle/_ref) and the oracle on one host core, the plain per-facet kernels behind the
    drop-in shim (inputs resident in HBM) and the batched plan, each timed over grid + degrid."""
    import torch
    from ddfacet_amd import configs
    from ddfacet_amd.facets import DeviceFacetSet
    from oracle import cpuref, refmod
    P = configs.cfg1(log=log)
    R, C = P.uvw.shape[0], P.freqs.shape[0]
    nvis = R * C
    rng = np.random.default_rng(8)
    model = (rng.standard_normal((1, 1, P.npix, P.npix)) * 1e-3).astype(np.complex64)
    cpu = {}
    if not a.no_cpu:
        impls = [("oracle", cpuref.grid, cpuref.degrid)]
        if refmod.available():
            m = refmod.module()
            impls.insert(0, ("reference", m.pyGridderWPol32, m.pyDeGridderWPol))
        for label, gf, df in impls:
            ts = []
            for _ in range(3):
                g, sw = P.new_grid(), np.zeros((1, 1))
                v = P.vis.copy()
                t0 = time.perf_counter()
                gf(*P.grid_args(0, g, sw))
                df(*P.degrid_args(0, model, v))
                ts.append(time.perf_counter() - t0)
            cpu[label] = {"value": nvis / float(np.median(ts)), "unit": "vis/s", "cores": 1,
                          "s_per_cycle": float(np.median(ts)), "cpu_model": cpu_model()}
        log("cpu: %s" % cpu)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    fs = DeviceFacetSet(P, device=dev)
    fs.models[0].copy_(torch.from_numpy(model[0]))
    vis0 = fs.t["vis"].clone()

    def timed(fn):
        for _ in range(a.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            fn()
        torch.cuda.synchronize()
        fs.t["vis"].copy_(vis0)
        return (time.perf_counter() - t0) * 1e3 / a.steps

    def plain():
        fs.zero_grids()
        fs.grid_plain()
        fs.degrid_plain()

    def batched():
        fs.zero_grids()
        fs.grid_batched()
        fs.degrid_batched()

    ms_plain = timed(plain)
    fs.plan_batched()
    ms_batched = timed(batched)
    out = {
        "metric": "vis/sec grid+degrid, single 512^2 facet",
        "value": nvis / (ms_batched * 1e-3), "unit": "vis/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_batched, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (fp32 grids, fp64 phase base)", "data": "synthetic",
        "config": {"workload": "single 512^2 facet at 2\", 27 antennas in a 20 km disk, 100 x 10 s, 16 ch "
                               "1.2-1.6 GHz, 8 w-planes, S = 7, OverS = 11, one BDA block per visibility",
                   "rows": R, "channels": C, "nvis": nvis},
        "plain_ms_per_step": ms_plain, "plain_value": nvis / (ms_plain * 1e-3),
        "cpu_one_core": cpu,
    }
    fs.destroy_plan()
    print(json.dumps(out))


def run_bda(a):
    """The 81-facet workload chunk's BDA map (mode 1, GridDecorr 0.02, facet field of view) built
    on the device from uvw / antennas resident in HBM; CPU baseline: the host restatement of the
    reference builder (oracle/bda.py) over the whole chunk on one core."""
    import torch
    from ddfacet_amd import bda as bdamod
    from ddfacet_amd import configs, synth
    freqs = np.linspace(0.9e9, 1.67e9, 64)
    obs = synth.make_observation("meerkat", 64, 600, 8.0, freqs, -30.0, -30.7, 2000, sort_by_baseline=True)
    cell = 1.5 * synth.ARCSEC
    nopt, _ = configs.estimate_npix(4096 / 9, 1.7)
    chan_map = np.zeros(64, np.int32)
    R = obs.nrow
    cpu, host = None, None
    if not a.no_cpu:
        from oracle import bda as obda
        t0 = time.perf_counter()
        host = obda.bda_map_mode1(obs.uvw, obs.a0, obs.a1, freqs, obs.dfreq, 0.02, cell, nopt, chan_map)
        dt = time.perf_counter() - t0
        cpu = {"value": R / dt, "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": "host restatement of GiveBlocksRowsListBL_old + collectSmearMapping (oracle/bda.py) on "
                         "the whole chunk (%d rows), %.1f s" % (R, dt)}
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    args = (torch.from_numpy(np.ascontiguousarray(obs.uvw)).cuda(), torch.from_numpy(np.int32(obs.a0)).cuda(),
            torch.from_numpy(np.int32(obs.a1)).cuda(), freqs, obs.dfreq, 0.02, cell, nopt,
            torch.from_numpy(chan_map).cuda())
    for _ in range(a.warmup):
        m = bdamod.bda_map_device(*args)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        m = bdamod.bda_map_device(*args)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / a.steps
    mh = m.cpu().numpy()
    out = {
        "metric": "rows/sec BDA block map build, 81-facet MeerKAT-like chunk",
        "value": R / (ms * 1e-3), "unit": "rows/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 builder arithmetic, int32 map", "data": "synthetic",
        "config": {"workload": "BDA map (mode 1, GridDecorr 0.02, facet FoV) of the 81-facet MeerKAT-like chunk: "
                               "64 antennas, 600 x 8 s, 64 ch", "rows": R, "blocks": bdamod.nblocks_of(mh),
                   "words": int(mh.size)},
        "equal_to_host_builder": None if host is None else bool(np.array_equal(mh, host)),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def run_pipeline(a, nchunks=4):
    """Major cycles over a stream of chunks whose visibilities arrive from (pinned) host memory:
    the upload of chunk n + 1 (copy stream) overlapped with the grid + degrid cycle of chunk n
    (compute stream), as the reference loads the next chunk while the imager works
    (ClassDeconvMachine.py:1320-1337).  Chunks: 300 timesteps of the 81-facet MeerKAT-like
    workload each (the metadata, BDA maps and plan of one such chunk, the visibilities of each
    chunk its own); timed serial (upload, then cycle) and pipelined."""
    import torch
    from ddfacet_amd import configs
    from ddfacet_amd.chunk import ChunkPipeline
    from ddfacet_amd.facets import DeviceFacetSet
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    P = configs.cfg2(with_vis=False, log=log, time_frac=0.5)
    R, C = P.uvw.shape[0], P.freqs.shape[0]
    P.vis = torch.zeros((R, C, 4), dtype=torch.complex64, device=dev)
    fs = DeviceFacetSet(P, device=dev)
    gen = torch.Generator(device=dev).manual_seed(9)
    fs.models.copy_(torch.randn(fs.models.shape, generator=gen, device=dev, dtype=torch.complex64) * 1e-3)
    fs.plan_batched()
    rng = np.random.default_rng(10)
    host = []
    for k in range(nchunks):
        v = torch.empty((R, C, 4), dtype=torch.complex64).pin_memory()
        v.view(torch.float32).copy_(torch.from_numpy(rng.standard_normal((R, C, 8), dtype=np.float32)))
        host.append({"vis": v})
    nbytes = host[0]["vis"].numel() * 8
    compute = torch.cuda.current_stream()

    def cycle(vis):
        fs.zero_grids()
        fs.grid_batched(vis=vis, stream=compute)
        fs.degrid_batched(vis=vis, stream=compute)

    def serial():
        for k in range(nchunks):
            fs.t["vis"].copy_(host[k]["vis"], non_blocking=True)
            cycle(fs.t["vis"])

    pipe = ChunkPipeline(dev)

    def pipelined():
        pipe.submit(0, host[0])
        for k in range(nchunks):
            if k + 1 < nchunks:
                pipe.submit(k + 1, host[k + 1])
            d = pipe.wait(k, compute)
            cycle(d["vis"])
            pipe.release(k, compute)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / a.steps

    # the parts alone
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        fs.t["vis"].copy_(host[0]["vis"], non_blocking=True)
    torch.cuda.synchronize()
    ms_copy = (time.perf_counter() - t0) * 1e3 / a.steps
    ms_cycle = timed(lambda: cycle(fs.t["vis"]))
    ms_serial = timed(serial)
    ms_pipe = timed(pipelined)
    nvis = R * C * nchunks
    out = {
        "metric": "vis/sec grid+degrid over a stream of chunks uploaded from host memory, 81 facets",
        "value": nvis / (ms_pipe * 1e-3), "unit": "vis/s", "n_gpus": 1, "steps": a.steps, "warmup": 1,
        "ms_per_step": ms_pipe, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (fp32 grids, fp64 phase base)", "data": "synthetic",
        "config": {"workload": "%d chunks of 300 x 8 s of the 81-facet MeerKAT-like workload (64 antennas, 64 ch, "
                               "81 facets), visibilities in pinned host memory" % nchunks,
                   "rows_per_chunk": R, "channels": C, "chunks": nchunks, "upload_bytes_per_chunk": nbytes},
        "upload_ms_per_chunk": ms_copy, "upload_GBps": nbytes / (ms_copy * 1e-3) / 1e9,
        "cycle_ms_per_chunk": ms_cycle, "serial_ms_per_step": ms_serial, "pipelined_ms_per_step": ms_pipe,
        "hidden_fraction_of_uploads": (ms_serial - ms_pipe) / (nchunks * ms_copy),
    }
    fs.destroy_plan()
    print(json.dumps(out))


def run_autograd(a, gd_steps=20):
    """chi^2 gradient descent through the differentiable degridder (forward: model cube ->
    visibilities, backward: the hand-written adjoint), 25 facets x 4 bands."""
    import torch
    from ddfacet_amd import configs
    from ddfacet_amd.autograd import DegridOperator, chi2
    from ddfacet_amd.cycle import MajorCycle
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    P, geom = configs.cfg5(device="cuda", log=log)
    cyc = MajorCycle(P, geom)
    op = DegridOperator(cyc)
    data = P.vis
    w = torch.from_numpy(P.weights).to(dev)
    shape = (P.nband, P.npol, geom.m, geom.m)
    gen = torch.Generator(device=dev).manual_seed(5003)
    # adjoint check on random vectors
    xr = torch.randn(shape, generator=gen, device=dev)
    yr = torch.randn(data.shape, generator=gen, device=dev, dtype=torch.complex64)
    axr = op.forward(xr).to(torch.complex128)
    lhs = torch.sum(axr.conj() * yr.to(torch.complex128)).real
    rhs = torch.sum(xr.double() * op.adjoint(yr).double())
    # relative to the Cauchy-Schwarz scale |A x| |y| (the inner product of random vectors is far below it)
    cs = torch.linalg.vector_norm(axr) * torch.linalg.vector_norm(yr.to(torch.complex128))
    dot_rel = float(abs(lhs - rhs) / max(float(cs), 1e-30))
    # step 1 / L, L = 2 lambda_max(A^H W A) by power iteration
    y = xr / xr.norm()
    lam = 0.0
    for _ in range(8):
        z = op.adjoint(w[..., None] * op.forward(y))
        lam = float(z.norm())
        y = z / lam
    alpha = 1.0 / (2.0 * lam)
    x = torch.zeros(shape, device=dev, requires_grad=True)

    def gd_step():
        loss = chi2(x, op, data, w)
        loss.backward()
        with torch.no_grad():
            x.sub_(alpha * x.grad)
        x.grad = None
        return loss

    first = float(gd_step().detach())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(gd_steps):
        last = gd_step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / gd_steps
    last = float(last.detach())
    R, C = P.uvw.shape[0], P.freqs.shape[0]
    out = {
        "metric": "chi^2 gradient-descent steps through the differentiable degridder, 25 facets x 4 bands",
        "value": R * C / (ms * 1e-3), "unit": "vis/s (one forward + one adjoint per step)", "n_gpus": 1,
        "steps": gd_steps, "warmup": 1, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 (fp32 grids, fp64 phase base)", "data": "synthetic",
        "config": {"workload": "25 facets (5 x 5, 405 -> 693 px at 2\"), model cube 4 x 2025 x 2025, 27 antennas "
                               "20 km disk, 300 x 10 s, 32 ch in 4 bands, 100 sources + noise 0.05",
                   "rows": R, "channels": C, "nvis": R * C},
        "chi2_first": first, "chi2_after": last, "dot_product_rel_mismatch": dot_rel, "step_size": alpha,
    }
    print(json.dumps(out))


def run_flags(a):
    """Flag preparation of the 81-facet workload's chunk on the device (autocorrelations, NaNs, 1e9
    fill, equalisation over the 4 correlations), chunk resident in HBM.  CPU baseline: the numpy
    restatement (oracle/flagprep.py) on every 8th row."""
    import torch
    from ddfacet_amd import configs
    from ddfacet_amd.flagprep import prepare_flags_device
    from oracle.flagprep import prepare_flags
    P = configs.cfg2(with_vis=False, with_jones=False, log=log)
    R, Cn, K = P.flags.shape
    rng = np.random.default_rng(77)
    cpu = None
    if not a.no_cpu:
        sl = slice(0, R, 8)
        d = (rng.standard_normal((P.flags[sl].shape)) + 1j * rng.standard_normal(P.flags[sl].shape)).astype(np.complex64)
        f = P.flags[sl].copy()
        t0 = time.perf_counter()
        prepare_flags(f, P.a0[sl], P.a1[sl], d)
        dt = time.perf_counter() - t0
        cpu = {"value": f.shape[0] * Cn / dt, "unit": "vis/s", "cores": 1, "kind": "port",
               "sample": "numpy restatement (oracle/flagprep.py) on every 8th row (%d visibilities), %.1f s"
                         % (f.shape[0] * Cn, dt)}
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    flags = torch.from_numpy(P.flags.astype(np.uint8)).to(dev)
    gen = torch.Generator(device=dev).manual_seed(78)
    data = torch.randn((R, Cn, K), dtype=torch.complex64, device=dev, generator=gen)
    ta0 = torch.from_numpy(P.a0.astype(np.int32)).to(dev)
    ta1 = torch.from_numpy(P.a1.astype(np.int32)).to(dev)
    for _ in range(a.warmup):
        prepare_flags_device(flags, ta0, ta1, data)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        prepare_flags_device(flags, ta0, ta1, data)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / a.steps
    nvis = R * Cn
    # algorithmic bytes: flags read and written (1 byte per correlation each), data read (8 per
    # correlation), antennas (8 per row); the 1e9 writes of flagged values are not counted
    nbytes = nvis * K * (1 + 1 + 8) + 8 * R
    ach = nbytes / (ms * 1e-3) / 1e9
    res = {
        "metric": "vis/sec chunk flag preparation, 81-facet MeerKAT-like chunk",
        "value": nvis / (ms * 1e-3), "unit": "vis/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8 flags, c64 data",
        "data": "synthetic",
        "config": {"workload": "flag preparation of the 81-facet MeerKAT-like chunk (64 antennas, 600 x 8 s, 64 ch, "
                               "4 corr)", "rows": R, "channels": Cn, "nvis": nvis},
        "roofline": {"bound": "hbm", "kernel": "prepare_flags_k", "achieved": ach, "peak": 8000.0, "unit": "GB/s",
                     "frac": ach / 8000.0, "traffic": None, "algorithmic_bytes_per_launch": nbytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


def run_beams(a):
    """Beam Jones producers on the device: the GMRT band-5 analytic E-Jones of 81 facet directions
    (9 x 9 over 1.8 deg) for 64 antennas, 64 channels 1.05-1.45 GHz into 64 Jones channels, beam
    intervals of 1 min over 600 x 8 s (80 intervals), CenterNorm on (ddf_beam_jones), then merged with
    a DD-solution table of 10-min intervals and 8 channels (ddf_jones_merge), tables in HBM.
    CPU baseline: the numpy restatement (oracle/beams.py) on the first 4 beam intervals."""
    import torch
    from ddfacet_amd import beams
    from oracle import beams as ob
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    na, nchan = 64, 64
    freqs = np.linspace(1.05e9, 1.45e9, nchan)
    width = freqs[1] - freqs[0]
    ra_c, dec_c = 1.2, -0.5
    off = np.deg2rad(np.linspace(-0.9, 0.9, 9))
    ra = (ra_c + off[:, None] / np.cos(dec_c) + 0 * off[None, :]).ravel()
    dec = (dec_c + 0 * off[:, None] + off[None, :]).ravel()
    times = np.arange(0.0, 600 * 8.0, 8.0)
    tb = beams.beam_sample_times(times, 1)
    t0d = np.arange(0.0, 600 * 8.0, 600.0)
    dd = {"Jones": None, "t0": t0d, "t1": t0d + 600.0, "tm": t0d + 300.0,
          "FreqDomains": beams.freq_domains(freqs, width, 8)}
    gen = torch.Generator(device=dev).manual_seed(81)
    dd["Jones"] = torch.randn((t0d.size, ra.size, na, 8, 2, 2), dtype=torch.complex64, device=dev, generator=gen)
    for _ in range(a.warmup):
        beam = beams.estimate_beam(tb, ra, dec, ra_c, dec_c, freqs, width, na, kind="GMRT", nband=nchan)
        merged = beams.merge_jones(dd, beam)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        beam = beams.estimate_beam(tb, ra, dec, ra_c, dec_c, freqs, width, na, kind="GMRT", nband=nchan)
        merged = beams.merge_jones(dd, beam)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / a.steps
    nb = beam["Jones"].numel() // 4
    nm = merged["Jones"].numel() // 4
    cpu = None
    if not a.no_cpu:
        tcpu = time.perf_counter()
        ob.estimate_beam(tb[:5], ra, dec, ra_c, dec_c, freqs, width, nchan, na, "GMRT", center_norm=True)
        dt = time.perf_counter() - tcpu
        cpu = {"value": (4 * ra.size * na * nchan) / dt, "unit": "Jones matrices/s", "cores": 1, "kind": "port",
               "sample": "numpy restatement (oracle/beams.py) of EstimateBeam on the first 4 beam intervals "
                         "(%d matrices), %.2f s; the domain merge not included" % (4 * ra.size * na * nchan, dt)}
    # algorithmic bytes: the beam table written once, both merge inputs read and the merged table
    # written (complex64 2 x 2 matrices, 32 bytes)
    nbytes = 32 * (nb + nb + dd["Jones"].numel() // 4 + nm)
    res = {
        "metric": "Jones matrices/s, beam E-Jones evaluation + merge with DD solutions",
        "value": (nb + nm) / (ms * 1e-3), "unit": "Jones matrices/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 beam model, c64 tables", "data": "synthetic",
        "config": {"workload": "GMRT band-5 beam of 81 directions x 64 antennas x 64 channels, 80 one-minute "
                               "intervals, CenterNorm, merged with 10-min x 8-channel DD solutions",
                   "beam_matrices": nb, "merged_matrices": nm},
        "roofline": {"bound": "hbm", "kernel": "beam_write_k + jones_merge_k", "achieved": nbytes / (ms * 1e-3) / 1e9,
                     "peak": 8000.0, "unit": "GB/s", "frac": nbytes / (ms * 1e-3) / 8e12, "traffic": None,
                     "algorithmic_bytes_per_launch": nbytes,
                     "note": "end-to-end host calls (domain logic, model upload, synchronisation) included"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


def run_weights(a):
    """Briggs imaging weights (robust 0) of the 81-facet workload's chunk on the device: uv-grid
    index, scatter-add, robust transform, renormalisation, inputs resident in HBM.  CPU baseline:
    the numpy oracle of the same steps on the whole chunk."""
    import torch
    from ddfacet_amd import configs
    from ddfacet_amd import weights as dw
    P = configs.cfg2(with_vis=False, with_jones=False, log=log)
    R, Cn = P.uvw.shape[0], P.freqs.shape[0]
    w = np.where(P.flags[..., 0], np.float32(0), P.weights).astype(np.float32)
    # weight grid of a 4096-pixel image at 1.5" (Super 1): cell = 1 / FOV in wavelengths
    fov = P.extra["cell_arcsec"] * np.pi / 180 / 3600 * P.extra["image_npix"]
    cell = 1.0 / fov
    uvmax = dw.uvmax_wavelengths(P.uvw, P.freqs)
    xymax, npixx, npix = dw.grid_geometry(uvmax, cell)
    cpu = None
    if not a.no_cpu:
        from oracle import weights as ow
        sl = slice(0, R)
        t0 = time.perf_counter()
        ow.imaging_weights(P.uvw[sl], w[sl], P.freqs, cell, xymax, npixx, npix, weighting="briggs", robust=0.0)
        dt = time.perf_counter() - t0
        nsamp = w[sl].size
        cpu = {"value": nsamp / dt, "unit": "vis/s", "cores": 1, "kind": "port",
               "sample": "numpy restatement (oracle/weights.py) on the whole chunk (%d weights), %.1f s" % (nsamp, dt)}
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    tu = torch.from_numpy(np.ascontiguousarray(P.uvw)).to(dev)
    tw = torch.from_numpy(w).to(dev)
    run = lambda: dw.imaging_weights(tu, tw, P.freqs, cell, uvmax, weighting="briggs", robust=0.0)
    for _ in range(a.warmup):
        run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        out = run()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / a.steps
    # algorithmic bytes: the weight copy (8 per visibility), index (uvw per row; weight in, index
    # out), scatter-add (weight, index; an FP64 read-modify-write per cell), robust transform (three
    # passes over the grid), renormalisation (weight, index, cell gather, weight out)
    nvis = R * Cn
    nbytes = 24 * R + nvis * (8 + 8 + 8 + 20) + npix * (16 + 3 * 8)
    ach = nbytes / (ms * 1e-3) / 1e9
    res = {
        "metric": "vis/sec Briggs imaging weights, 81-facet MeerKAT-like chunk",
        "value": nvis / (ms * 1e-3), "unit": "vis/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64 grid, f32 weights",
        "data": "synthetic",
        "config": {"workload": "imaging weights of the 81-facet MeerKAT-like chunk (64 antennas, 600 x 8 s, 64 ch), "
                               "Briggs robust 0, weight grid %d x %d" % (npixx, xymax + 1),
                   "rows": R, "channels": Cn, "nvis": nvis, "npix": npix},
        "roofline": {"bound": "hbm", "kernel": "whole pipeline", "achieved": ach, "peak": 8000.0, "unit": "GB/s",
                     "frac": ach / 8000.0, "traffic": None, "algorithmic_bytes_per_launch": nbytes},
        "cpu_baseline": cpu,
        "finite_weights": int(torch.isfinite(out).sum()),
    }
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="batched", choices=["batched", "plain"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--facets", type=int, default=0, help="use only the first N facets (debug)")
    ap.add_argument("--workload", default="meerkat81",
                    choices=["meerkat81", "single512", "uneven121", "autograd25", "weights", "flags", "bda",
                             "pipeline", "beams"],
                    help="meerkat81: the 81-facet MeerKAT-like cycle (the headline line); single512: the "
                         "plain single-facet workload, CPU and GPU plain / batched; uneven121: "
                         "121 facets x 4 bands with baseline-dependent averaging, plain per-facet "
                         "launches against the autotuned batched cycle (eager and graph-replayed); "
                         "autograd25: 20 chi^2 gradient-descent steps through the differentiable "
                         "degridder on 25 facets x 4 bands; weights: Briggs imaging weights of the "
                         "81-facet workload's chunk; flags: flag preparation of that chunk; bda: the "
                         "chunk's BDA block map built on the device; pipeline: chunks uploaded from host memory "
                         "on a copy stream while the previous chunk's cycle runs; beams: analytic beam "
                         "E-Jones tables and their merge with DD solutions")
    ap.add_argument("--tile", type=int, default=0, help="uv tile edge of the batched plan (0: the plan default)")
    ap.add_argument("--blocks-per-item", type=int, default=0,
                    help="BDA blocks per pass-A work item of the batched plan (0: the plan default)")
    ap.add_argument("--group", type=int, default=0,
                    help="facets per pass-A lane group of the batched plan (0: the plan default)")
    ap.add_argument("--graph", action="store_true",
                    help="also time the batched cycle replayed from a HIP graph")
    ap.add_argument("--no-fp64-check", action="store_true",
                    help="skip the FP64-precision cycle that measures the error (and the time) against the timed one")
    ap.add_argument("--no-image", action="store_true",
                    help="skip the separately reported image side (facet FFTs, stitch, model cut)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only an elementwise copy of 1 GiB (reads 1 GiB, writes 1 GiB), to scale "
                         "the FETCH_SIZE / WRITE_SIZE counters of a profiler run")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch
        x = torch.ones(2 ** 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            torch.mul(x, 1.0, out=y)  # an elementwise kernel: 1 GiB read, 1 GiB written per launch
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4, "launches": 3,
                          "kernel": "elementwise mul by 1 (3 launches after the fill of x)"}))
        return
    if a.gpus != 1:
        print("bench.py runs on one GPU: the library has no multi-GPU path (the work would shard by "
              "visibility rows, with one all-reduce of the facet grids); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if a.workload == "uneven121":
        return run_uneven(a)
    if a.workload == "single512":
        return run_single(a)
    if a.workload == "autograd25":
        return run_autograd(a)
    if a.workload == "weights":
        return run_weights(a)
    if a.workload == "flags":
        return run_flags(a)
    if a.workload == "bda":
        return run_bda(a)
    if a.workload == "pipeline":
        return run_pipeline(a)
    if a.workload == "beams":
        return run_beams(a)
    from ddfacet_amd import configs
    t0 = time.time()
    # geometry, maps, CFs and Jones; the CPU baseline runs before the GPU is touched, so with it
    # the BDA map comes from the host restatement of the builder (the same map word for word as
    # the device builder's, which is then checked and timed once the GPU is up)
    host_bda = None
    if not a.no_cpu:
        from oracle import bda as obda
        host_bda = obda.bda_map_mode1
    P = configs.cfg2(with_vis=False, log=log, bda_builder=host_bda)
    if a.facets:
        P.facets = P.facets[:a.facets]
    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline(P)
            log("cpu baseline: %s" % (cpu,))
        except Exception as e:  # the baseline must not hide the GPU result
            log("cpu baseline failed: %r" % (e,))
    import torch
    from ddfacet_amd.facets import DeviceFacetSet
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    P.vis = configs.cfg2_visibilities(P.uvw, P.freqs, P.extra["sky"], device="cuda")
    bda_dev = device_bda_map(P, a.steps)
    if host_bda is not None:
        bda_dev["equal_to_host_builder"] = bool(np.array_equal(bda_dev.pop("map"), P.bda_grid))
    else:
        bda_dev.pop("map")
    log("device BDA map: %s" % bda_dev)
    fs = DeviceFacetSet(P, device=dev)
    nf = fs.nf
    gen = torch.Generator(device=dev).manual_seed(7)
    fs.models.copy_(torch.randn(fs.models.shape, generator=gen, device=dev, dtype=torch.complex64) * 1e-3)
    vis0 = fs.t["vis"].clone()
    log("setup %.1fs" % (time.time() - t0))

    stream = torch.cuda.current_stream()
    plan = None
    if a.mode == "batched":
        torch.cuda.synchronize()
        tp = time.perf_counter()
        fs.plan_batched(tile=a.tile, group=a.group, blocks_per_item=a.blocks_per_item, stream=stream)
        torch.cuda.synchronize()
        plan_ms = (time.perf_counter() - tp) * 1e3
        fs.plan_timing(True)
        info = fs.plan_info()
        plan = {"ms": plan_ms, "device_bytes": info["device_bytes"],
                "note": "placements, bins and stamp records of both maps; once per chunk, untimed in the cycle"}
        log("plan: %s, built in %.1f ms" % (info, plan_ms))
    ev = {"grid": [], "degrid": []}
    passes = []

    def step(record):
        fs.zero_grids()
        if record:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e2 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
        if a.mode == "plain":
            fs.grid_plain(stream=stream)
        else:
            fs.grid_batched(stream=stream)
        if record:
            e1.record(stream)
        if a.mode == "plain":
            fs.degrid_plain(stream=stream)
        else:
            fs.degrid_batched(stream=stream)
        if record:
            e2.record(stream)
            ev["grid"].append((e0, e1))
            ev["degrid"].append((e1, e2))

    for _ in range(a.warmup):
        step(False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(a.steps):
        step(True)
        if a.mode == "batched":
            passes.append(fs.pass_ms())  # waits for this step's pass events
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    ms = (t2 - t1) * 1e3 / a.steps
    fs.t["vis"].copy_(vis0)
    R, C = P.uvw.shape[0], P.freqs.shape[0]
    nvis = R * C
    value = nvis / (ms * 1e-3)
    tg = float(np.mean([x.elapsed_time(y) for x, y in ev["grid"]]))
    td = float(np.mean([x.elapsed_time(y) for x, y in ev["degrid"]]))
    log("grid %.2f ms, degrid %.2f ms per cycle (%s)" % (tg, td, a.mode))
    if a.mode == "batched":
        ns = len(P.stokes_codes)
        pms = {k: float(np.mean([d[k] for d in passes])) for k in passes[0] if passes[0][k] is not None}
        pb = pass_bytes(P, nf, info, ns)
        sb = survey_pass_bytes(P, nf)
        dom = max(pms, key=pms.get)
        launch_ms, abytes = pms[dom], sb[dom]
        extra = {"passes_ms": pms,
                 "passes_frac": {k: sb[k] / (pms[k] * 1e-3) / HBM_PEAK for k in pms if k in sb},
                 "unique_bytes_per_launch": pb[dom],
                 "unique_frac": pb[dom] / (launch_ms * 1e-3) / HBM_PEAK,
                 "passes_unique_gbps": {k: pb[k] / (pms[k] * 1e-3) / 1e9 for k in pms if k in pb}}
    else:
        dom = "grid_plain" if tg >= td else "degrid_plain"
        launch_ms = (tg if dom == "grid_plain" else td) / nf
        abytes = plain_bytes(P, "grid" if dom == "grid_plain" else "degrid")
        extra = {}
    ach = abytes / (launch_ms * 1e-3)
    traffic = measured_traffic(dom) if (a.mode == "batched" and not a.facets) else None
    cb = cycle_bytes(P, nf)
    graph_ms = None
    if a.graph and a.mode == "batched":
        graph_ms = time_graph(fs, a.steps)
        fs.t["vis"].copy_(vis0)
    err64 = None
    image = None
    if a.mode == "batched" and not a.no_image:
        image = image_side_ms(P, fs, a.steps)
        log("image side: %s" % (image,))
    if a.mode == "batched" and not a.no_fp64_check:
        err64 = error_vs_fp64(fs, vis0, a.tile, a.group, a.blocks_per_item, stream, a.steps)
    cfg = {"workload": "81-facet MeerKAT-like: 9x9 facets (455 -> 735 px at 1.5\"), 64 antennas, "
                       "600 x 8 s, 64 ch 0.9-1.67 GHz, 4 corr, per-visibility DD Jones, default BDA (0.02)",
           "rows": R, "channels": C, "nvis": nvis, "facets": nf,
           "bda_blocks": fs.nblocks_grid, "mode": a.mode, "grid_ms": tg, "degrid_ms": td}
    if a.mode == "batched":
        cfg.update(grid_records=info["grid_records"], degrid_records=info["degrid_records"])
    out = {
        "metric": "vis/sec grid+degrid, 81-facet MeerKAT-like; % HBM roofline; rel err vs FP64",
        "value": value, "unit": "vis/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64 (fp32 grids, fp64 phase base)", "data": "synthetic",
        "config": cfg,
        "roofline": dict({"bound": "hbm", "kernel": dom, "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9,
                          "unit": "GB/s", "frac": ach / HBM_PEAK, "traffic": traffic, "launch_ms": launch_ms,
                          "algorithmic_bytes_per_launch": abytes,
                          "cycle": {"bytes": cb, "ms": ms, "achieved": cb / (ms * 1e-3) / 1e9,
                                    "frac": cb / (ms * 1e-3) / HBM_PEAK}}, **extra),
        "cpu_baseline": cpu,
    }
    if graph_ms is not None:
        out["graph_ms_per_step"] = graph_ms
    if err64 is not None:
        out["fp64_ms_per_step"] = err64.pop("fp64_ms_per_step")
        out["rel_err_vs_fp64"] = err64
    if plan is not None:
        out["plan"] = plan
    if image is not None:
        out["image_side_ms"] = image
    out["bda_map"] = bda_dev
    print(json.dumps(out))


if __name__ == "__main__":
    main()
